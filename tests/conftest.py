import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))


def pytest_configure(config):
  config.addinivalue_line(
      "markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def device():
  import torch
  if not torch.cuda.is_available():
    pytest.fail("gpu test selected but no GPU is visible")
  return torch.device("cuda:0")
