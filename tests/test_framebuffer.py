"""Framebuffer validation (framebuffer_test.py): host logic, no GPU."""
import pytest
import torch

from diffren_amd import framebuffer as fb


def test_initialize_framebuffer_with_wrong_rank():
  with pytest.raises(ValueError, match="Expected all input shapes to match"):
    fb.Framebuffer(barycentrics=torch.ones([1, 1, 4, 4, 1]),
                   triangle_id=torch.ones([4, 3]))


def test_initialize_framebuffer_with_wrong_shapes():
  with pytest.raises(ValueError, match="Expected all input shapes to match"):
    fb.Framebuffer(barycentrics=torch.ones([1, 1, 4, 4, 3]),
                   triangle_id=torch.ones([1, 1, 4, 4, 1]),
                   attributes={"an_attr": torch.ones([1, 1, 4, 3, 4])})


def test_derived_buffers():
  ids = torch.tensor([[[[-1], [1]], [[0], [-1]]]], dtype=torch.int32)
  tris = torch.tensor([[0, 1, 2], [3, 4, 5]], dtype=torch.int32)
  f = fb.Framebuffer(barycentrics=torch.zeros([1, 2, 2, 3]), triangle_id=ids,
                     triangles=tris)
  assert f.foreground_mask[..., 0].tolist() == [[[0.0, 1.0], [1.0, 0.0]]]
  # Background wraps to the last triangle, as the reference's gather does.
  assert f.vertex_ids[0, 0, 0].tolist() == [3, 4, 5]
  assert f.vertex_ids[0, 1, 0].tolist() == [0, 1, 2]
  assert not f.is_batched
