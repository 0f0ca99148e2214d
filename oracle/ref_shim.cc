// TEST INFRASTRUCTURE ONLY.  A C-ABI shim over the reference's own CPU
// rasterizer (diffren::RasterizeTrianglesImpl, rasterize_triangles_impl.h:53-58),
// compiled together with the unmodified reference source by oracle/Makefile into
// oracle/_ref/.  It performs the output initialisation of the reference's CPU
// FFI handler (rasterize_triangles_cpu.cc:55-60: ids -1, z 1, barycentrics 0)
// and then calls the reference function.  Used to pin the oracle restatement
// and as the reference CPU baseline.
#include <cstdint>
#include <cstring>

#include "diffren/common/kernels/rasterize_triangles_impl.h"

extern "C" void ref_rasterize(const float* vertices, const int32_t* triangles,
                              int32_t triangle_count, int32_t width,
                              int32_t height, int32_t layers,
                              int32_t culling_mode, int32_t* ids, float* z,
                              float* barys) {
  const int64_t n = int64_t(width) * height * layers;
  for (int64_t i = 0; i < n; ++i) {
    ids[i] = -1;
    z[i] = 1.0f;
  }
  if (barys) std::memset(barys, 0, sizeof(float) * 3 * n);
  diffren::RasterizeTrianglesImpl(
      vertices, triangles, triangle_count, width, height, layers,
      static_cast<diffren::FaceCullingMode>(culling_mode), ids, z, barys);
}
