// Shared host/device helpers for the diffren_amd HIP library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

namespace dr {

// Per-thread last error message returned by dr_last_error().
void set_error(const std::string& msg);
const char* get_error();

// Status codes returned by every C-ABI entry point.
enum Status : int {
  kOk = 0,
  kInvalidArgument = 1,
  kWorkspaceTooSmall = 2,
  kLaunchFailed = 3,
  kUnsupported = 4,
};

inline int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

}  // namespace dr

#define DR_CHECK_ARG(cond, msg)                                   \
  do {                                                            \
    if (!(cond)) return dr::fail(dr::kInvalidArgument, (msg));    \
  } while (0)

#define DR_CHECK_LAUNCH(what)                                                 \
  do {                                                                        \
    hipError_t e_ = hipGetLastError();                                        \
    if (e_ != hipSuccess)                                                     \
      return dr::fail(dr::kLaunchFailed,                                      \
                      std::string(what) + ": " + hipGetErrorString(e_));      \
  } while (0)

#define DR_CHECK_HIP(call)                                                    \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess)                                                     \
      return dr::fail(dr::kLaunchFailed,                                      \
                      std::string(#call) + ": " + hipGetErrorString(e_));     \
  } while (0)

// Wave (64 lanes on CDNA) helpers.
namespace dr {

// Load through a global-address-space pointer: pointers read out of job
// tables are generic, and generic (flat) loads also count against the LDS
// wait counter, so a wait for LDS results would wait for them too.
template <typename T>
__device__ __forceinline__ T ldg(const T* p) {
  return *(const __attribute__((address_space(1))) T*)(p);
}

// 16-byte global loads at p (16-byte aligned).
__device__ __forceinline__ float4 ldg4(const float* p) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  const f4 q = *(const __attribute__((address_space(1))) f4*)(p);
  return make_float4(q.x, q.y, q.z, q.w);
}
__device__ __forceinline__ uint4 ldgu4(const uint32_t* p) {
  typedef uint32_t u4 __attribute__((ext_vector_type(4)));
  const u4 q = *(const __attribute__((address_space(1))) u4*)(p);
  return make_uint4(q.x, q.y, q.z, q.w);
}

// *p for a wave-uniform p, word by word through readfirstlane, so the value
// lives in scalar registers.
template <typename T>
__device__ __forceinline__ T load_uniform(const T* p) {
  static_assert(sizeof(T) % 4 == 0, "whole 32-bit words");
  T out;
  const uint32_t* src = reinterpret_cast<const uint32_t*>(p);
  uint32_t* dst = reinterpret_cast<uint32_t*>(&out);
#pragma unroll
  for (int i = 0; i < int(sizeof(T) / 4); ++i)
    dst[i] = __builtin_amdgcn_readfirstlane(ldg(src + i));
  return out;
}

// Stores through global-address-space pointers: a generic pointer (read out
// of a job table) would make them flat stores, which also count against the
// LDS wait counter, so every later LDS wait or barrier would wait for them.
template <typename T>
__device__ __forceinline__ void gst(T* p, T v) {
  *(__attribute__((address_space(1))) T*)(p) = v;
}
// 16-byte stores at p (16-byte aligned).
__device__ __forceinline__ void gst4(float* p, float4 v) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  f4 q;
  q.x = v.x; q.y = v.y; q.z = v.z; q.w = v.w;
  *(__attribute__((address_space(1))) f4*)(p) = q;
}
__device__ __forceinline__ void gstu4(uint32_t* p, uint4 v) {
  typedef uint32_t u4 __attribute__((ext_vector_type(4)));
  u4 q;
  q.x = v.x; q.y = v.y; q.z = v.z; q.w = v.w;
  *(__attribute__((address_space(1))) u4*)(p) = q;
}

// Float atomic add through a global-address-space pointer (hardware
// global_atomic_add_f32, relaxed, device scope; a generic pointer would make
// it a flat atomic, which also counts against the LDS wait counter).
__device__ __forceinline__ void gatomic_add(float* p, float v) {
  __hip_atomic_fetch_add((__attribute__((address_space(1))) float*)(p), v,
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// 64-bit integer atomic add through a global-address-space pointer.
__device__ __forceinline__ void gatomic_add_u64(unsigned long long* p,
                                                unsigned long long v) {
  __hip_atomic_fetch_add(
      (__attribute__((address_space(1))) unsigned long long*)(p), v,
      __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline int lane_id() { return __lane_id(); }

// Work item (screen tile, or chunk of triangles) of this block, XCD-aware.
// The dispatcher deals workgroups to the chip's 8 XCDs round-robin (block b
// to XCD b % 8), and each XCD has its own L2, so consecutive items -- which
// share vertices, triangles and halo pixels -- would land in 8 different
// L2s.  This maps each XCD's blocks onto runs of kRun consecutive items
// (fine-grained, so the XCDs' loads stay balanced): block 8k + x takes item
// ((k / kRun) * 8 + x) * kRun + k % kRun.  A bijection on the largest
// multiple of 8 kRun blocks; the remaining tail keeps its own index.
template <int kRun>
__device__ inline int64_t xcd_order(int64_t b, int64_t n) {
  const int64_t full = n / (8 * kRun) * (8 * kRun);
  if (b >= full) return b;
  const int64_t x = b % 8, k = b / 8;
  return ((k / kRun) * 8 + x) * kRun + k % kRun;
}
template <int kRun>
__device__ inline int64_t xcd_block() {
  return xcd_order<kRun>(blockIdx.x, gridDim.x);
}
// The same for a 2-D grid of tiles, dispatched in row-major block order:
// the (x, y) block this workgroup takes.
template <int kRun>
__device__ inline int2 xcd_block_2d() {
  const int64_t w = gridDim.x;
  const int64_t l = xcd_order<kRun>(int64_t(blockIdx.y) * w + blockIdx.x,
                                    w * gridDim.y);
  return make_int2(int(l % w), int(l / w));
}
__device__ inline uint64_t lanemask_lt() {
  return (uint64_t(1) << lane_id()) - 1;
}
}  // namespace dr
