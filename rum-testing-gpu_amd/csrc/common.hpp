// Shared host/device helpers of librumgpu: status codes, the context (one device, one stream),
// the last-error buffer, and the subset-lattice bit helpers used by every kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

typedef int64_t i64;
typedef uint64_t u64;
typedef uint32_t u32;

enum {
  RG_OK = 0,
  RG_EINVAL = 1,     // invalid argument
  RG_ENOMEM = 2,     // out of device memory
  RG_EHIP = 3,       // any other HIP runtime error
  RG_EUNSUP = 4,     // valid but unsupported here (size limits)
};

void rg_set_error(const char* fmt, ...);

// Return-on-error helpers for host code; every C-ABI entry returns an int status.
#define RG_HIP(call)                                                                       \
  do {                                                                                     \
    hipError_t _e = (call);                                                                \
    if (_e != hipSuccess) {                                                                \
      rg_set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), __FILE__, __LINE__); \
      return (_e == hipErrorOutOfMemory || _e == hipErrorMemoryAllocation) ? RG_ENOMEM : RG_EHIP; \
    }                                                                                      \
  } while (0)

#define RG_TRY(expr)              \
  do {                            \
    int _s = (expr);              \
    if (_s != RG_OK) return _s;   \
  } while (0)

#define RG_CHECK(cond, code, ...) \
  do {                            \
    if (!(cond)) {                \
      rg_set_error(__VA_ARGS__);  \
      return (code);              \
    }                             \
  } while (0)

struct rumgpu_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  size_t lds_per_block = 65536;   // the device's LDS per workgroup (160 KiB on gfx950)
};

// Device allocation that records a readable error (with the size) on failure.
int rg_malloc(void** p, size_t bytes, const char* what);
void rg_free(void* p);

// Launch-error check after a kernel launch.
#define RG_LAUNCH_CHECK()                                                                  \
  do {                                                                                     \
    hipError_t _e = hipGetLastError();                                                     \
    if (_e != hipSuccess) {                                                                \
      rg_set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(_e), __FILE__, __LINE__); \
      return RG_EHIP;                                                                      \
    }                                                                                      \
  } while (0)

// ------------------------------------------------------------------ lattice bit helpers
// Slab position of (item i, set S containing i): i * 2^(n-1) + compress(S minus i, i).
__host__ __device__ inline u64 compress_bit(u64 m, int i) {
  return (m & ((1ull << i) - 1ull)) | ((m >> (i + 1)) << i);
}
// Inverse: the set (without i) whose compressed index is rest.
__host__ __device__ inline u64 expand_bit(u64 rest, int i) {
  return (rest & ((1ull << i) - 1ull)) | ((rest >> i) << (i + 1));
}

// Stream-ordered copy that returns when the copy is done: the engines' streams are non-blocking,
// so a plain hipMemcpy (null stream) would not wait for their kernels.
inline hipError_t rg_copy(hipStream_t st, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, st);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(st);
}

inline unsigned grid_for(i64 nthreads, int block) { return (unsigned)((nthreads + block - 1) / block); }
