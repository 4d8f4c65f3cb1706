// Context, device queries, allocation and the last-error buffer of librumgpu.
#include "rumgpu_internal.hpp"

#include <cstdarg>
#include <mutex>

static std::mutex g_err_mu;
static char g_err[1024] = "";

void rg_set_error(const char* fmt, ...) {
  std::lock_guard<std::mutex> lk(g_err_mu);
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int rg_malloc(void** p, size_t bytes, const char* what) {
  *p = nullptr;
  if (bytes == 0) bytes = 1;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    size_t fr = 0, tot = 0;
    (void)hipMemGetInfo(&fr, &tot);
    rg_set_error("out of memory: %s needs %.3f GB, %.3f GB free of %.3f GB (%s)", what, bytes / 1073741824.0,
                 fr / 1073741824.0, tot / 1073741824.0, hipGetErrorString(e));
    *p = nullptr;
    return (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? RG_ENOMEM : RG_EHIP;
  }
  return RG_OK;
}

void rg_free(void* p) {
  if (p) (void)hipFree(p);
}

extern "C" int rumgpu_version(void) { return 1; }

extern "C" int rumgpu_last_error(char* buf, int len) {
  std::lock_guard<std::mutex> lk(g_err_mu);
  if (buf && len > 0) {
    strncpy(buf, g_err, (size_t)len - 1);
    buf[len - 1] = 0;
  }
  return RG_OK;
}

extern "C" int rumgpu_device_count(int* count) {
  RG_CHECK(count, RG_EINVAL, "device_count: null argument");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) {
    *count = 0;
    rg_set_error("hipGetDeviceCount: %s", hipGetErrorString(e));
    return RG_EHIP;
  }
  *count = c;
  return RG_OK;
}

extern "C" int rumgpu_ctx_create(int device, rumgpu_ctx** out) {
  RG_CHECK(out, RG_EINVAL, "ctx_create: null argument");
  int c = 0;
  RG_HIP(hipGetDeviceCount(&c));
  RG_CHECK(device >= 0 && device < c, RG_EINVAL, "ctx_create: device %d of %d", device, c);
  RG_HIP(hipSetDevice(device));
  rumgpu_ctx* ctx = new rumgpu_ctx();
  ctx->device = device;
  hipDeviceProp_t pr;
  if (hipGetDeviceProperties(&pr, device) == hipSuccess && pr.sharedMemPerBlock > 0)
    ctx->lds_per_block = pr.sharedMemPerBlock;
  // the no-allocation entry point's scratch address on this device (a failure is reported by
  // that entry point; every other entry point works without it)
  (void)rg_noalloc_setup(device);
  hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete ctx;
    rg_set_error("hipStreamCreateWithFlags: %s", hipGetErrorString(e));
    return RG_EHIP;
  }
  *out = ctx;
  return RG_OK;
}

extern "C" int rumgpu_ctx_destroy(rumgpu_ctx* ctx) {
  if (!ctx) return RG_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return RG_OK;
}

extern "C" int rumgpu_mem_info(int device, size_t* free_bytes, size_t* total_bytes) {
  RG_CHECK(free_bytes && total_bytes, RG_EINVAL, "mem_info: null argument");
  RG_HIP(hipSetDevice(device));
  RG_HIP(hipMemGetInfo(free_bytes, total_bytes));
  return RG_OK;
}

extern "C" int rumgpu_device_props(int device, char* name, int name_len, char* arch, int arch_len, size_t* total_mem,
                                   int* cu_count) {
  hipDeviceProp_t pr;
  RG_HIP(hipGetDeviceProperties(&pr, device));
  if (name && name_len > 0) {
    // some boxes report an empty marketing name: fall back to the architecture and CU count
    char fallback[96];
    const char* nm = pr.name;
    if (!nm[0]) {
      snprintf(fallback, sizeof(fallback), "AMD %s (%d CUs)", pr.gcnArchName, pr.multiProcessorCount);
      nm = fallback;
    }
    strncpy(name, nm, (size_t)name_len - 1);
    name[name_len - 1] = 0;
  }
  if (arch && arch_len > 0) {
    strncpy(arch, pr.gcnArchName, (size_t)arch_len - 1);
    arch[arch_len - 1] = 0;
  }
  if (total_mem) *total_mem = pr.totalGlobalMem;
  if (cu_count) *cu_count = pr.multiProcessorCount;
  return RG_OK;
}
