// Context lifetime, error reporting and the host check entry points.
#include <cstdarg>

#include "smh_common.h"
#include "smh_special.h"

// Fixed scratch of the entry points that allocate nothing (smh_op_matvec_noalloc):
// one copy per device, alive until the process ends. Host code uses only the
// device address that smh_init resolves into ctx->scratch, never this name.
static __device__ __align__(256) char scratch[kScratchBytes];

int smh_fail(smh_ctx* ctx, int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (ctx) ctx->last_error = buf;
  return code;
}

extern "C" int smh_init(int device, smh_ctx** out) {
  if (!out) return SMH_ERR_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return SMH_ERR_HIP;
  if (device < 0 || device >= count) return SMH_ERR_ARG;
  smh_ctx* ctx = new smh_ctx();
  ctx->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&ctx->ev0);
  if (e == hipSuccess) e = hipEventCreate(&ctx->ev1);
  hipDeviceProp_t prop;
  if (e == hipSuccess) e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) {
    delete ctx;
    return SMH_ERR_HIP;
  }
  // this device's copy of the scratch array (the device is current)
  void* sp = nullptr;
  if (e == hipSuccess) e = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
  if (e != hipSuccess || sp == nullptr) {
    hipEventDestroy(ctx->ev0);
    hipEventDestroy(ctx->ev1);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return SMH_ERR_HIP;
  }
  ctx->scratch = static_cast<char*>(sp);
  ctx->num_cu = prop.multiProcessorCount;
  ctx->lds_per_block = prop.sharedMemPerBlock;
  snprintf(ctx->arch, sizeof(ctx->arch), "%s", prop.gcnArchName);
  if (strncmp(ctx->arch, "gfx950", 6) != 0) {
    fprintf(stderr, "smh: device %d is %s, kernels are built for gfx950\n", device, ctx->arch);
    hipEventDestroy(ctx->ev0);
    hipEventDestroy(ctx->ev1);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return SMH_ERR_UNSUPPORTED;
  }
  *out = ctx;
  return SMH_OK;
}

void smh_ctx_release_if_idle(smh_ctx* ctx) {
  if (!ctx || !ctx->destroy_requested || ctx->live_ops > 0) return;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  hipEventDestroy(ctx->ev0);
  hipEventDestroy(ctx->ev1);
  hipStreamDestroy(ctx->stream);
  delete ctx;
}

extern "C" int smh_destroy(smh_ctx* ctx) {
  if (!ctx) return SMH_OK;
  if (ctx->live_ops > 0) {  // freed when the last operator is destroyed
    ctx->destroy_requested = true;
    return SMH_OK;
  }
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  hipEventDestroy(ctx->ev0);
  hipEventDestroy(ctx->ev1);
  hipStreamDestroy(ctx->stream);
  delete ctx;
  return SMH_OK;
}

extern "C" const char* smh_last_error(const smh_ctx* ctx) { return ctx ? ctx->last_error.c_str() : "no context"; }

extern "C" int smh_device_info(smh_ctx* ctx, char* name, int name_len, int* num_cu) {
  if (!ctx) return SMH_ERR_ARG;
  if (name && name_len > 0) snprintf(name, name_len, "%s", ctx->arch);
  if (num_cu) *num_cu = ctx->num_cu;
  return SMH_OK;
}

extern "C" int smh_device_alloc(smh_ctx* ctx, int64_t bytes, void** out) {
  if (!ctx || !out) return SMH_ERR_ARG;
  *out = nullptr;
  SMH_CHECK(ctx, bytes > 0, SMH_ERR_ARG, "allocation size must be positive");
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  SMH_HIP(ctx, hipMalloc(out, (size_t)bytes));
  return SMH_OK;
}

extern "C" int smh_device_free(smh_ctx* ctx, void* p) {
  if (!ctx) return SMH_ERR_ARG;
  if (!p) return SMH_OK;
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  SMH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  SMH_HIP(ctx, hipFree(p));
  return SMH_OK;
}

extern "C" int smh_device_copy(smh_ctx* ctx, void* dst, const void* src, int64_t bytes) {
  if (!ctx) return SMH_ERR_ARG;
  SMH_CHECK(ctx, bytes >= 0 && (bytes == 0 || (dst && src)), SMH_ERR_ARG, "invalid copy arguments");
  if (bytes == 0) return SMH_OK;
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  SMH_HIP(ctx, hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDefault, ctx->stream));
  SMH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SMH_OK;
}

extern "C" int smh_host_bessel_j(int nmax, int64_t count, const double* z, double* out) {
  if (nmax < 0 || nmax > smh::kMaxBesselOrder || count < 0 || (!z && count) || (!out && count)) return SMH_ERR_ARG;
  smh::zc buf[smh::kMaxBesselOrder + 1];
  for (int64_t j = 0; j < count; ++j) {
    smh::bessel_j_all(nmax, smh::zc{z[2 * j], z[2 * j + 1]}, buf);
    for (int n = 0; n <= nmax; ++n) {
      out[2 * (j * (nmax + 1) + n)] = buf[n].re;
      out[2 * (j * (nmax + 1) + n) + 1] = buf[n].im;
    }
  }
  return SMH_OK;
}

extern "C" int smh_host_spherical_hankel(int nmax, int64_t count, const double* z, double* out) {
  if (nmax < 0 || nmax > smh::kMaxBesselOrder || count < 0 || (!z && count) || (!out && count)) return SMH_ERR_ARG;
  smh::zc buf[smh::kMaxBesselOrder + 1];
  for (int64_t j = 0; j < count; ++j) {
    smh::zc zz{z[2 * j], z[2 * j + 1]};
    if (zz.re == 0.0 && zz.im == 0.0) return SMH_ERR_ARG;
    smh::spherical_hankel_all(nmax, zz, buf);
    for (int n = 0; n <= nmax; ++n) {
      out[2 * (j * (nmax + 1) + n)] = buf[n].re;
      out[2 * (j * (nmax + 1) + n) + 1] = buf[n].im;
    }
  }
  return SMH_OK;
}

namespace {
__global__ void copy16_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x)
    dst[j] = src[j];
}
}  // namespace

extern "C" int smh_bench_copy(smh_ctx* ctx, int64_t bytes, double* ms) {
  if (!ctx) return SMH_ERR_ARG;
  SMH_CHECK(ctx, bytes >= 16 && bytes % 16 == 0, SMH_ERR_ARG, "copy size must be a positive multiple of 16 bytes");
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  void *a = nullptr, *b = nullptr;
  SMH_HIP(ctx, hipMalloc(&a, bytes));
  if (hipMalloc(&b, bytes) != hipSuccess) {
    hipFree(a);
    return smh_fail(ctx, SMH_ERR_HIP, "copy buffer allocation failed");
  }
  hipStream_t s = ctx->stream;
  const int64_t n = bytes / 16;
  hipMemsetAsync(a, 1, bytes, s);
  hipEventRecord(ctx->ev0, s);
  hipLaunchKernelGGL(copy16_kernel, dim3(4096), dim3(256), 0, s, (const uint4*)a, (uint4*)b, n);
  hipEventRecord(ctx->ev1, s);
  const hipError_t e = hipEventSynchronize(ctx->ev1);
  float t = 0;
  hipEventElapsedTime(&t, ctx->ev0, ctx->ev1);
  hipFree(a);
  hipFree(b);
  SMH_CHECK(ctx, e == hipSuccess, SMH_ERR_HIP, "copy failed: %s", hipGetErrorString(e));
  if (ms) *ms = t;
  return SMH_OK;
}
