// Context lifetime, workspace and memory helpers of the C-ABI.
#include "hamx_internal.h"

#include <cstdio>
#include <cstring>
#include <new>

namespace hamx {

int ws_get(hamx_ctx *ctx, int slot, size_t bytes, void **out) {
  if (slot < 0 || slot >= hamx_ctx::kSlots)
    return set_err(ctx, HAMX_ERR_ARG, "workspace slot %d out of range", slot);
  if (bytes == 0)
    bytes = 256;
  if (ctx->ws_bytes[slot] < bytes) {
    if (ctx->ws[slot]) {
      HAMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
      HAMX_HIP(ctx, hipFree(ctx->ws[slot]));
      ctx->ws[slot] = nullptr;
      ctx->ws_bytes[slot] = 0;
    }
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return set_err(ctx, HAMX_ERR_NOMEM,
                     "device allocation of %zu bytes failed", bytes);
    }
    ctx->ws[slot] = p;
    ctx->ws_bytes[slot] = bytes;
  }
  *out = ctx->ws[slot];
  return HAMX_OK;
}

int stage_in(hamx_ctx *ctx, int slot, const void *src, size_t bytes,
             bool in_device, const void **dev) {
  if (!src) {
    *dev = nullptr;
    return HAMX_OK;
  }
  if (in_device) {
    *dev = src;
    return HAMX_OK;
  }
  void *p = nullptr;
  HAMX_TRY(ws_get(ctx, slot, bytes, &p));
  HAMX_HIP(ctx, hipMemcpyAsync(p, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  *dev = p;
  return HAMX_OK;
}

__global__ void k_fill_f64(double *p, size_t n, double v) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    p[i] = v;
}

int fill_f64(hamx_ctx *ctx, double *p, size_t n, double v) {
  if (n == 0)
    return HAMX_OK;
  size_t blocks = (n + 255) / 256;
  if (blocks > 8192)
    blocks = 8192;
  hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)blocks), dim3(256), 0,
                     ctx->stream, p, n, v);
  HAMX_HIP(ctx, hipGetLastError());
  return HAMX_OK;
}

static hipEvent_t pool_event(hamx_ctx *ctx) {
  if (!ctx->event_pool.empty()) {
    hipEvent_t e = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return e;
}

ProfMark prof_begin(hamx_ctx *ctx) {
  ProfMark m;
  if (!ctx->profiling)
    return m;
  m.start = pool_event(ctx);
  if (m.start && hipEventRecord(m.start, ctx->stream) == hipSuccess)
    m.active = 1;
  return m;
}

void prof_end(hamx_ctx *ctx, ProfMark &m, const char *kernel, uint32_t shell,
              uint64_t work) {
  if (!m.active)
    return;
  hamx_ctx::Pending p;
  p.start = m.start;
  p.stop = pool_event(ctx);
  if (!p.stop || hipEventRecord(p.stop, ctx->stream) != hipSuccess) {
    ctx->event_pool.push_back(p.start);
    m.active = 0;
    return;
  }
  std::memset(&p.rec, 0, sizeof(p.rec));
  std::snprintf(p.rec.kernel, sizeof(p.rec.kernel), "%s", kernel);
  p.rec.shell = shell;
  p.rec.work = work;
  ctx->pending.push_back(p);
  m.active = 0;
}

void prof_resolve(hamx_ctx *ctx) {
  for (auto &p : ctx->pending) {
    float ms = 0.f;
    if (hipEventSynchronize(p.stop) == hipSuccess &&
        hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
      p.rec.ms = ms;
      ctx->times.push_back(p.rec);
    } else {
      (void)hipGetLastError();
    }
    ctx->event_pool.push_back(p.start);
    ctx->event_pool.push_back(p.stop);
  }
  ctx->pending.clear();
}

} // namespace hamx

using namespace hamx;

extern "C" {

int hamx_abi_version(void) { return HAMX_ABI_VERSION; }

int hamx_profile(hamx_ctx *ctx, int enable) {
  if (!ctx)
    return HAMX_ERR_ARG;
  prof_resolve(ctx);
  ctx->times.clear();
  ctx->profiling = enable != 0;
  return HAMX_OK;
}

int hamx_profile_read(hamx_ctx *ctx, hamx_kernel_time *out, int max, int *count) {
  if (!ctx || !count || (max > 0 && !out))
    return HAMX_ERR_ARG;
  prof_resolve(ctx);
  const int n = static_cast<int>(ctx->times.size());
  const int c = n < max ? n : max;
  for (int i = 0; i < c; ++i)
    out[i] = ctx->times[i];
  *count = n;
  return HAMX_OK;
}

int hamx_open(int device, hamx_ctx **out) {
  if (!out)
    return HAMX_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return HAMX_ERR_DEVICE;
  }
  if (device < 0 || device >= n)
    return HAMX_ERR_ARG;
  hamx_ctx *ctx = new (std::nothrow) hamx_ctx;
  if (!ctx)
    return HAMX_ERR_NOMEM;
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    delete ctx;
    return HAMX_ERR_DEVICE;
  }
  static_scratch_init(ctx);
  *out = ctx;
  return HAMX_OK;
}

void hamx_close(hamx_ctx *ctx) {
  if (!ctx)
    return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream)
    (void)hipStreamSynchronize(ctx->stream);
  release_twiddles(ctx);
  prof_resolve(ctx);
  for (hipEvent_t e : ctx->event_pool)
    (void)hipEventDestroy(e);
  for (int i = 0; i < hamx_ctx::kSlots; ++i)
    if (ctx->ws[i])
      (void)hipFree(ctx->ws[i]);
  if (ctx->pinned)
    (void)hipHostFree(ctx->pinned);
  if (ctx->stream)
    (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char *hamx_last_error(const hamx_ctx *ctx) {
  if (!ctx)
    return "null hamx context";
  return ctx->err.c_str();
}

int hamx_sync(hamx_ctx *ctx) {
  if (!ctx)
    return HAMX_ERR_ARG;
  HAMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HAMX_OK;
}

int hamx_dev_alloc(hamx_ctx *ctx, size_t bytes, void **dptr) {
  if (!ctx || !dptr)
    return HAMX_ERR_ARG;
  HAMX_HIP(ctx, hipSetDevice(ctx->device));
  if (hipMalloc(dptr, bytes ? bytes : 1) != hipSuccess) {
    (void)hipGetLastError();
    *dptr = nullptr;
    return set_err(ctx, HAMX_ERR_NOMEM, "device allocation of %zu bytes failed",
                   bytes);
  }
  return HAMX_OK;
}

int hamx_dev_free(hamx_ctx *ctx, void *dptr) {
  if (!ctx)
    return HAMX_ERR_ARG;
  if (!dptr)
    return HAMX_OK;
  HAMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  HAMX_HIP(ctx, hipFree(dptr));
  return HAMX_OK;
}

int hamx_memcpy_h2d(hamx_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx || (!dst && bytes) || (!src && bytes))
    return HAMX_ERR_ARG;
  HAMX_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  HAMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HAMX_OK;
}

int hamx_memcpy_d2h(hamx_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx || (!dst && bytes) || (!src && bytes))
    return HAMX_ERR_ARG;
  HAMX_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HAMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HAMX_OK;
}

} // extern "C"
