// In-place 3D complex128 FFT for gfx950: one kernel launch per axis, each
// workgroup transforming a tile of C lines of length n held in LDS with the
// Stockham autosort formulation (radix-4 passes, one radix-2 pass when
// log2 n is odd).  Unnormalised, exponent sign * 2 pi i jk / n: sign = +1 is
// the reference's FFTW_BACKWARD plan, -1 its FFTW_FORWARD plan
// (grid_brnd.cc:42-54, grid_ternd.cc:35-37).
//
// Global access: along the contiguous z axis each line is a contiguous run;
// along y and x the tile's C lines are C adjacent columns, so every tile row
// is C*16 contiguous bytes (C = 16 -> 256 B rows).  Twiddles come from a
// per-length table exp(2 pi i m / n) computed on the host in long double.
#include <hip/hip_runtime.h>

#include <cmath>
#include <map>
#include <vector>

#include "hamx_internal.h"

namespace hamx {

struct FftAxis {
  uint64_t s;      // element stride of the transformed axis
  uint64_t inner;  // lines with consecutive base addresses
  uint64_t outer;  // base-address step between groups of `inner` lines
  uint32_t n;      // line length (power of 2)
  uint32_t C;      // lines per workgroup
};

__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

template <int SIGN>
__device__ inline double2 twiddle(const double2 *tw, uint32_t j) {
  double2 w = tw[j];
  if (SIGN < 0)
    w.y = -w.y;
  return w;
}

template <int SIGN>
__global__ __launch_bounds__(256) void k_fft_axis(double2 *__restrict__ data,
                                                  FftAxis ax,
                                                  const double2 *__restrict__ tw) {
  extern __shared__ double2 lds[];
  const uint32_t n = ax.n, C = ax.C, total = C * n;
  const uint64_t L0 = static_cast<uint64_t>(blockIdx.x) * C;
  for (uint32_t e = threadIdx.x; e < total; e += blockDim.x) {
    uint32_t c, t;
    if (ax.s == 1) {
      c = e / n;
      t = e % n;
    } else {
      c = e % C;
      t = e / C;
    }
    const uint64_t L = L0 + c;
    const uint64_t base = (L / ax.inner) * ax.outer + (L % ax.inner);
    lds[c * n + t] = data[base + static_cast<uint64_t>(t) * ax.s];
  }
  __syncthreads();
  // Every pass stages its inputs in registers (at most 16 complex values
  // per thread: C*n <= 4096 and 256 threads), then writes the outputs.
  uint32_t p = 1;
  // radix-4 Stockham passes
  while (p * 4 <= n) {
    const uint32_t q4 = n >> 2;
    const uint32_t nbf = C * q4;
    double2 x[4][4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t b = threadIdx.x + q * 256;
      if (b < nbf) {
        const uint32_t c = b / q4, i = b % q4;
        const double2 *line = lds + c * n;
#pragma unroll
        for (int m = 0; m < 4; ++m)
          x[q][m] = line[i + m * q4];
      }
    }
    __syncthreads();
    const uint32_t tstep = n / (4 * p);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t b = threadIdx.x + q * 256;
      if (b < nbf) {
        const uint32_t c = b / q4, i = b % q4;
        const uint32_t k = i & (p - 1);
        const double2 x0 = x[q][0];
        const double2 x1 = cmul(x[q][1], twiddle<SIGN>(tw, k * tstep));
        const double2 x2 = cmul(x[q][2], twiddle<SIGN>(tw, 2 * k * tstep));
        const double2 x3 = cmul(x[q][3], twiddle<SIGN>(tw, 3 * k * tstep));
        const double2 a0 = cadd(x0, x2), a1 = csub(x0, x2), a2 = cadd(x1, x3);
        const double2 d13 = csub(x1, x3);
        // (x1 - x3) * SIGN * i
        const double2 a3 =
            SIGN > 0 ? make_double2(-d13.y, d13.x) : make_double2(d13.y, -d13.x);
        double2 *line = lds + c * n;
        const uint32_t j = ((i - k) << 2) + k;
        line[j] = cadd(a0, a2);
        line[j + p] = cadd(a1, a3);
        line[j + 2 * p] = csub(a0, a2);
        line[j + 3 * p] = csub(a1, a3);
      }
    }
    __syncthreads();
    p <<= 2;
  }
  // final radix-2 pass (p = n/2)
  if (p * 2 == n) {
    const uint32_t q2 = n >> 1;
    const uint32_t nbf = C * q2;
    double2 x[8][2];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const uint32_t b = threadIdx.x + q * 256;
      if (b < nbf) {
        const uint32_t c = b / q2, i = b % q2;
        x[q][0] = lds[c * n + i];
        x[q][1] = lds[c * n + i + q2];
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const uint32_t b = threadIdx.x + q * 256;
      if (b < nbf) {
        const uint32_t c = b / q2, i = b % q2;
        const double2 x1 = cmul(x[q][1], twiddle<SIGN>(tw, i));
        lds[c * n + i] = cadd(x[q][0], x1);
        lds[c * n + i + p] = csub(x[q][0], x1);
      }
    }
    __syncthreads();
  }
  for (uint32_t e = threadIdx.x; e < total; e += blockDim.x) {
    uint32_t c, t;
    if (ax.s == 1) {
      c = e / n;
      t = e % n;
    } else {
      c = e % C;
      t = e / C;
    }
    const uint64_t L = L0 + c;
    const uint64_t base = (L / ax.inner) * ax.outer + (L % ax.inner);
    data[base + static_cast<uint64_t>(t) * ax.s] = lds[c * n + t];
  }
}

// exp(+2 pi i m / n), m in [0, n), cached per context and length.
static std::map<std::pair<hamx_ctx *, uint64_t>, double2 *> g_twiddles;

static int twiddle_table(hamx_ctx *ctx, uint64_t n, const double2 **out) {
  auto key = std::make_pair(ctx, n);
  auto it = g_twiddles.find(key);
  if (it != g_twiddles.end()) {
    *out = it->second;
    return HAMX_OK;
  }
  std::vector<double2> h(n);
  const long double twopi = 6.283185307179586476925286766559005768394L;
  for (uint64_t m = 0; m < n; ++m) {
    const long double a = twopi * static_cast<long double>(m) / static_cast<long double>(n);
    h[m] = make_double2(static_cast<double>(cosl(a)), static_cast<double>(sinl(a)));
  }
  double2 *d = nullptr;
  HAMX_HIP(ctx, hipMalloc(&d, n * sizeof(double2)));
  HAMX_HIP(ctx, hipMemcpy(d, h.data(), n * sizeof(double2), hipMemcpyHostToDevice));
  g_twiddles[key] = d;
  *out = d;
  return HAMX_OK;
}

void release_twiddles(hamx_ctx *ctx) {
  for (auto it = g_twiddles.begin(); it != g_twiddles.end();) {
    if (it->first.first == ctx) {
      (void)hipFree(it->second);
      it = g_twiddles.erase(it);
    } else {
      ++it;
    }
  }
}

int fft3d_dev(hamx_ctx *ctx, double2 *data, uint64_t nx, uint64_t ny,
              uint64_t nz, int sign) {
  const uint64_t dims[3] = {nx, ny, nz};
  for (int a = 0; a < 3; ++a)
    if (!is_pow2(dims[a]) || dims[a] > 4096)
      return set_err(ctx, HAMX_ERR_UNSUPPORTED,
                     "FFT axis length %llu is not a power of two in [1, 4096]",
                     static_cast<unsigned long long>(dims[a]));
  const uint64_t total = nx * ny * nz;
  // axis order z, y, x (any order gives the same transform)
  for (int a = 2; a >= 0; --a) {
    const uint64_t n = dims[a];
    if (n < 2)
      continue;
    FftAxis ax;
    ax.n = static_cast<uint32_t>(n);
    if (a == 2) {
      ax.s = 1;
      ax.inner = 1;
      ax.outer = nz;
    } else if (a == 1) {
      ax.s = nz;
      ax.inner = nz;
      ax.outer = ny * nz;
    } else {
      ax.s = ny * nz;
      ax.inner = ny * nz;
      ax.outer = total;
    }
    // lines per workgroup: fill 4096 complex of LDS (64 KB); for strided
    // axes the lines must be adjacent columns, so C divides `inner`.
    uint64_t C = 4096 / n;
    if (C > 16 && ax.s != 1)
      C = 16;
    if (ax.s != 1 && C > ax.inner)
      C = ax.inner;
    const uint64_t lines = total / n;
    if (C > lines)
      C = lines;
    ax.C = static_cast<uint32_t>(C);
    const double2 *tw;
    HAMX_TRY(twiddle_table(ctx, n, &tw));
    const size_t lds = static_cast<size_t>(C * n) * sizeof(double2);
    const unsigned blocks = static_cast<unsigned>(lines / C);
    if (sign > 0)
      hipLaunchKernelGGL(k_fft_axis<1>, dim3(blocks), dim3(256), lds, ctx->stream,
                         data, ax, tw);
    else
      hipLaunchKernelGGL(k_fft_axis<-1>, dim3(blocks), dim3(256), lds, ctx->stream,
                         data, ax, tw);
    HAMX_HIP(ctx, hipGetLastError());
  }
  return HAMX_OK;
}

} // namespace hamx

using namespace hamx;

extern "C" int hamx_fft3d(hamx_ctx *ctx, double *data, uint64_t nx, uint64_t ny,
                          uint64_t nz, int sign, uint32_t flags) {
  if (!ctx)
    return HAMX_ERR_ARG;
  if (!data || (sign != 1 && sign != -1))
    return set_err(ctx, HAMX_ERR_ARG, "hamx_fft3d: null data or sign not +-1");
  HAMX_HIP(ctx, hipSetDevice(ctx->device));
  const uint64_t total = nx * ny * nz;
  const size_t bytes = total * sizeof(double2);
  double2 *d = reinterpret_cast<double2 *>(data);
  const bool in_dev = (flags & HAMX_IN_DEVICE) != 0;
  if (!in_dev) {
    void *w;
    HAMX_TRY(ws_get(ctx, kSlotFft, bytes, &w));
    HAMX_HIP(ctx, hipMemcpyAsync(w, data, bytes, hipMemcpyHostToDevice, ctx->stream));
    d = static_cast<double2 *>(w);
  }
  HAMX_TRY(fft3d_dev(ctx, d, nx, ny, nz, sign));
  if (!in_dev)
    HAMX_HIP(ctx, hipMemcpyAsync(data, d, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HAMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HAMX_OK;
}
