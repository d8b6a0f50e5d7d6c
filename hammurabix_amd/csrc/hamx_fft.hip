// In-place 3D complex128 FFT for gfx950: one kernel launch per axis, each
// workgroup transforming a tile of C lines held in LDS.  Unnormalised,
// exponent sign * 2 pi i jk / n: sign = +1 is the reference's FFTW_BACKWARD
// plan, -1 its FFTW_FORWARD plan (grid_brnd.cc:42-54, grid_ternd.cc:35-37).
// FFTW plans any length; so does this file:
//
// * n = 2^a 3^b 5^c 7^d <= 8192: mixed-radix Stockham autosort passes in
//   LDS (radix 4 while 4 divides what is left, then 2, 3, 5, 7).  A pass of
//   radix R stages its inputs in registers, synchronises, then writes its
//   outputs, so one LDS buffer of C * n values suffices.
// * any other n <= 4096: Bluestein's algorithm.  With c_j = exp(sign i pi j^2 / n),
//   X_k = c_k sum_j (x_j c_j) conj(c_{k-j}): the line times the chirp, zero
//   padded to a power of two M >= 2n - 1, is convolved with conj(c) by an
//   M-point forward transform, a product with the transform of conj(c)
//   (tabulated once per length and sign), and an M-point inverse transform,
//   all in LDS; the outputs are multiplied by the chirp and 1/M.
//
// Tiles: 2048 values (32 KB of LDS, 5 waves per SIMD) for lines of up to
// 2048 points, 4096 and 8192 for longer ones.  Global access: along the
// contiguous z axis each line is a contiguous run; along y and x the tile's
// C lines are C adjacent columns, so every tile row is C*16 contiguous bytes
// (C = 8 -> 128 B rows for 256-point lines).  Twiddles and chirps come from
// per-length tables computed on the host in long double; the power-of-two
// kernels of the 2048 tile copy theirs into LDS behind the tile.
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "hamx_internal.h"

namespace hamx {

constexpr uint32_t kMaxPass = 16;
constexpr uint32_t kTileMini = 2048;  // complex128 per workgroup: 32 KB of LDS
constexpr uint32_t kTileSmall = 4096; // 64 KB: lines of 2049 to 4096 points
constexpr uint32_t kTileLarge = 8192; // 128 KB: lines of 4097 to 8192 points

struct FftPlan {
  uint32_t n;     // transform length
  uint32_t npass; // Stockham passes
  uint8_t radix[kMaxPass];
};

struct FftAxis {
  uint64_t s;      // element stride of the transformed axis
  uint64_t inner;  // lines with consecutive base addresses
  uint64_t outer;  // base-address step between groups of `inner` lines
  uint64_t lines;  // lines along this axis
  uint32_t n;      // line length
  uint32_t M;      // 0: direct; else the Bluestein convolution length
  uint32_t C;      // lines per workgroup
  // addressing of a tile's C lines: 1 = one contiguous run of C n elements
  // (the contiguous axis); 2 = C adjacent columns (C a power of two, 2^cshift,
  // dividing `inner`: a tile never straddles two groups of `inner` lines);
  // 0 = per-element line bases
  uint32_t fast, cshift;
  FftPlan plan;    // passes of n (direct) or of M (Bluestein)
};

__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

// exp(SIGN 2 pi i j / n) from the table of exp(+2 pi i m / n)
template <int SIGN>
__device__ inline double2 twiddle(const double2 *tw, uint32_t j) {
  double2 w = tw[j];
  if (SIGN < 0)
    w.y = -w.y;
  return w;
}

// y_t = sum_m x_m W^(t m), W = exp(SIGN 2 pi i / R); w[t] = W^t.
template <int R, int SIGN>
__device__ inline void small_dft(const double2 (&x)[R], const double2 (&w)[R], double2 (&y)[R]) {
  if constexpr (R == 2) {
    y[0] = cadd(x[0], x[1]);
    y[1] = csub(x[0], x[1]);
  } else if constexpr (R == 4) {
    const double2 a0 = cadd(x[0], x[2]), a1 = csub(x[0], x[2]), a2 = cadd(x[1], x[3]);
    const double2 d13 = csub(x[1], x[3]);
    // (x1 - x3) * SIGN * i
    const double2 a3 = SIGN > 0 ? make_double2(-d13.y, d13.x) : make_double2(d13.y, -d13.x);
    y[0] = cadd(a0, a2);
    y[1] = cadd(a1, a3);
    y[2] = csub(a0, a2);
    y[3] = csub(a1, a3);
  } else {
#pragma unroll
    for (int t = 0; t < R; ++t) {
      double2 acc = x[0];
#pragma unroll
      for (int m = 1; m < R; ++m)
        acc = cadd(acc, cmul(x[m], w[(t * m) % R]));
      y[t] = acc;
    }
  }
}

// One radix-R Stockham pass over C lines of length n in LDS: p is the
// product of the radices before it, tw the table of exp(2 pi i m / n).
// Line c's point t sits at lds[c n + t] (blocked) or, when IL, at
// lds[t C + c] (interleaved, C = 2^cshift): the strided axes load C adjacent
// columns a row at a time, and with the interleaved layout those rows, and
// every pass's butterfly operands (consecutive threads take consecutive
// lines), are contiguous runs of LDS instead of addresses n apart in one bank.
template <int R, int SIGN, uint32_t TILE, bool IL>
__device__ inline void stockham_pass(double2 *lds, uint32_t n, uint32_t C, uint32_t cshift,
                                     uint32_t p, const double2 *__restrict__ tw) {
  constexpr uint32_t NB = (TILE / R + 255) / 256; // butterflies per thread
  const uint32_t q = n / R, nbf = C * q, tstep = n / (R * p);
  const uint32_t ts = IL ? C : 1u; // distance between a line's points
  double2 x[NB][R];
#pragma unroll
  for (uint32_t u = 0; u < NB; ++u) {
    const uint32_t b = threadIdx.x + u * 256;
    if (b < nbf) {
      const uint32_t c = IL ? (b & (C - 1)) : b / q, i = IL ? (b >> cshift) : b - c * q;
      const double2 *line = lds + (IL ? c : c * n);
#pragma unroll
      for (int m = 0; m < R; ++m)
        x[u][m] = line[(i + m * q) * ts];
    }
  }
  __syncthreads();
  double2 w[R];
#pragma unroll
  for (int t = 0; t < R; ++t)
    w[t] = twiddle<SIGN>(tw, t * q); // exp(SIGN 2 pi i t / R)
#pragma unroll
  for (uint32_t u = 0; u < NB; ++u) {
    const uint32_t b = threadIdx.x + u * 256;
    if (b < nbf) {
      const uint32_t c = IL ? (b & (C - 1)) : b / q, i = IL ? (b >> cshift) : b - c * q;
      const uint32_t k = i % p;
      double2 xs[R], y[R];
      xs[0] = x[u][0];
#pragma unroll
      for (int m = 1; m < R; ++m)
        xs[m] = cmul(x[u][m], twiddle<SIGN>(tw, m * k * tstep));
      small_dft<R, SIGN>(xs, w, y);
      double2 *line = lds + (IL ? c : c * n);
      const uint32_t j = (i - k) * R + k;
#pragma unroll
      for (int m = 0; m < R; ++m)
        line[(j + m * p) * ts] = y[m];
    }
  }
  __syncthreads();
}

// Every pass of a plan, in order.  Entered and left with LDS synchronised.
// POW2 kernels (plans of radix 4 and 2 only) leave the odd radices out, so
// their register staging stays at the radix-4 / radix-2 size.
template <int SIGN, uint32_t TILE, bool POW2, bool IL>
__device__ inline void run_plan(double2 *lds, const FftPlan &pl, uint32_t C, uint32_t cshift,
                                const double2 *__restrict__ tw) {
  uint32_t p = 1;
  for (uint32_t s = 0; s < pl.npass; ++s) {
    const uint32_t r = pl.radix[s];
    if (r == 4) {
      stockham_pass<4, SIGN, TILE, IL>(lds, pl.n, C, cshift, p, tw);
    } else if (r == 2) {
      stockham_pass<2, SIGN, TILE, IL>(lds, pl.n, C, cshift, p, tw);
    } else if constexpr (!POW2) {
      if (r == 3)
        stockham_pass<3, SIGN, TILE, IL>(lds, pl.n, C, cshift, p, tw);
      else if (r == 5)
        stockham_pass<5, SIGN, TILE, IL>(lds, pl.n, C, cshift, p, tw);
      else
        stockham_pass<7, SIGN, TILE, IL>(lds, pl.n, C, cshift, p, tw);
    }
    p *= r;
  }
}

// Tile element e -> (line c of the tile, position t on the line): along the
// contiguous axis a line is a run; along strided axes the C lines of a tile
// are adjacent columns, so consecutive threads take consecutive columns.
__device__ inline void tile_coords(const FftAxis &ax, uint32_t len, uint32_t e, uint32_t &c,
                                   uint32_t &t) {
  if (ax.s == 1) {
    c = e / len;
    t = e - c * len;
  } else {
    c = e % ax.C;
    t = e / ax.C;
  }
}

__device__ inline uint64_t line_base(const FftAxis &ax, uint64_t L) {
  return (L / ax.inner) * ax.outer + (L % ax.inner);
}

template <int SIGN, uint32_t TILE, bool POW2>
__global__ __launch_bounds__(256) void k_fft_axis(double2 *__restrict__ data0,
                                                  double2 *__restrict__ data1, FftAxis ax,
                                                  const double2 *__restrict__ tw) {
  // blockIdx.y picks one of two arrays transformed in the same launch
  double2 *__restrict__ data = blockIdx.y ? data1 : data0;
  // power-of-two lines in the mini tile: the twiddle table is copied into
  // LDS after the tile (dynamic LDS of C n + n values), so the passes read
  // it with LDS latency (the mixed-radix kernels keep it in memory: their
  // extra registers would cost a wave per SIMD)
  constexpr bool kLdsTw = POW2 && TILE <= kTileMini;
  extern __shared__ double2 dyn_lds[];
  __shared__ double2 static_lds[kLdsTw ? 1 : TILE];
  double2 *lds = kLdsTw ? dyn_lds : static_lds;
  const uint32_t n = ax.n, C = ax.C, total = C * n;
  const uint64_t L0 = static_cast<uint64_t>(blockIdx.x) * C;
  // the tile's first element: one line_base (and its integer divisions) per
  // workgroup on the two fast paths instead of one per element
  const uint64_t base0 = ax.fast ? line_base(ax, L0) : 0;
  const uint64_t left = ax.lines - L0; // lines of this tile that exist (>= 1)
  // the fast paths issue all of a thread's loads before the first wait
  // (a loop over the runtime tile size waited for each load before the next
  // one, leaving one 16-byte load in flight per thread); 256 threads
  // (in groups of 8, the 2048-point tile's whole share)
  constexpr uint32_t kPer = TILE / 256, kGroup = kPer < 8 ? kPer : 8;
  if (ax.fast == 1) {
    const uint32_t valid = static_cast<uint32_t>(left < C ? left : C) * n;
    for (uint32_t g = 0; g < kPer; g += kGroup) {
      double2 v[kGroup];
#pragma unroll
      for (uint32_t k = 0; k < kGroup; ++k) {
        const uint32_t e = threadIdx.x + 256 * (g + k);
        v[k] = e < valid ? data[base0 + e] : make_double2(0., 0.);
      }
#pragma unroll
      for (uint32_t k = 0; k < kGroup; ++k) {
        const uint32_t e = threadIdx.x + 256 * (g + k);
        if (e < total)
          lds[e] = v[k];
      }
    }
  } else if (ax.fast == 2) {
    for (uint32_t g = 0; g < kPer; g += kGroup) {
      double2 v[kGroup];
#pragma unroll
      for (uint32_t k = 0; k < kGroup; ++k) {
        const uint32_t e = threadIdx.x + 256 * (g + k);
        const uint32_t c = e & (C - 1), t = e >> ax.cshift;
        v[k] = (e < total && c < left) ? data[base0 + c + static_cast<uint64_t>(t) * ax.s]
                                       : make_double2(0., 0.);
      }
#pragma unroll
      for (uint32_t k = 0; k < kGroup; ++k) {
        const uint32_t e = threadIdx.x + 256 * (g + k);
        if (e < total)
          lds[e] = v[k]; // interleaved: e = t C + c
      }
    }
  } else {
    for (uint32_t e = threadIdx.x; e < total; e += blockDim.x) {
      uint32_t c, t;
      tile_coords(ax, n, e, c, t);
      const uint64_t L = L0 + c;
      lds[c * n + t] = L < ax.lines ? data[line_base(ax, L) + static_cast<uint64_t>(t) * ax.s]
                                    : make_double2(0., 0.);
    }
  }
  const double2 *twl = tw;
  if constexpr (kLdsTw) {
    double2 *t = lds + total;
    for (uint32_t e = threadIdx.x; e < n; e += blockDim.x)
      t[e] = tw[e];
    twl = t;
  }
  __syncthreads();
  if (ax.fast == 2)
    run_plan<SIGN, TILE, POW2, true>(lds, ax.plan, C, ax.cshift, twl);
  else
    run_plan<SIGN, TILE, POW2, false>(lds, ax.plan, C, 0u, twl);
  if (ax.fast == 1) {
    const uint32_t valid = static_cast<uint32_t>(left < C ? left : C) * n;
    for (uint32_t e = threadIdx.x; e < valid; e += blockDim.x)
      data[base0 + e] = lds[e];
  } else if (ax.fast == 2) {
    for (uint32_t e = threadIdx.x; e < total; e += blockDim.x) {
      const uint32_t c = e & (C - 1), t = e >> ax.cshift;
      if (c < left)
        data[base0 + c + static_cast<uint64_t>(t) * ax.s] = lds[e];
    }
  } else {
    for (uint32_t e = threadIdx.x; e < total; e += blockDim.x) {
      uint32_t c, t;
      tile_coords(ax, n, e, c, t);
      const uint64_t L = L0 + c;
      if (L < ax.lines)
        data[line_base(ax, L) + static_cast<uint64_t>(t) * ax.s] = lds[c * n + t];
    }
  }
}

// Bluestein: chirp[j] = exp(SIGN i pi (j^2 mod 2n) / n) for j < n; bhat the
// M-point forward transform of conj(chirp) laid out circularly; twm the table
// of exp(2 pi i m / M).
template <int SIGN, uint32_t TILE>
__global__ __launch_bounds__(256) void k_fft_axis_bluestein(double2 *__restrict__ data0,
                                                            double2 *__restrict__ data1,
                                                            FftAxis ax,
                                                            const double2 *__restrict__ twm,
                                                            const double2 *__restrict__ chirp,
                                                            const double2 *__restrict__ bhat) {
  double2 *__restrict__ data = blockIdx.y ? data1 : data0;
  __shared__ double2 lds[TILE];
  const uint32_t n = ax.n, M = ax.M, C = ax.C, total = C * M;
  const uint64_t L0 = static_cast<uint64_t>(blockIdx.x) * C;
  for (uint32_t e = threadIdx.x; e < total; e += blockDim.x) {
    uint32_t c, t;
    tile_coords(ax, M, e, c, t);
    const uint64_t L = L0 + c;
    double2 v = make_double2(0., 0.);
    if (t < n && L < ax.lines)
      v = cmul(data[line_base(ax, L) + static_cast<uint64_t>(t) * ax.s], chirp[t]);
    lds[c * M + t] = v;
  }
  __syncthreads();
  run_plan<-1, TILE, true, false>(lds, ax.plan, C, 0u, twm);
  for (uint32_t e = threadIdx.x; e < total; e += blockDim.x)
    lds[e] = cmul(lds[e], bhat[e % M]);
  __syncthreads();
  run_plan<1, TILE, true, false>(lds, ax.plan, C, 0u, twm);
  const double inv_m = 1.0 / static_cast<double>(M);
  for (uint32_t e = threadIdx.x; e < total; e += blockDim.x) {
    uint32_t c, t;
    tile_coords(ax, M, e, c, t);
    const uint64_t L = L0 + c;
    if (t < n && L < ax.lines) {
      const double2 v = cmul(lds[c * M + t], chirp[t]);
      data[line_base(ax, L) + static_cast<uint64_t>(t) * ax.s] =
          make_double2(v.x * inv_m, v.y * inv_m);
    }
  }
}

// ------------------------------------------------------------------ host --
namespace {

using cld = std::complex<long double>;
const long double kTwoPiL = 6.283185307179586476925286766559005768394L;

// Radices of a Stockham plan for n: 4 while 4 divides, then 2, 3, 5, 7.
bool make_plan(uint32_t n, FftPlan &pl) {
  pl.n = n;
  pl.npass = 0;
  uint32_t m = n;
  auto push = [&](uint32_t r) {
    if (pl.npass >= kMaxPass)
      return false;
    pl.radix[pl.npass++] = static_cast<uint8_t>(r);
    m /= r;
    return true;
  };
  while (m % 4 == 0)
    if (!push(4))
      return false;
  if (m % 2 == 0 && !push(2))
    return false;
  for (uint32_t r : {3u, 5u, 7u})
    while (m % r == 0)
      if (!push(r))
        return false;
  return m == 1;
}

// Device tables cached per context: kind 0 = twiddles exp(2 pi i m / n),
// kind 1 / 2 = Bluestein chirp / kernel transform of sign (kind - 1.5 > 0).
std::map<std::tuple<hamx_ctx *, uint64_t, int>, double2 *> g_tables;

int upload(hamx_ctx *ctx, const std::vector<double2> &h, std::tuple<hamx_ctx *, uint64_t, int> key,
           const double2 **out) {
  double2 *d = nullptr;
  HAMX_HIP(ctx, hipMalloc(&d, h.size() * sizeof(double2)));
  HAMX_HIP(ctx, hipMemcpy(d, h.data(), h.size() * sizeof(double2), hipMemcpyHostToDevice));
  g_tables[key] = d;
  *out = d;
  return HAMX_OK;
}

int twiddle_table(hamx_ctx *ctx, uint64_t n, const double2 **out) {
  auto key = std::make_tuple(ctx, n, 0);
  auto it = g_tables.find(key);
  if (it != g_tables.end()) {
    *out = it->second;
    return HAMX_OK;
  }
  std::vector<double2> h(n);
  for (uint64_t m = 0; m < n; ++m) {
    const long double a = kTwoPiL * static_cast<long double>(m) / static_cast<long double>(n);
    h[m] = make_double2(static_cast<double>(cosl(a)), static_cast<double>(sinl(a)));
  }
  return upload(ctx, h, key, out);
}

// In-place radix-2 transform of a power-of-two length, exponent sign (host,
// long double; for the Bluestein kernel table only).
void host_fft_pow2(std::vector<cld> &a, int sign) {
  const size_t n = a.size();
  for (size_t i = 1, j = 0; i < n; ++i) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1)
      j ^= bit;
    j ^= bit;
    if (i < j)
      std::swap(a[i], a[j]);
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    for (size_t k = 0; k < len / 2; ++k) {
      const long double ang = sign * kTwoPiL * static_cast<long double>(k) /
                              static_cast<long double>(len);
      const cld w(cosl(ang), sinl(ang));
      for (size_t s = 0; s < n; s += len) {
        const cld u = a[s + k], v = a[s + k + len / 2] * w;
        a[s + k] = u + v;
        a[s + k + len / 2] = u - v;
      }
    }
  }
}

int bluestein_tables(hamx_ctx *ctx, uint64_t n, uint64_t M, int sign, const double2 **chirp,
                     const double2 **bhat) {
  const int kc = sign > 0 ? 2 : 1;
  auto kchirp = std::make_tuple(ctx, n, 2 * kc), kbhat = std::make_tuple(ctx, n, 2 * kc + 1);
  auto ic = g_tables.find(kchirp), ib = g_tables.find(kbhat);
  if (ic != g_tables.end() && ib != g_tables.end()) {
    *chirp = ic->second;
    *bhat = ib->second;
    return HAMX_OK;
  }
  // c_j = exp(sign i pi (j^2 mod 2n) / n): the reduction keeps the angle exact
  std::vector<cld> c(n);
  for (uint64_t j = 0; j < n; ++j) {
    const uint64_t r = (j * j) % (2 * n);
    const long double a = sign * (kTwoPiL / 2) * static_cast<long double>(r) /
                          static_cast<long double>(n);
    c[j] = cld(cosl(a), sinl(a));
  }
  std::vector<cld> b(M, cld(0, 0));
  b[0] = std::conj(c[0]);
  for (uint64_t j = 1; j < n; ++j)
    b[j] = b[M - j] = std::conj(c[j]);
  host_fft_pow2(b, -1);
  std::vector<double2> hc(n), hb(M);
  for (uint64_t j = 0; j < n; ++j)
    hc[j] = make_double2(static_cast<double>(c[j].real()), static_cast<double>(c[j].imag()));
  for (uint64_t j = 0; j < M; ++j)
    hb[j] = make_double2(static_cast<double>(b[j].real()), static_cast<double>(b[j].imag()));
  HAMX_TRY(upload(ctx, hc, kchirp, chirp));
  return upload(ctx, hb, kbhat, bhat);
}

uint64_t next_pow2(uint64_t v) {
  uint64_t p = 1;
  while (p < v)
    p <<= 1;
  return p;
}

} // namespace

void release_twiddles(hamx_ctx *ctx) {
  for (auto it = g_tables.begin(); it != g_tables.end();) {
    if (std::get<0>(it->first) == ctx) {
      (void)hipFree(it->second);
      it = g_tables.erase(it);
    } else {
      ++it;
    }
  }
}

bool fft_length_ok(uint64_t n) {
  FftPlan pl;
  if (n >= 1 && n <= kTileLarge && make_plan(static_cast<uint32_t>(n), pl))
    return true;
  return n >= 1 && n <= kTileSmall;
}

int fft3d_dev(hamx_ctx *ctx, double2 *data, uint64_t nx, uint64_t ny,
              uint64_t nz, int sign) {
  return fft3d_dev2(ctx, data, nullptr, nx, ny, nz, sign);
}

int fft3d_dev2(hamx_ctx *ctx, double2 *data, double2 *data2, uint64_t nx, uint64_t ny,
               uint64_t nz, int sign) {
  const unsigned arrays = data2 ? 2u : 1u;
  const uint64_t dims[3] = {nx, ny, nz};
  for (int a = 0; a < 3; ++a)
    if (!fft_length_ok(dims[a]))
      return set_err(ctx, HAMX_ERR_UNSUPPORTED,
                     "FFT axis length %llu: lengths of the form 2^a 3^b 5^c 7^d up to %u "
                     "and any other length up to %u are supported",
                     static_cast<unsigned long long>(dims[a]), kTileLarge, kTileSmall);
  const uint64_t total = nx * ny * nz;
  // axis order z, y, x (any order gives the same transform)
  for (int a = 2; a >= 0; --a) {
    const uint64_t n = dims[a];
    if (n < 2)
      continue;
    FftAxis ax;
    std::memset(&ax, 0, sizeof(ax));
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
    ax.lines = total / n;
    const bool direct = make_plan(ax.n, ax.plan);
    uint64_t len = n;
    if (!direct) {
      ax.M = static_cast<uint32_t>(next_pow2(2 * n - 1));
      len = ax.M;
      make_plan(ax.M, ax.plan);
    }
    const uint32_t tile = len <= kTileMini ? kTileMini : (len <= kTileSmall ? kTileSmall : kTileLarge);
    // lines per workgroup: fill the tile; for strided axes the lines are
    // adjacent columns (at most 16, 256 B rows)
    uint64_t C = tile / len;
    if (ax.s != 1 && C > 16)
      C = 16;
    if (ax.s != 1 && C > ax.inner)
      C = ax.inner;
    if (C > ax.lines)
      C = ax.lines;
    ax.C = static_cast<uint32_t>(C);
    if (ax.s == 1 && ax.inner == 1 && ax.outer == n) {
      ax.fast = 1;
    } else if (ax.s != 1 && is_pow2(C) && ax.inner % C == 0) {
      ax.fast = 2;
      while ((1ull << ax.cshift) < C)
        ++ax.cshift;
    }
    const unsigned blocks = static_cast<unsigned>((ax.lines + C - 1) / C);
    if (direct) {
      const double2 *tw;
      HAMX_TRY(twiddle_table(ctx, n, &tw));
      const bool pow2 = is_pow2(n);
#define HAMX_FFT_LAUNCH(S, T, P)                                                           \
  hipLaunchKernelGGL((k_fft_axis<S, T, P>), dim3(blocks, arrays), dim3(256),               \
                     (P) && (T) <= kTileMini ? (ax.C + 1) * ax.n * sizeof(double2) : 0,        \
                     ctx->stream,                                                          \
                     data, data2, ax, tw)
#define HAMX_FFT_LAUNCH_T(S, T)                                                            \
  do {                                                                                     \
    if (pow2)                                                                              \
      HAMX_FFT_LAUNCH(S, T, true);                                                         \
    else                                                                                   \
      HAMX_FFT_LAUNCH(S, T, false);                                                        \
  } while (0)
#define HAMX_FFT_LAUNCH_S(S)                                                               \
  do {                                                                                     \
    if (tile == kTileMini)                                                                 \
      HAMX_FFT_LAUNCH_T(S, kTileMini);                                                     \
    else if (tile == kTileSmall)                                                           \
      HAMX_FFT_LAUNCH_T(S, kTileSmall);                                                    \
    else                                                                                   \
      HAMX_FFT_LAUNCH_T(S, kTileLarge);                                                    \
  } while (0)
      if (sign > 0)
        HAMX_FFT_LAUNCH_S(1);
      else
        HAMX_FFT_LAUNCH_S(-1);
#undef HAMX_FFT_LAUNCH_S
#undef HAMX_FFT_LAUNCH_T
#undef HAMX_FFT_LAUNCH
    } else {
      const double2 *twm, *chirp, *bhat;
      HAMX_TRY(twiddle_table(ctx, ax.M, &twm));
      HAMX_TRY(bluestein_tables(ctx, n, ax.M, sign, &chirp, &bhat));
#define HAMX_FFT_LAUNCH(S, T)                                                              \
  hipLaunchKernelGGL((k_fft_axis_bluestein<S, T>), dim3(blocks, arrays), dim3(256), 0,     \
                     ctx->stream, data, data2, ax, twm, chirp, bhat)
#define HAMX_FFT_LAUNCH_S(S)                                                               \
  do {                                                                                     \
    if (tile == kTileMini)                                                                 \
      HAMX_FFT_LAUNCH(S, kTileMini);                                                       \
    else if (tile == kTileSmall)                                                           \
      HAMX_FFT_LAUNCH(S, kTileSmall);                                                      \
    else                                                                                   \
      HAMX_FFT_LAUNCH(S, kTileLarge);                                                      \
  } while (0)
      if (sign > 0)
        HAMX_FFT_LAUNCH_S(1);
      else
        HAMX_FFT_LAUNCH_S(-1);
#undef HAMX_FFT_LAUNCH_S
#undef HAMX_FFT_LAUNCH
    }
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
