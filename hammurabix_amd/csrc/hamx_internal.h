// Context, workspace and error plumbing shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "hamx.h"

struct hamx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // kernel timing (hamx_profile)
  bool profiling = false;
  struct Pending {
    hipEvent_t start, stop;
    hamx_kernel_time rec;
  };
  std::vector<Pending> pending;      // recorded, not yet resolved
  std::vector<hipEvent_t> event_pool;
  std::vector<hamx_kernel_time> times;
  // Grow-only device workspace, one buffer per slot.  Slots are owned by the
  // entry point that uses them; a call reuses the previous call's memory.
  static constexpr int kSlots = 56;
  void *ws[kSlots] = {};
  size_t ws_bytes[kSlots] = {};
  // What the tape in workspace slots kSlotAdjCkpt / kSlotTapeFd was recorded
  // for (hamx_los_forward with HAMX_KEEP_TAPE).
  bool tape_valid = false;
  hamx_los_desc tape_desc{};
  hamx_fields tape_fields{};
  // Host copy of small per-call tables uploaded asynchronously; it lives
  // until the next call so the copy never reads freed memory.
  std::vector<double> host_tab;
  // Pinned host staging buffer for host-resident outputs.
  void *pinned = nullptr;
  size_t pinned_bytes = 0;
};

namespace hamx {

// Workspace slot indices.
enum Slot : int {
  kSlotBrndX = 0,
  kSlotBrndY,
  kSlotBrndZ,
  kSlotTernd,
  kSlotBregX,
  kSlotBregY,
  kSlotBregZ,
  kSlotTereg,
  kSlotPacked,
  kSlotTmpDm,
  kSlotTmpFd,
  kSlotTmpIs,
  kSlotTmpQs,
  kSlotTmpUs,
  kSlotOutDm,
  kSlotOutFd,
  kSlotOutIs,
  kSlotOutQs,
  kSlotOutUs,
  kSlotRndC0,
  kSlotRndC1,
  kSlotRndOut0,
  kSlotRndOut1,
  kSlotRndOut2,
  kSlotReduce,
  kSlotFft,
  kSlotAux,
  kSlotAdjCkpt,   // FD map before each shell
  kSlotAdjIs,
  kSlotAdjQs,
  kSlotAdjUs,
  kSlotAdjDm,
  kSlotAdjFd,
  kSlotAdjShell0, // five shell-map adjoints
  kSlotAdjShell1,
  kSlotAdjShell2,
  kSlotAdjShell3,
  kSlotAdjShell4,
  kSlotAdjG,      // FD-map adjoint
  kSlotAdjFdin,   // dL/d(fd cache) per shell pixel
  kSlotAdjRayPar, // per-ray parameter partials
  kSlotAdjPar,    // reduced parameter gradient
  kSlotAdjGrid,   // packed grid gradient
  kSlotTapeFd,    // every ray's FD over each shell (forward tape)
  kSlotMaskPivot, // the caller's mask map (host input staged)
  kSlotMaskShell, // the mask duplicated to the current shell's nside
  kSlotCre,       // the caller's CRE flux grid (host input staged)
  kSlotCreTab,    // energy tables of the CRE grid emissivities
  kSlotSyncIO,    // hamx_sync_kernels staging (x, F, G)
  kSlotCreOut,    // hamx_cre_write_grid output staging
  kSlotFlag,      // hamx_los_forward's non-finite output flag
  kSlotCount
};
static_assert(kSlotCount <= hamx_ctx::kSlots, "workspace slots");

inline int set_err(hamx_ctx *ctx, int code, const char *fmt, ...) {
  if (ctx) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    ctx->err = buf;
  }
  return code;
}

#define HAMX_HIP(ctx, call)                                                    \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      return ::hamx::set_err((ctx), HAMX_ERR_DEVICE, "%s failed: %s (%s:%d)",  \
                             #call, hipGetErrorString(e_), __FILE__,          \
                             __LINE__);                                        \
  } while (0)

#define HAMX_TRY(expr)                                                         \
  do {                                                                         \
    int r_ = (expr);                                                           \
    if (r_ != HAMX_OK)                                                         \
      return r_;                                                               \
  } while (0)

// Device buffer of at least `bytes` in workspace slot `slot`.
int ws_get(hamx_ctx *ctx, int slot, size_t bytes, void **out);

// Make a device-resident copy of a host input (or pass a device pointer
// through when `in_device`).
int stage_in(hamx_ctx *ctx, int slot, const void *src, size_t bytes,
             bool in_device, const void **dev);

// Fill n doubles with v on the context stream (a kernel, so that the call
// can be captured into a HIP graph).
int fill_f64(hamx_ctx *ctx, double *p, size_t n, double v);

inline bool is_pow2(uint64_t n) { return n != 0 && (n & (n - 1)) == 0; }

// True when fft3d_dev transforms lines of length n: 2^a 3^b 5^c 7^d up to
// 8192, any other length up to 4096 (hamx_fft.hip).
bool fft_length_ok(uint64_t n);
// In-place unnormalised 3D c2c FFT of device data (hamx_fft.hip).
int fft3d_dev(hamx_ctx *ctx, double2 *data, uint64_t nx, uint64_t ny,
              uint64_t nz, int sign);
// The same transform of two arrays of one shape, both in each axis launch
// (data2 may be NULL).
int fft3d_dev2(hamx_ctx *ctx, double2 *data, double2 *data2, uint64_t nx, uint64_t ny,
               uint64_t nz, int sign);
// Free the per-context twiddle tables (called by hamx_close).
void release_twiddles(hamx_ctx *ctx);
// Look up, once per device, the device address of hamx_los_forward_static's
// scratch array (called by hamx_open with the context's device current).
int static_scratch_init(hamx_ctx *ctx);

// Kernel timing: prof_begin records a start event before a launch when
// profiling is on; prof_end records the stop event and queues the record,
// resolved (elapsed time read) by prof_resolve after a stream sync.
struct ProfMark {
  int active = 0;
  hipEvent_t start = nullptr;
};
ProfMark prof_begin(hamx_ctx *ctx);
void prof_end(hamx_ctx *ctx, ProfMark &m, const char *kernel, uint32_t shell,
              uint64_t work);
void prof_resolve(hamx_ctx *ctx);

} // namespace hamx
