// xg_internal.hpp -- host-side internals shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/xg_abi.h"

struct xg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // grow-only scratch (block sums, scan partials, staging)
  void *ws = nullptr;
  size_t ws_size = 0;
  // sticky device error word: set by kernels that meet a condition on which
  // the reference throws; read and cleared by xg_sync()
  int32_t *d_err = nullptr;
  int32_t *h_err = nullptr;  // pinned host mirror
  // exact-rule subdivision tables, built on first use: vertices of every
  // refinement-tree node of depth <= 5, and their degree-2p moments (p = 1..3)
  void *sdx_verts = nullptr;
  double *sdx_mom[4] = {nullptr, nullptr, nullptr, nullptr};
  // device address of the no-alloc marking scratch on this device
  int32_t *noalloc_scratch = nullptr;
  // grow-only lists of the staged subdivision kernels (node ids per element)
  void *sdx_ws = nullptr;
  size_t sdx_ws_size = 0;
  // per-kernel timing (xg_timing_enable): event pairs by kernel name
  bool timing = false;
  std::vector<std::string> tnames;
  std::vector<std::vector<std::pair<hipEvent_t, hipEvent_t>>> tevents;
};

namespace xg {

int set_error(int code, const std::string &msg);
int check_hip(hipError_t e, const char *what);
// records a HIP event pair on the context's stream around the launches in
// its scope while the context's timing is enabled
struct KernelTimer {
  xg_ctx *ctx;
  int slot = -1;
  hipEvent_t a = nullptr;
  KernelTimer(xg_ctx *c, const char *name);
  ~KernelTimer();
};
void timing_clear(xg_ctx *ctx);
// make sure ctx->ws holds at least `bytes`
int ensure_ws(xg_ctx *ctx, size_t bytes);
// the same for another grow-only buffer of the context
int ensure_buf(xg_ctx *ctx, void **buf, size_t *size, size_t bytes);
// exclusive scan of n int32 values in place (device), total written to
// data[n] when `with_total` (data must then hold n+1 entries)
int scan_exclusive_i32(xg_ctx *ctx, int32_t *data, int n, bool with_total);
// graph-capturable marking + compaction (kernel launches only)
size_t mark_scratch_ints(int ne);
int bits_and_cut_list_from_dt(xg_ctx *ctx, int ne, const uint8_t *dt, uint32_t *bits, int32_t *cut_list,
                              int32_t *n_cut);
int mark_compact(hipStream_t st, int32_t *d_err, int D, int ne, const int32_t *el2v, const double *lset,
                 uint8_t *dt, double *cr, uint32_t *bits, int32_t *scratch, int32_t *cut_list, int32_t *n_cut);
// fused cut rule + element matrices in moment form (xg_moment.hip); dts packed
// 2 bits per domain type
int launch_moment_pipeline(xg_ctx *ctx, int D, int p, int order, int nel, const int32_t *elems,
                           const int32_t *el2v, const double *coords, const double *lset, int ndt, int dts,
                           int mask, bool fp64_out, void *out);

// device address of the no-alloc marking scratch on `device` (xg_mark.hip)
int noalloc_scratch(int device, int32_t **out);
// degree-p (tetrahedra) mass / stiffness matrices from degree-2p moments
// stored element-major with ndt sets per element (xg_moment.hip)
int elmat_from_moments(xg_ctx *ctx, int p, int nel, int ndt, const int32_t *elems, const int32_t *el2v,
                       const double *coords, const double *mom, int mask, bool fp64_out, void *out);

}  // namespace xg

#define XG_CHECK(expr)                                  \
  do {                                                  \
    hipError_t _e = (expr);                             \
    if (_e != hipSuccess) return xg::check_hip(_e, #expr); \
  } while (0)

#define XG_LAUNCH_CHECK() XG_CHECK(hipGetLastError())
