// xg_moment.hip -- fused cut rule + element matrices in moment form.
//
// Same quadrature as k_pipeline (xg_elmat.hip): for every (element, domain
// type) the straight cut rule is built in FP64 (cut points
// straightcutrule.cpp:50, decomposition :131-196, sub-simplex volumes
// :66-72, interface factor :558-578) and mapped point by point
// (x = ow V0 + sum xi_m V_{m+1}, :80-100).  Instead of the B^T W B
// contraction of symboliccutbfi.cpp:229-287, each lane accumulates the
// homogeneous barycentric moments m_alpha = sum_i w_i lam(x_i)^alpha,
// |alpha| = 2P, in FP64, and the mass and stiffness entries are formed from
// them by the generated identities of moments_table.h:
//
//     M_ab = sum_alpha C^M_ab,alpha m_alpha
//     K_ab = sum_{d<=e} H_de sum_beta C^K_ab,de,beta mk_beta,  H = F^{-1} F^{-T}
//
// This is the same quadrature sum, reassociated through the polynomial
// identities phi_a phi_b = sum C lam^alpha, which hold at every point; it
// holds for any rule, not only exact ones.  All arithmetic is FP64 (CDNA4
// issues v_fma_f64 at the FP32 rate), the output is cast to T at the store.
//
// Mapping: one lane per (element, domain type); a wave holds 64 elements of
// one domain type, so the number of base-rule points per sub-simplex is
// wave-uniform and the base rule is read with scalar loads.  No LDS, no
// barriers: occupancy is set by the VGPR count of the moment accumulators.
#include "xg_device.hpp"
#include "xg_elmat_core.hpp"
#include "xg_internal.hpp"
#include "xg_moment_core.hpp"

using namespace xg;

namespace {

using namespace xg::mcore;

template <int D, int P, typename T>
__global__ __launch_bounds__(256) void k_moment_pipeline(int nel, const int32_t *__restrict__ elems,
                                                         const int32_t *__restrict__ el2v,
                                                         const double *__restrict__ coords,
                                                         const double *__restrict__ lset, int order, int ndt,
                                                         int dts, int mask, T *__restrict__ out,
                                                         int32_t *err) {
  constexpr int N = ndof(D, P);
  extern __shared__ char smem[];
  const MomLds S = stage_lds<D>(smem, order);  // block barrier: before any exit
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int ntiles = (nel + 63) >> 6;
  const int t = wid / ntiles;
  if (t >= ndt) return;
  const int i = (wid - t * ntiles) * 64 + lane;
  if (i >= nel) return;  // no cross-lane operations below
  const int dt = (dts >> (2 * t)) & 3;
  const int e = elems ? elems[i] : i;
  double V[D + 1][3], l[D + 1];
  ecore::load_geom<D>(el2v, coords, lset, e, V, l);
  double F[D * D], Finv[D * D], absdet;
  jacobian<D>(V, F, Finv, absdet);
  if (mcore::moment_elmat<D, P, T>(S, l, Finv, absdet, dt, mask, out + ((size_t)i * ndt + t) * (size_t)(N * N)))
    atomicOr(err, 1);
}

// element matrices from given moments (one lane per element): the rule is
// whatever produced the moments (subdivision rule, learned rule)
template <int D, int P, typename T>
__global__ __launch_bounds__(256) void k_elmat_from_moments(int nel, const int32_t *__restrict__ elems,
                                                            const int32_t *__restrict__ el2v,
                                                            const double *__restrict__ coords,
                                                            const double *__restrict__ mom, int mask,
                                                            T *__restrict__ out) {
  using MM = Moments<D, P>;
  constexpr int N = ndof(D, P);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nel) return;
  double Finv[D * D], absdet = 1.0;
#pragma unroll
  for (int k = 0; k < D * D; ++k) Finv[k] = (k % (D + 1) == 0) ? 1.0 : 0.0;
  if (coords) {
    const int e = elems ? elems[i] : i;
    double V[D + 1][3], F[D * D];
    ecore::load_geom<D>(el2v, coords, nullptr, e, V, nullptr);
    jacobian<D>(V, F, Finv, absdet);
  }
  double m[MM::NM];
#pragma unroll
  for (int k = 0; k < MM::NM; ++k) m[k] = mom[(size_t)i * MM::NM + k];
  double mk[MM::NK];
  MM::lower(m, mk);
  double Hs[D * (D + 1) / 2];
  int h = 0;
#pragma unroll
  for (int d = 0; d < D; ++d)
#pragma unroll
    for (int q = d; q < D; ++q) {
      double acc = 0.0;
#pragma unroll
      for (int r = 0; r < D; ++r) acc = fma(Finv[d * D + r], Finv[q * D + r], acc);
      Hs[h++] = acc;
    }
  const double sm = (mask & 1) ? absdet : 0.0, sk = (mask & 2) ? absdet : 0.0;
  MM::template forms<T>(m, mk, Hs, sm, sk, out + (size_t)i * (N * N));
}

template <int D, int P, typename T>
int launch_from_moments(xg_ctx *ctx, int nel, const int32_t *elems, const int32_t *el2v, const double *coords,
                        const double *mom, int mask, void *out) {
  const int blocks = (nel + 255) / 256;
  k_elmat_from_moments<D, P, T><<<blocks, 256, 0, ctx->stream>>>(nel, elems, el2v, coords, mom, mask, (T *)out);
  XG_LAUNCH_CHECK();
  return XG_OK;
}

template <int D, int P, typename T>
int launch(xg_ctx *ctx, int order, int nel, const int32_t *elems, const int32_t *el2v, const double *coords,
           const double *lset, int ndt, int dts, int mask, void *out) {
  const long long waves = (long long)ndt * ((nel + 63) / 64);
  const long long blocks = (waves + 3) / 4;
  if (blocks > 0x7fffffffLL) return set_error(XG_ERR_ARG, "xg_cut_pipeline: too many elements");
  k_moment_pipeline<D, P, T><<<(unsigned)blocks, 256, mom_lds_bytes(D, order), ctx->stream>>>(nel, elems, el2v, coords, lset, order,
                                                                      ndt, dts, mask, (T *)out, ctx->d_err);
  XG_LAUNCH_CHECK();
  return XG_OK;
}

}  // namespace

namespace xg {

int launch_moment_pipeline(xg_ctx *ctx, int D, int p, int order, int nel, const int32_t *elems,
                           const int32_t *el2v, const double *coords, const double *lset, int ndt, int dts,
                           int mask, bool fp64_out, void *out) {
#define XG_MCALL(DD, PP)                                                                                   \
  if (D == DD && p == PP)                                                                                  \
    return fp64_out ? launch<DD, PP, double>(ctx, order, nel, elems, el2v, coords, lset, ndt, dts, mask, out) \
                    : launch<DD, PP, float>(ctx, order, nel, elems, el2v, coords, lset, ndt, dts, mask, out);
  XG_MCALL(1, 1) XG_MCALL(1, 2) XG_MCALL(1, 3)
  XG_MCALL(2, 1) XG_MCALL(2, 2) XG_MCALL(2, 3)
  XG_MCALL(3, 1) XG_MCALL(3, 2) XG_MCALL(3, 3)
#undef XG_MCALL
  return set_error(XG_ERR_ARG, "xg_cut_pipeline: bad D or p");
}

}  // namespace xg

extern "C" int xg_elmat_moments(xg_ctx *ctx, int D, int p, int nel, const int32_t *elems, const int32_t *el2v,
                                const double *coords, const double *moments, int mask, int precision, void *out) {
  if (!ctx || D < 1 || D > 3 || p < 1 || p > 3 || nel < 0 || mask < 1 || mask > 3)
    return set_error(XG_ERR_ARG, "xg_elmat_moments: bad argument");
  if (precision != XG_PREC_FP32 && precision != XG_PREC_FP64)
    return set_error(XG_ERR_UNSUPPORTED, "xg_elmat_moments: FP32 or FP64 output");
  if (nel == 0) return XG_OK;
  if (!moments || !out || (coords && !el2v)) return set_error(XG_ERR_ARG, "xg_elmat_moments: null pointer");
  const bool f64 = precision == XG_PREC_FP64;
#define XG_ECALL(DD, PP)                                                                          \
  if (D == DD && p == PP)                                                                         \
    return f64 ? launch_from_moments<DD, PP, double>(ctx, nel, elems, el2v, coords, moments, mask, out) \
               : launch_from_moments<DD, PP, float>(ctx, nel, elems, el2v, coords, moments, mask, out);
  XG_ECALL(1, 1) XG_ECALL(1, 2) XG_ECALL(1, 3)
  XG_ECALL(2, 1) XG_ECALL(2, 2) XG_ECALL(2, 3)
  XG_ECALL(3, 1) XG_ECALL(3, 2) XG_ECALL(3, 3)
#undef XG_ECALL
  return set_error(XG_ERR_ARG, "xg_elmat_moments: bad D or p");
}
