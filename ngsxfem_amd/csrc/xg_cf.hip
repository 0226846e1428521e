// xg_cf.hip -- coefficient functions at the points of a materialised rule,
// and element vectors on such a rule.
//
// xg_rule_scale_cf multiplies every point's weight by the coefficient's value
// at the physical point (the coefficient is a postfix program, xg_cf.hpp);
// the element matrices (xg_elmat), element vectors (xg_elvec) and integrals
// (xg_integrate with f = 1) of the scaled rule are then those of
// c u v / c grad u . grad v, of f v and of f, as NGSolve evaluates them: the
// coefficient at each mapped integration point of the rule
// (symboliccutbfi.cpp:229-287, symboliccutlfi.cpp:57-140,
// python_cutint.cpp:66-98).
//
// Mapping: one wave per element, lanes over the element's points (the rules
// hold tens of points per cut element); the program is a kernel argument,
// so its words are read through the scalar cache and every lane takes the
// same branch of the interpreter.
#include "xg_device.hpp"
#include "xg_internal.hpp"
#include "xg_cf.hpp"

using namespace xg;

namespace {

template <int D>
__device__ __forceinline__ void element_geometry(const int32_t *__restrict__ el2v, const double *__restrict__ coords,
                                                 int e, int32_t *vv, double (*V)[3]) {
#pragma unroll
  for (int k = 0; k <= D; ++k) {
    vv[k] = el2v[(size_t)e * (D + 1) + k];
#pragma unroll
    for (int d = 0; d < 3; ++d) V[k][d] = d < D ? coords[(size_t)vv[k] * D + d] : 0.0;
  }
}

// w_q <- w_q * cf(x_q) for every point of every listed element; the
// element's vertices have A coordinates (A = D, or D + 1 for boundary
// elements given in the ambient space), x_q = sum_k lam_k X_k
template <int D, int A>
__global__ __launch_bounds__(256) void k_rule_scale_cf(int nel, const int32_t *__restrict__ elems,
                                                       const int32_t *__restrict__ el2v,
                                                       const double *__restrict__ coords,
                                                       const int32_t *__restrict__ offsets, double4 *pts,
                                                       const cf::Prog P) {
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  for (int i = wid; i < nel; i += nwaves) {
    const int e = elems ? elems[i] : i;
    int32_t vv[D + 1];
    double V[D + 1][3];
#pragma unroll
    for (int k = 0; k <= D; ++k) {
      vv[k] = el2v[(size_t)e * (D + 1) + k];
#pragma unroll
      for (int d = 0; d < 3; ++d) V[k][d] = d < A ? coords[(size_t)vv[k] * A + d] : 0.0;
    }
    for (int q = offsets[i] + lane; q < offsets[i + 1]; q += 64) {
      double4 p = pts[q];
      const double xi[3] = {p.x, p.y, p.z};
      // x = V_D + sum_c xi_c (V_c - V_D); lam_c = xi_c, lam_D = 1 - sum xi
      double x[3] = {0.0, 0.0, 0.0}, lam[D + 1];
      double sl = 1.0;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        lam[c] = xi[c];
        sl -= xi[c];
      }
      lam[D] = sl;
#pragma unroll
      for (int r = 0; r < A; ++r) {
        double a = V[D][r];
#pragma unroll
        for (int c = 0; c < D; ++c) a += (V[c][r] - V[D][r]) * xi[c];
        x[r] = a;
      }
      p.w *= cf::eval<D>(P, x, lam, vv, e);
      pts[q] = p;
    }
  }
}

// out[i][a] = |det F| sum_q w_q phi_a(xi_q): element vectors of the source
// f = 1 on the rule (f is already in the weights after xg_rule_scale_cf)
template <int D, int P>
__global__ __launch_bounds__(256) void k_elvec_rule(int nel, const int32_t *__restrict__ elems,
                                                    const int32_t *__restrict__ el2v,
                                                    const double *__restrict__ coords,
                                                    const int32_t *__restrict__ offsets,
                                                    const double4 *__restrict__ pts, double *__restrict__ out) {
  constexpr int N = ndof(D, P);
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  for (int i = wid; i < nel; i += nwaves) {
    const int e = elems ? elems[i] : i;
    int32_t vv[D + 1];
    double V[D + 1][3];
    element_geometry<D>(el2v, coords, e, vv, V);
    double F[D * D], Finv[D * D], absdet;
    jacobian<D>(V, F, Finv, absdet);
    double s[N];
#pragma unroll
    for (int a = 0; a < N; ++a) s[a] = 0.0;
    for (int q = offsets[i] + lane; q < offsets[i + 1]; q += 64) {
      const double4 p = pts[q];
      const double xi[3] = {p.x, p.y, p.z};
      double lam[D + 1], phi[N];
      double sl = 1.0;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        lam[c] = xi[c];
        sl -= xi[c];
      }
      lam[D] = sl;
      h1_shape<D, P, double>(lam, phi, (double(*)[D]) nullptr);
#pragma unroll
      for (int a = 0; a < N; ++a) s[a] = fma(p.w, phi[a], s[a]);
    }
#pragma unroll
    for (int a = 0; a < N; ++a) {
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) s[a] += __shfl_xor(s[a], o, 64);
    }
#pragma unroll
    for (int a = 0; a < N; ++a)
      if (lane == a) out[(size_t)i * N + a] = absdet * s[a];
  }
}

// InterpolateP1::Do (utils/p1interpol.cpp:27-113): the coefficient at every
// vertex, then the perturbation |v| < eps -> eps (:104-107); one lane per
// vertex (a grid function opcode reads the vertex's own value)
template <int D>
__global__ __launch_bounds__(256) void k_p1_interpolate_cf(int nv, const double *__restrict__ coords, const cf::Prog P,
                                                           double eps, double *__restrict__ out) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  double x[3] = {0.0, 0.0, 0.0}, lam[D + 1];
  int32_t vv[D + 1];
#pragma unroll
  for (int d = 0; d < D; ++d) x[d] = coords[(size_t)v * D + d];
#pragma unroll
  for (int k = 0; k <= D; ++k) {
    lam[k] = k == 0 ? 1.0 : 0.0;
    vv[k] = v;
  }
  double val = cf::eval<D>(P, x, lam, vv, -1);
  if (fabs(val) < eps) val = eps;
  out[v] = val;
}

// Element-boundary rules (SymbolicCutBFI with element_boundary,
// symboliccutbfi.cpp:488-640): the cut rules of the D + 1 facets of every
// listed element, computed on a submesh of the facets (facet k of element i
// is submesh element i (D + 1) + k: the element's vertices without vertex k,
// in ascending order), mapped in place into the element's reference
// coordinates (facet barycentrics mu_j -> element barycentrics
// lam_{loc_j}), the weights rescaled from the facet's measure to the
// element's (w |det F_facet| / |det F_element|, the element kernels
// multiply by |det F_element|).  One wave per element, lanes over points;
// eoff[i] = foff[i (D + 1)].
template <int D>
__global__ __launch_bounds__(256) void k_facet_rule_to_element(int nel, const int32_t *__restrict__ elems,
                                                               const int32_t *__restrict__ el2v,
                                                               const double *__restrict__ coords,
                                                               const int32_t *__restrict__ foff, double4 *pts,
                                                               int32_t *__restrict__ eoff) {
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  if (wid == 0 && lane == 0) eoff[nel] = foff[nel * (D + 1)];
  for (int i = wid; i < nel; i += nwaves) {
    const int e = elems ? elems[i] : i;
    if (lane == 0) eoff[i] = foff[i * (D + 1)];
    int32_t vv[D + 1];
    double V[D + 1][3];
    element_geometry<D>(el2v, coords, e, vv, V);
    double F[D * D], Finv[D * D], absdet;
    jacobian<D>(V, F, Finv, absdet);
#pragma unroll
    for (int k = 0; k <= D; ++k) {
      // facet k: the element's vertices without k, ascending
      int loc[D];
      int c = 0;
#pragma unroll
      for (int j = 0; j <= D; ++j)
        if (j != k) loc[c++] = j;
      // |det F_facet|: length (D = 2) or twice the area (D = 3) of the facet
      double fdet = 1.0;
      if constexpr (D == 2) {
        const double a0 = V[loc[1]][0] - V[loc[0]][0], a1 = V[loc[1]][1] - V[loc[0]][1];
        fdet = sqrt(a0 * a0 + a1 * a1);
      } else if constexpr (D == 3) {
        double a[3], b[3], cr[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          a[d] = V[loc[1]][d] - V[loc[0]][d];
          b[d] = V[loc[2]][d] - V[loc[0]][d];
        }
        cr[0] = a[1] * b[2] - a[2] * b[1];
        cr[1] = a[2] * b[0] - a[0] * b[2];
        cr[2] = a[0] * b[1] - a[1] * b[0];
        fdet = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
      }
      const double scale = fdet / absdet;
      const int q0 = foff[i * (D + 1) + k], q1 = foff[i * (D + 1) + k + 1];
      for (int q = q0 + lane; q < q1; q += 64) {
        double4 p = pts[q];
        const double xi[3] = {p.x, p.y, p.z};
        double mu[D], s = 1.0;
#pragma unroll
        for (int j = 0; j < D - 1; ++j) {
          mu[j] = xi[j];
          s -= xi[j];
        }
        mu[D - 1] = s;
        double lam[D + 1];
#pragma unroll
        for (int j = 0; j <= D; ++j) lam[j] = 0.0;
#pragma unroll
        for (int j = 0; j < D; ++j) lam[loc[j]] = mu[j];
        p.x = lam[0];
        p.y = D > 1 ? lam[1] : 0.0;
        p.z = D > 2 ? lam[2] : 0.0;
        p.w *= scale;
        pts[q] = p;
      }
    }
  }
}

int cf_grid(int nel) {
  long long blocks = ((long long)nel * 64 + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  return (int)blocks;
}

}  // namespace

extern "C" int xg_rule_scale_cf(xg_ctx *ctx, int D, int A, int nel, const int32_t *elems, const int32_t *el2v,
                                const double *coords, const int32_t *offsets, double *pts, const xg_cf *cf) {
  if (!ctx || D < 1 || D > 3 || (A != D && A != D + 1) || A > 3 || nel < 0 || !cf)
    return set_error(XG_ERR_ARG, "xg_rule_scale_cf: bad argument");
  cf::Prog P;
  if (const char *msg = cf::validate(cf->ncode, cf->code, cf->nconst, cf->consts, cf->ngf, cf->gf, P))
    return set_error(XG_ERR_ARG, std::string("xg_rule_scale_cf: ") + msg);
  if (nel == 0) return XG_OK;
  if (!el2v || !coords || !offsets || !pts) return set_error(XG_ERR_ARG, "xg_rule_scale_cf: null pointer");
  const int grid = cf_grid(nel);
  double4 *p = reinterpret_cast<double4 *>(pts);
#define XG_SC(DD, AA) k_rule_scale_cf<DD, AA><<<grid, 256, 0, ctx->stream>>>(nel, elems, el2v, coords, offsets, p, P)
  switch (D * 4 + A) {
    case 5: XG_SC(1, 1); break;
    case 6: XG_SC(1, 2); break;
    case 10: XG_SC(2, 2); break;
    case 11: XG_SC(2, 3); break;
    default: XG_SC(3, 3); break;
  }
#undef XG_SC
  XG_LAUNCH_CHECK();
  return XG_OK;
}

extern "C" int xg_p1_interpolate_cf(xg_ctx *ctx, int D, int nv, const double *coords, const xg_cf *cf, double eps,
                                    double *out) {
  if (!ctx || D < 1 || D > 3 || nv < 0 || !cf) return set_error(XG_ERR_ARG, "xg_p1_interpolate_cf: bad argument");
  cf::Prog P;
  if (const char *msg = cf::validate(cf->ncode, cf->code, cf->nconst, cf->consts, cf->ngf, cf->gf, P))
    return set_error(XG_ERR_ARG, std::string("xg_p1_interpolate_cf: ") + msg);
  if (cf::uses_elements(P))
    return set_error(XG_ERR_ARG, "xg_p1_interpolate_cf: piecewise constant grid functions have no vertex values");
  if (nv == 0) return XG_OK;
  if (!coords || !out) return set_error(XG_ERR_ARG, "xg_p1_interpolate_cf: null pointer");
  const int grid = (nv + 255) / 256;
  switch (D) {
    case 1: k_p1_interpolate_cf<1><<<grid, 256, 0, ctx->stream>>>(nv, coords, P, eps, out); break;
    case 2: k_p1_interpolate_cf<2><<<grid, 256, 0, ctx->stream>>>(nv, coords, P, eps, out); break;
    default: k_p1_interpolate_cf<3><<<grid, 256, 0, ctx->stream>>>(nv, coords, P, eps, out); break;
  }
  XG_LAUNCH_CHECK();
  return XG_OK;
}

extern "C" int xg_facet_rule_to_element(xg_ctx *ctx, int D, int nel, const int32_t *elems, const int32_t *el2v,
                                        const double *coords, const int32_t *foffsets, double *pts,
                                        int32_t *eoffsets) {
  if (!ctx || D < 2 || D > 3 || nel < 0 || !foffsets || !eoffsets)
    return set_error(XG_ERR_ARG, "xg_facet_rule_to_element: bad argument");
  if (nel == 0) {
    XG_CHECK(hipMemcpyAsync(eoffsets, foffsets, sizeof(int32_t), hipMemcpyDeviceToDevice, ctx->stream));
    return XG_OK;
  }
  if (!el2v || !coords || !pts) return set_error(XG_ERR_ARG, "xg_facet_rule_to_element: null pointer");
  const int grid = cf_grid(nel);
  double4 *q = reinterpret_cast<double4 *>(pts);
  if (D == 2)
    k_facet_rule_to_element<2><<<grid, 256, 0, ctx->stream>>>(nel, elems, el2v, coords, foffsets, q, eoffsets);
  else
    k_facet_rule_to_element<3><<<grid, 256, 0, ctx->stream>>>(nel, elems, el2v, coords, foffsets, q, eoffsets);
  XG_LAUNCH_CHECK();
  return XG_OK;
}

extern "C" int xg_elvec(xg_ctx *ctx, int D, int p, int nel, const int32_t *elems, const int32_t *el2v,
                        const double *coords, const int32_t *offsets, const double *pts, double *out) {
  if (!ctx || D < 1 || D > 3 || p < 1 || p > 3 || nel < 0) return set_error(XG_ERR_ARG, "xg_elvec: bad argument");
  if (nel == 0) return XG_OK;
  if (!el2v || !coords || !offsets || !pts || !out) return set_error(XG_ERR_ARG, "xg_elvec: null pointer");
  const int grid = cf_grid(nel);
  const double4 *q = reinterpret_cast<const double4 *>(pts);
#define XG_EV(DD, PP) \
  k_elvec_rule<DD, PP><<<grid, 256, 0, ctx->stream>>>(nel, elems, el2v, coords, offsets, q, out)
  switch (D * 4 + p) {
    case 5: XG_EV(1, 1); break;
    case 6: XG_EV(1, 2); break;
    case 7: XG_EV(1, 3); break;
    case 9: XG_EV(2, 1); break;
    case 10: XG_EV(2, 2); break;
    case 11: XG_EV(2, 3); break;
    case 13: XG_EV(3, 1); break;
    case 14: XG_EV(3, 2); break;
    default: XG_EV(3, 3); break;
  }
#undef XG_EV
  XG_LAUNCH_CHECK();
  return XG_OK;
}
