// xg_rule.hip -- materialised straight-cut rules (count + fill) and IntegrateX.
//
// Reference: cutint/straightcutrule.cpp:582-677 (StraightCutIntegrationRule),
//            :80-100 (GetPlainIntegrationRule), :558-578 (interface weights),
//            cutint/python_cutint.cpp:66-98 (IntegrateX sums).
//
// count: one thread per element (topology only, integer work).
// fill : one wavefront per element, lanes = quadrature points; each lane
//        writes one double4 (x, y, z, w), so a wave stores 2 KB contiguously.
//        The decomposition is wave-uniform and recomputed by every lane
//        (a few dozen FP64 operations) instead of being staged through HBM.
#include "xg_device.hpp"
#include "xg_internal.hpp"

using namespace xg;

namespace {

template <int D>
__device__ __forceinline__ void load_elem(const int32_t *__restrict__ el2v, const double *__restrict__ lset,
                                          const double *__restrict__ coords, int e, double *l,
                                          double (*V)[3]) {
#pragma unroll
  for (int k = 0; k <= D; ++k) {
    const int v = el2v[(size_t)e * (D + 1) + k];
    l[k] = lset[v];
    if (V) {
#pragma unroll
      for (int d = 0; d < 3; ++d) V[k][d] = d < D ? coords[(size_t)v * D + d] : 0.0;
    }
  }
}

template <int D>
__global__ void k_rule_count(int nel, const int32_t *__restrict__ elems, const int32_t *__restrict__ el2v,
                             const double *__restrict__ lset, int dt, int order, int32_t *__restrict__ cnt,
                             uint8_t *__restrict__ nsub, int32_t *err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nel) return;
  const int e = elems ? elems[i] : i;
  double l[D + 1];
  load_elem<D>(el2v, lset, nullptr, e, l, nullptr);
  const int eldom = check_if_straight_cut<D>(l);
  int n = 0, ns = 0;
  if (eldom != IF) {
    n = (eldom == dt) ? rule_cnt(D, order) : 0;
  } else {
    Decomp dc;
    if (!decompose<D>(l, dt, dc)) {
      atomicOr(err, 1);
    } else {
      ns = dc.ns;
      n = dc.ns * rule_cnt(dc.nv - 1, order);
    }
  }
  cnt[i] = n;
  if (nsub) nsub[i] = (uint8_t)ns;
}

template <int D>
__global__ __launch_bounds__(256) void k_rule_fill(int nel, const int32_t *__restrict__ elems,
                                                   const int32_t *__restrict__ el2v,
                                                   const double *__restrict__ coords,
                                                   const double *__restrict__ lset, int dt, int order,
                                                   const int32_t *__restrict__ offsets,
                                                   double4 *__restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  for (int i = wid; i < nel; i += nwaves) {
    const int e = elems ? elems[i] : i;
    double l[D + 1], V[D + 1][3];
    load_elem<D>(el2v, lset, coords, e, l, V);
    const int base = offsets[i];
    const int npts = offsets[i + 1] - base;
    if (npts == 0) continue;
    const int eldom = check_if_straight_cut<D>(l);
    if (eldom != IF) {
      const int off = rule_off(D, order);
      for (int q = lane; q < npts; q += 64) {
        const double *r = xg_rule_data_d[off + q];
        out[base + q] = make_double4(r[0], D > 1 ? r[1] : 0.0, D > 2 ? r[2] : 0.0, r[3]);
      }
      continue;
    }
    Decomp dc;
    decompose<D>(l, dt, dc);
    double P[D + 1 + 4][3];
    cut_point_array<D>(l, dc, P);
    const int Ds = dc.nv - 1;
    const int nq = rule_cnt(Ds, order), off = rule_off(Ds, order);
    double ifac = 1.0;
    if (dt == IF) {
      double F[D * D], Finv[D * D], absdet;
      jacobian<D>(V, F, Finv, absdet);
      ifac = interface_factor<D>(l, Finv);
    }
    for (int g = lane; g < npts; g += 64) {
      const int s = g / nq, k = g - s * nq;
      const int8_t *idx = dc.v[s];
      const double tf = simplex_volume(dc.nv, idx, P);
      const double *r = xg_rule_data_d[off + k];
      double ow = 1.0;
      for (int m = 0; m < Ds; ++m) ow -= r[m];
      double x[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) x[d] = ow * P[idx[0]][d];
      for (int m = 0; m < Ds; ++m)
#pragma unroll
        for (int d = 0; d < 3; ++d) x[d] += r[m] * P[idx[m + 1]][d];
      double w = r[3] * tf;
      if (dt == IF) w = w * ifac;
      out[base + g] = make_double4(x[0], x[1], x[2], w);
    }
  }
}

template <int D>
__global__ __launch_bounds__(256) void k_integrate(int nel, const int32_t *__restrict__ elems,
                                                   const int32_t *__restrict__ el2v,
                                                   const double *__restrict__ coords,
                                                   const int32_t *__restrict__ offsets,
                                                   const double4 *__restrict__ pts, Poly f,
                                                   double *__restrict__ per_elem, double *total) {
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  double acc_all = 0.0;
  for (int i = wid; i < nel; i += nwaves) {
    const int e = elems ? elems[i] : i;
    double V[D + 1][3];
#pragma unroll
    for (int k = 0; k <= D; ++k) {
      const int v = el2v[(size_t)e * (D + 1) + k];
#pragma unroll
      for (int d = 0; d < 3; ++d) V[k][d] = d < D ? coords[(size_t)v * D + d] : 0.0;
    }
    double F[D * D], Finv[D * D], absdet;
    jacobian<D>(V, F, Finv, absdet);
    double s = 0.0;
    for (int q = offsets[i] + lane; q < offsets[i + 1]; q += 64) {
      const double4 p = pts[q];
      const double xi[3] = {p.x, p.y, p.z};
      double x[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int r = 0; r < D; ++r) {
        double a = V[D][r];
#pragma unroll
        for (int c = 0; c < D; ++c) a += F[r * D + c] * xi[c];
        x[r] = a;
      }
      double fv = 0.0;
      for (int t = 0; t < f.n; ++t) {
        double m = f.c[t];
        for (int d = 0; d < 3; ++d)
          for (int k = 0; k < f.ex[t][d]; ++k) m *= x[d];
        fv += m;
      }
      s += absdet * p.w * fv;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0 && per_elem) per_elem[i] = s;
    acc_all += s;
  }
  if (lane == 0 && total) atomicAdd(total, acc_all);
}

// integrals on fixed-node rules: W (nel x N) weights at the N nodes of the
// degree-Q Lagrange lattice of the reference tetrahedron (xg_subdiv_weights);
// one thread per element
template <int Q>
__global__ __launch_bounds__(256) void k_integrate_nodes(int nel, const int32_t *__restrict__ elems,
                                                         const int32_t *__restrict__ el2v,
                                                         const double *__restrict__ coords,
                                                         const double *__restrict__ W, Poly f,
                                                         double *__restrict__ per_elem, double *total) {
  constexpr Lattice<3, Q> LAT{};
  constexpr int N = Lattice<3, Q>::N;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  if (i < nel) {
    const int e = elems ? elems[i] : i;
    double V[4][3];
#pragma unroll
    for (int k = 0; k <= 3; ++k) {
      const int v = el2v[(size_t)e * 4 + k];
#pragma unroll
      for (int d = 0; d < 3; ++d) V[k][d] = coords[(size_t)v * 3 + d];
    }
    double F[9], Finv[9], absdet;
    jacobian<3>(V, F, Finv, absdet);
    for (int k = 0; k < N; ++k) {
      double x[3];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        double a = V[3][r];
#pragma unroll
        for (int c = 0; c < 3; ++c) a += F[r * 3 + c] * ((double)LAT.a[k][c] / Q);
        x[r] = a;
      }
      double fv = 0.0;
      for (int t = 0; t < f.n; ++t) {
        double m = f.c[t];
        for (int d = 0; d < 3; ++d)
          for (int j = 0; j < f.ex[t][d]; ++j) m *= x[d];
        fv += m;
      }
      s += absdet * W[(size_t)i * N + k] * fv;
    }
    if (per_elem) per_elem[i] = s;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0 && total) atomicAdd(total, s);
}

int grid_for_waves(int nel) {
  long long blocks = ((long long)nel * 64 + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  return (int)blocks;
}

}  // namespace

extern "C" int xg_rule_npts(int D, int order) {
  if (D < 0 || D > 3 || order < 0 || order / 2 + 1 > XG_RULE_NMAX) return XG_ERR_ARG;
  return xg_rule_cnt[D][order / 2 + 1];
}

extern "C" int xg_cut_rule_count(xg_ctx *ctx, int D, int nel, const int32_t *elems, const int32_t *el2v,
                                 const double *lset_v, int dt, int order, int32_t *offsets, uint8_t *nsub) {
  if (!ctx || D < 1 || D > 3 || nel < 0 || !offsets || dt < 0 || dt > 2 || order < 0 ||
      order / 2 + 1 > XG_RULE_NMAX)
    return set_error(XG_ERR_ARG, "xg_cut_rule_count: bad argument");
  if (nel > 0 && (!el2v || !lset_v)) return set_error(XG_ERR_ARG, "xg_cut_rule_count: null mesh");
  if (nel == 0) {
    XG_CHECK(hipMemsetAsync(offsets, 0, sizeof(int32_t), ctx->stream));
    return XG_OK;
  }
  const int grid = (nel + 255) / 256;
  switch (D) {
    case 1: k_rule_count<1><<<grid, 256, 0, ctx->stream>>>(nel, elems, el2v, lset_v, dt, order, offsets, nsub, ctx->d_err); break;
    case 2: k_rule_count<2><<<grid, 256, 0, ctx->stream>>>(nel, elems, el2v, lset_v, dt, order, offsets, nsub, ctx->d_err); break;
    default: k_rule_count<3><<<grid, 256, 0, ctx->stream>>>(nel, elems, el2v, lset_v, dt, order, offsets, nsub, ctx->d_err); break;
  }
  XG_LAUNCH_CHECK();
  return scan_exclusive_i32(ctx, offsets, nel, true);
}

extern "C" int xg_cut_rule_fill(xg_ctx *ctx, int D, int nel, const int32_t *elems, const int32_t *el2v,
                                const double *coords, const double *lset_v, int dt, int order,
                                const int32_t *offsets, double *pts) {
  if (!ctx || D < 1 || D > 3 || nel < 0 || !offsets || dt < 0 || dt > 2 || order < 0 ||
      order / 2 + 1 > XG_RULE_NMAX)
    return set_error(XG_ERR_ARG, "xg_cut_rule_fill: bad argument");
  if (nel == 0) return XG_OK;
  if (!el2v || !lset_v || !coords || !pts) return set_error(XG_ERR_ARG, "xg_cut_rule_fill: null pointer");
  const int grid = grid_for_waves(nel);
  double4 *o = reinterpret_cast<double4 *>(pts);
  switch (D) {
    case 1: k_rule_fill<1><<<grid, 256, 0, ctx->stream>>>(nel, elems, el2v, coords, lset_v, dt, order, offsets, o); break;
    case 2: k_rule_fill<2><<<grid, 256, 0, ctx->stream>>>(nel, elems, el2v, coords, lset_v, dt, order, offsets, o); break;
    default: k_rule_fill<3><<<grid, 256, 0, ctx->stream>>>(nel, elems, el2v, coords, lset_v, dt, order, offsets, o); break;
  }
  XG_LAUNCH_CHECK();
  return XG_OK;
}

extern "C" int xg_integrate(xg_ctx *ctx, int D, int nel, const int32_t *elems, const int32_t *el2v,
                            const double *coords, const int32_t *offsets, const double *pts, int nterms,
                            const double *coef, const int32_t *expo, double *per_elem, double *total) {
  if (!ctx || D < 1 || D > 3 || nel < 0 || nterms < 0 || nterms > 32 || (nterms > 0 && (!coef || !expo)))
    return set_error(XG_ERR_ARG, "xg_integrate: bad argument");
  if (total) XG_CHECK(hipMemsetAsync(total, 0, sizeof(double), ctx->stream));
  if (nel == 0) return XG_OK;
  if (!el2v || !coords || !offsets || !pts) return set_error(XG_ERR_ARG, "xg_integrate: null pointer");
  Poly f;
  f.n = nterms;
  for (int t = 0; t < nterms; ++t) {
    f.c[t] = coef[t];
    for (int d = 0; d < 3; ++d) {
      if (expo[3 * t + d] < 0 || expo[3 * t + d] > 64) return set_error(XG_ERR_ARG, "xg_integrate: bad exponent");
      f.ex[t][d] = expo[3 * t + d];
    }
  }
  const int grid = grid_for_waves(nel);
  const double4 *p = reinterpret_cast<const double4 *>(pts);
  switch (D) {
    case 1: k_integrate<1><<<grid, 256, 0, ctx->stream>>>(nel, elems, el2v, coords, offsets, p, f, per_elem, total); break;
    case 2: k_integrate<2><<<grid, 256, 0, ctx->stream>>>(nel, elems, el2v, coords, offsets, p, f, per_elem, total); break;
    default: k_integrate<3><<<grid, 256, 0, ctx->stream>>>(nel, elems, el2v, coords, offsets, p, f, per_elem, total); break;
  }
  XG_LAUNCH_CHECK();
  return XG_OK;
}

extern "C" int xg_integrate_nodes(xg_ctx *ctx, int nel, const int32_t *elems, const int32_t *el2v,
                                  const double *coords, int q, const double *weights, int nterms,
                                  const double *coef, const int32_t *expo, double *per_elem, double *total) {
  if (!ctx || nel < 0 || q != 4 || nterms < 0 || nterms > 32 || (nterms > 0 && (!coef || !expo)))
    return set_error(XG_ERR_ARG, "xg_integrate_nodes: bad argument (q == 4)");
  if (total) XG_CHECK(hipMemsetAsync(total, 0, sizeof(double), ctx->stream));
  if (nel == 0) return XG_OK;
  if (!el2v || !coords || !weights) return set_error(XG_ERR_ARG, "xg_integrate_nodes: null pointer");
  Poly f;
  f.n = nterms;
  for (int t = 0; t < nterms; ++t) {
    f.c[t] = coef[t];
    for (int d = 0; d < 3; ++d) {
      if (expo[3 * t + d] < 0 || expo[3 * t + d] > 64)
        return set_error(XG_ERR_ARG, "xg_integrate_nodes: bad exponent");
      f.ex[t][d] = expo[3 * t + d];
    }
  }
  k_integrate_nodes<4><<<(nel + 255) / 256, 256, 0, ctx->stream>>>(nel, elems, el2v, coords, weights, f, per_elem,
                                                                   total);
  XG_LAUNCH_CHECK();
  return XG_OK;
}
