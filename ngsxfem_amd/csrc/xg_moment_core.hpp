// xg_moment_core.hpp -- device building blocks of the moment-form element
// matrices (shared by the single-mesh kernel in xg_moment.hip and the
// batched persistent kernels in xg_batch.hip).  See xg_moment.hip.
#pragma once
#include "moments_table.h"
#include "xg_device.hpp"

namespace xg {
namespace mcore {

// vertex `code` of a decomposition entry in reference coordinates (FP64,
// same operation order as cut_point_array)
template <int D>
__device__ __forceinline__ void piece_vertex(int code, const DecompEntry *E, const double *t, double *p) {
  if (code <= D) {
#pragma unroll
    for (int d = 0; d < 3; ++d) p[d] = (d == code && code < D) ? 1.0 : 0.0;
    return;
  }
  const int c = code - (D + 1);
  const int i = E->ce[c][0], j = E->ce[c][1];
  double tc = t[0];
#pragma unroll
  for (int k = 1; k < 4; ++k) tc = (c == k) ? t[k] : tc;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double pi = (d == i && i < D) ? 1.0 : 0.0;
    const double pj = (d == j && j < D) ? 1.0 : 0.0;
    p[d] = pi + tc * (pj - pi);
  }
}

// SimpleX::GetVolume (straightcutrule.cpp:66-72) on explicit vertices
__device__ __forceinline__ double piece_volume(int nv, const double (*W)[3]) {
  if (nv == 1) return 1.0;
  if (nv == 2) {
    double a[3] = {W[1][0] - W[0][0], W[1][1] - W[0][1], W[1][2] - W[0][2]};
    return norm3(a);
  }
  if (nv == 3) {
    double a[3], b[3], c[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      a[d] = W[2][d] - W[0][d];
      b[d] = W[1][d] - W[0][d];
    }
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
    return norm3(c);
  }
  double a[3], b[3], c[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    a[d] = W[3][d] - W[0][d];
    b[d] = W[2][d] - W[0][d];
    c[d] = W[1][d] - W[0][d];
  }
  return fabs(det3(a, b, c));
}

// moments of one sub-simplex: base rule (NC reference coordinates per
// point) mapped by x = ow W0 + sum_q xi_q W_{q+1}
template <int D, int P, int NC>
__device__ __forceinline__ void accum_piece(int roff, int nq, const double (*W)[3], double wsc, double *m) {
  // the next point's rule entry is loaded (scalar loads, wave-uniform)
  // before the current point is processed, so the load latency overlaps
  // the FP64 work
  double nx[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) nx[c] = xg_rule_data_d[roff][c];
  for (int k = 0; k < nq; ++k) {
    double r[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) r[c] = nx[c];
    if (k + 1 < nq) {
#pragma unroll
      for (int c = 0; c < 4; ++c) nx[c] = xg_rule_data_d[roff + k + 1][c];
    }
    double ow = 1.0;
#pragma unroll
    for (int q = 0; q < NC; ++q) ow -= r[q];
    double x[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) x[d] = ow * W[0][d];
#pragma unroll
    for (int q = 0; q < NC; ++q)
#pragma unroll
      for (int d = 0; d < 3; ++d) x[d] = fma(r[q], W[q + 1][d], x[d]);
    double lam[D + 1];
    double sl = 1.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      lam[d] = x[d];
      sl -= x[d];
    }
    lam[D] = sl;
    Moments<D, P>::accum(lam, r[3] * wsc, m);
  }
}

// One (element, domain type): rule topology, moments, identities; writes the
// ndof x ndof matrix to out.  Returns true where the reference throws
// (straightcutrule.cpp:45-49, :148, :193).
template <int D, int P, typename T>
__device__ __forceinline__ bool moment_elmat(const double *l, const double *Finv, double absdet, int dt, int order,
                                             int mask, T *out) {
  using MM = Moments<D, P>;
  // topology of the (element, dt) rule
  const int nvs = (dt == IF) ? D : D + 1;  // vertices per sub-simplex (wave-uniform)
  const int eldom = check_if_straight_cut<D>(l);
  const DecompEntry *E = nullptr;
  int ns = 0;
  bool bad = false;
  double tcut[4] = {0.0, 0.0, 0.0, 0.0};
  if (eldom != IF) {
    ns = (eldom == dt) ? 1 : 0;  // uncut: the base rule on the element itself
  } else if (!cut_is_valid<D>(l)) {
    bad = true;
  } else {
    E = &xg_decomp_d.e[D - 1][dt][sign_mask<D>(l)];
    if (!E->ok) {
      bad = true;
    } else {
      ns = E->ns;
      for (int c = 0; c < E->nc; ++c) {
        const int a = E->ce[c][0], b = E->ce[c][1];
        double la = l[0], lb = l[0];
#pragma unroll
        for (int k = 1; k <= D; ++k) {
          la = (k == a) ? l[k] : la;
          lb = (k == b) ? l[k] : lb;
        }
        const double tc = la / (la - lb);
#pragma unroll
        for (int k = 0; k < 4; ++k) tcut[k] = (c == k) ? tc : tcut[k];
      }
    }
  }
  const double ifac = (dt == IF && ns > 0) ? interface_factor<D>(l, Finv) : 1.0;
  const int rd = (dt == IF) ? D - 1 : D;
  const int nq = rule_cnt(rd, order), roff = rule_off(rd, order);

  double m[MM::NM];
#pragma unroll
  for (int k = 0; k < MM::NM; ++k) m[k] = 0.0;
  for (int s = 0; s < ns; ++s) {
    double W[D + 1][3];
    double wsc = 1.0;
    if (E == nullptr) {
      // uncut element: origin, e_0, .., e_{D-1} (identity map)
#pragma unroll
      for (int k = 0; k <= D; ++k)
#pragma unroll
        for (int d = 0; d < 3; ++d) W[k][d] = (k > 0 && d == k - 1) ? 1.0 : 0.0;
    } else {
#pragma unroll
      for (int k = 0; k <= D; ++k) {
        if (k < nvs) {
          piece_vertex<D>(E->v[s][k], E, tcut, W[k]);
        } else {
#pragma unroll
          for (int d = 0; d < 3; ++d) W[k][d] = 0.0;
        }
      }
      wsc = piece_volume(nvs, W) * ifac;
    }
    // one code path for volume and interface pieces (a single instantiation
    // keeps the register allocation of the two from adding up): the rule
    // coordinates past the piece's dimension are exact zeros in the rule
    // table and the unused vertices are zero, so the extra terms add exact
    // zeros
    accum_piece<D, P, D>(roff, nq, W, wsc, m);
  }
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
  MM::template forms<T>(m, mk, Hs, sm, sk, out);
  return bad;
}

}  // namespace mcore
}  // namespace xg
