// xg_mark.hip -- element marking (CutInformation::Update), compaction of the
// cut elements, P1 perturbation, facet marking, device scan.
//
// Reference: xfem/cutinfo.cpp:61-133 (marking), :476-538 (facets),
//            utils/p1interpol.cpp:104-107 (perturbation).
//
// Layout: one thread per element, 256-thread blocks (4 waves).  The element
// row of el2v is read as one int4 (tets) -- 16 B per lane, coalesced -- and
// the D+1 level-set values are gathered (FP64, L2-resident for meshes of a
// few million vertices).  The 8 COMBINED_DOMAIN_TYPE bit arrays are built
// with one wave ballot per array: lane 0 stores the 64-bit mask, i.e. two
// uint32 words, so no atomics touch the bit arrays.
#include <mutex>

#include "xg_device.hpp"
#include "xg_internal.hpp"

using namespace xg;

namespace {

constexpr int BLK = 256;

template <int D>
__device__ __forceinline__ void load_lset(const int32_t *__restrict__ el2v,
                                          const double *__restrict__ lset, int e, double *l) {
  if constexpr (D == 3) {
    const int4 v = reinterpret_cast<const int4 *>(el2v)[e];
    l[0] = lset[v.x];
    l[1] = lset[v.y];
    l[2] = lset[v.z];
    l[3] = lset[v.w];
  } else {
#pragma unroll
    for (int k = 0; k <= D; ++k) l[k] = lset[el2v[(size_t)e * (D + 1) + k]];
  }
}

// Coordinate d of point idx of the cut point array: reference vertex k < D
// is e_k, vertex D the origin, D+1+c the c-th cut point (the layout of
// cut_point_array).  Selects over registers, no indexed private arrays.
template <int D>
__device__ __forceinline__ double point_coord(int idx, int d, const double (&cp)[4][3]) {
  double v = (idx == d && d < D) ? 1.0 : 0.0;
#pragma unroll
  for (int c = 0; c < 4; ++c) v = (idx == D + 1 + c) ? cp[c][d] : v;
  return v;
}

// part_vol[NEG], part_vol[POS] from the order-0 rules (cutinfo.cpp:87-101):
// the weight sums of StraightCutIntegrationRule(..., np, order 0).  Returns
// false when the reference would throw.  Same decomposition (the compile-time
// table of decompose(), by sign pattern), same cut points
// (straightcutrule.cpp:50) and the same volume arithmetic
// (SimpleX::GetVolume, :66-72) as cut_point_array / simplex_volume, with the
// point array held in registers: the indexed arrays of the generic helpers
// would live in per-lane scratch memory, which every wave of this
// memory-bound kernel would then have to set up.
template <int D>
__device__ __forceinline__ bool part_volumes(const double *l, double *pv) {
  const double w0 = xg_rule_data_d[xg_rule_off_d[D][1]][3];  // order-0 weight = |ref simplex|
  pv[NEG] = 0.0;
  pv[POS] = 0.0;
  const int eldom = check_if_straight_cut<D>(l);
  if (eldom != IF) {
    pv[eldom] = w0;
    return true;
  }
  // decompose(): every sign-changing pair must be a straight cut (no exact
  // zero next to a negative value, straightcutrule.cpp:45-49)
  bool hz = false, hn = false;
#pragma unroll
  for (int i = 0; i <= D; ++i) {
    hz |= l[i] == 0.0;
    hn |= l[i] < 0.0;
  }
  const bool pairs_ok = !(hz && hn);
  const int msk = sign_mask<D>(l);
  // cut points of the sign-changing edges (same list for both parts)
  const DecompEntry &E0 = xg_decomp_d.e[D - 1][POS][msk];
  double cp[4][3];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int i = E0.ce[c][0], j = E0.ce[c][1];
    double li = l[0], lj = l[0];
#pragma unroll
    for (int k = 1; k <= D; ++k) {
      li = (k == i) ? l[k] : li;
      lj = (k == j) ? l[k] : lj;
    }
    const double t = (c < E0.nc) ? li / (li - lj) : 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double pi = (d == i && i < D) ? 1.0 : 0.0;
      const double pj = (d == j && j < D) ? 1.0 : 0.0;
      cp[c][d] = pi + t * (pj - pi);
    }
  }
  bool ok = true;
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int np = t == 0 ? POS : NEG;
    const DecompEntry &E = xg_decomp_d.e[D - 1][np][msk];
    if (!(pairs_ok && E.ok)) {
      ok = false;
      continue;
    }
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (k >= E.ns) break;
      double Q[4][3];
#pragma unroll
      for (int m = 0; m <= D; ++m)
#pragma unroll
        for (int d = 0; d < 3; ++d) Q[m][d] = point_coord<D>(E.v[k][m], d, cp);
      double vol;
      if constexpr (D == 1) {
        const double a[3] = {Q[1][0] - Q[0][0], Q[1][1] - Q[0][1], Q[1][2] - Q[0][2]};
        vol = norm3(a);
      } else if constexpr (D == 2) {
        double a[3], b[3], c[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          a[d] = Q[2][d] - Q[0][d];
          b[d] = Q[1][d] - Q[0][d];
        }
        c[0] = a[1] * b[2] - a[2] * b[1];
        c[1] = a[2] * b[0] - a[0] * b[2];
        c[2] = a[0] * b[1] - a[1] * b[0];
        vol = norm3(c);
      } else {
        double a[3], b[3], c[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          a[d] = Q[3][d] - Q[0][d];
          b[d] = Q[2][d] - Q[0][d];
          c[d] = Q[1][d] - Q[0][d];
        }
        vol = fabs(det3(a, b, c));
      }
      s += w0 * vol;
    }
    pv[np] = s;
  }
  return ok;
}

// Marking kernel.  A block of NB threads marks NB * EPT consecutive
// elements, EPT per thread (element base + k * NB + tid, so every wave's
// elements stay 64 consecutive ones for the bit-array ballots): a thread
// first issues the element rows of all its elements, then all level-set
// gathers, so that EPT rows and (D+1) * EPT values per lane are in flight at
// once -- the kernel is bound by memory latency otherwise.  Elements with a
// sign change (the only ones that need the cut decomposition) are queued in
// LDS per wave and processed densely by the wave's first lanes: one pass of
// the decomposition code for up to 64 of them instead of one per group of 64
// elements holding any, and no block barrier that would hold the other waves
// while one decomposes.  The 8 bit-array words of each group of 64 elements
// (two 32-bit halves of 8 ballots) go out in one store instruction per wave,
// one word per lane.
constexpr int MARK_BLK = 256;
constexpr int MARK_EPT = 4;

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// (launch bounds: at most 80 VGPRs, 6 waves per SIMD; the decomposition
// code of the queued elements otherwise sets the register count and leaves
// 5, and 8 waves need spills -- 6 measured fastest)
template <int D, int NB, int EPT>
__global__ __launch_bounds__(NB, 6) void k_mark(int ne, const int32_t *__restrict__ el2v,
                                             const double *__restrict__ lset,
                                             uint8_t *__restrict__ dt_out,
                                             double *__restrict__ cut_ratio,
                                             uint32_t *__restrict__ bits, int nw,
                                             int32_t *__restrict__ blk_if, int32_t *err) {
  static_assert(EPT * 16 <= 64, "one bit-array word per lane: 8 arrays x 2 halves x EPT groups");
  constexpr int NE = NB * EPT;  // elements per block
  constexpr int NW = NB / 64;   // waves per block
  __shared__ int s_cnt[NW];
  __shared__ int16_t s_qi[NW][64 * EPT];  // the wave's queued element slots (k * 64 + lane)
  __shared__ uint8_t s_cdt[NW][64 * EPT];  // combined domain bits of the queued slots
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int base = blockIdx.x * NE;
  // element rows of all EPT elements, then their level-set values.  The
  // loads are unconditional (lanes past the end read the last element
  // again): a load under a branch is waited for at the branch's end, which
  // would serialise the EPT round trips.
  int vi[EPT][D + 1];
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const int ec = min(base + k * NB + tid, ne - 1);
    if constexpr (D == 3) {
      const int4 v = reinterpret_cast<const int4 *>(el2v)[ec];
      vi[k][0] = v.x;
      vi[k][1] = v.y;
      vi[k][2] = v.z;
      vi[k][3] = v.w;
    } else {
#pragma unroll
      for (int i = 0; i <= D; ++i) vi[k][i] = el2v[(size_t)ec * (D + 1) + i];
    }
  }
  double l[EPT][D + 1];
#pragma unroll
  for (int k = 0; k < EPT; ++k)
#pragma unroll
    for (int i = 0; i <= D; ++i) l[k][i] = lset[vi[k][i]];
  int cdt[EPT];
  bool cand[EPT];
  int nq = 0;  // wave-uniform
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const int e = base + k * NB + tid;
    cdt[k] = 0;
    cand[k] = false;
    if (e < ne) {
      const int eldom = check_if_straight_cut<D>(l[k]);
      if (eldom != IF) {
        // one part holds the whole order-0 weight: cut ratio w0/(w0+0) = 1
        // or 0/(0+w0) = 0 (cutinfo.cpp:87-124)
        cdt[k] = eldom == NEG ? 1 : 2;
        if (cut_ratio) cut_ratio[e] = eldom == NEG ? 1.0 : 0.0;
        if (dt_out) dt_out[e] = (uint8_t)eldom;
      } else {
        cand[k] = true;
      }
    }
    const unsigned long long mc = __ballot(cand[k]);
    if (cand[k])
      s_qi[wv][nq + __builtin_amdgcn_mbcnt_hi((uint32_t)(mc >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mc, 0))] =
          (int16_t)(k * 64 + lane);
    nq += __popcll(mc);
  }
  wave_sync_lds();
  for (int q0 = 0; q0 < nq; q0 += 64) {  // wave-uniform
    const int q = q0 + lane;
    if (q < nq) {
      const int slot = s_qi[wv][q];
      const int eq = base + (slot >> 6) * NB + wv * 64 + (slot & 63);
      double lq[D + 1];
      load_lset<D>(el2v, lset, eq, lq);  // cache hits: gathered above
      double pv[2];
      if (!part_volumes<D>(lq, pv)) atomicOr(err, 1);
      if (cut_ratio) cut_ratio[eq] = pv[NEG] / (pv[NEG] + pv[POS]);
      // cutinfo.cpp:104-124
      int c;
      if (pv[NEG] > 0.0)
        c = (pv[POS] > 0.0) ? 4 : 1;
      else
        c = 2;
      if (dt_out) dt_out[eq] = (uint8_t)(c == 4 ? IF : (c == 1 ? NEG : POS));
      s_cdt[wv][slot] = (uint8_t)c;
    }
  }
  wave_sync_lds();
  // combined arrays (cutinfo.cpp:72-78,127-132): bit set iff (c & cdt) != 0,
  // CDOM_ANY everywhere, CDOM_NO nowhere.  Lane L stores half (L & 1) of
  // array (L >> 1) & 7 for group L >> 4.
  const int mk = lane >> 4, mc = (lane >> 1) & 7, mh = lane & 1;
  uint32_t word = 0;
  int nif = 0;
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const bool valid = base + k * NB + tid < ne;
    if (cand[k]) cdt[k] = s_cdt[wv][k * 64 + lane];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const unsigned long long m = __ballot(valid && ((c == 7) || (c & cdt[k])));
      const uint32_t h = mh ? (uint32_t)(m >> 32) : (uint32_t)(m & 0xffffffffull);
      word = (mk == k && mc == c) ? h : word;
    }
    nif += __popcll(__ballot(valid && cdt[k] == 4));
  }
  {
    const int w = ((base + mk * NB + wv * 64) >> 5) + mh;
    if (lane < 16 * EPT && w < nw) bits[(size_t)mc * nw + w] = word;
  }
  if (lane == 0) s_cnt[wv] = nif;
  __syncthreads();
  if (tid == 0 && blk_if) {
    int s = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) s += s_cnt[k];
    blk_if[blockIdx.x] = s;
  }
}

// writes the IF elements of block b (NB * EPT elements per block, as in the
// marking pass that counted them) to cut_list[blk_off[b] ...] in ascending
// order
template <int NB, int EPT>
__global__ __launch_bounds__(NB) void k_compact(int ne, const uint32_t *__restrict__ if_bits,
                                                const int32_t *__restrict__ blk_off,
                                                int32_t *__restrict__ cut_list) {
  __shared__ int s_cnt[EPT][NB / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  bool on[EPT];
  int below[EPT];
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const int e = blockIdx.x * NB * EPT + k * NB + threadIdx.x;
    on[k] = e < ne && ((if_bits[e >> 5] >> (e & 31)) & 1u);
    const unsigned long long m = __ballot(on[k]);
    below[k] = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
    if (lane == 0) s_cnt[k][wv] = __popcll(m);
  }
  __syncthreads();
  int off = blk_off[blockIdx.x];
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    int o = off, tot = 0;
#pragma unroll
    for (int j = 0; j < NB / 64; ++j) {
      o += (j < wv) ? s_cnt[k][j] : 0;
      tot += s_cnt[k][j];
    }
    if (on[k]) cut_list[o + below[k]] = blockIdx.x * NB * EPT + k * NB + threadIdx.x;
    off += tot;
  }
}

// ---- marking for P2 level sets (per-element nodal values, Lattice<3,2> order)
//
// The reference classifies such an element by the lattice check of its
// subdivision rule (xintegration.cpp:617-752): the level set at the
// C(2^L+3, 3) points of the level-L lattice, IF if both v >= 0 and v < 0
// occur.  Pass 1 (one lane per element) settles most elements without the
// lattice: the Bernstein-Bezier coefficients of a quadratic on a simplex
// (vertex values, and 2 phi_ij - (phi_i + phi_j) / 2 on the edges) bound it
// from both sides (convex hull property), so all coefficients
// > 1e-12 max|b| means v > 0 at every lattice point (POS), all < -1e-12
// max|b| means v < 0 (NEG).  Pass 2 runs the lattice check, one wave per
// element, for the rest.
constexpr uint8_t XG_UNDECIDED = 3;

__global__ __launch_bounds__(BLK) void k_mark_p2_bern(int ne, const double *__restrict__ lset_nodal,
                                                      uint8_t *__restrict__ dt_out) {
  const int e = blockIdx.x * BLK + threadIdx.x;
  if (e >= ne) return;
  double v[10];
#pragma unroll
  for (int k = 0; k < 10; ++k) v[k] = lset_nodal[(size_t)e * 10 + k];
  constexpr int EI[6] = {0, 0, 0, 1, 1, 2}, EJ[6] = {1, 2, 3, 2, 3, 3};
  double b[10];
#pragma unroll
  for (int k = 0; k < 4; ++k) b[k] = v[k];
#pragma unroll
  for (int k = 0; k < 6; ++k) b[4 + k] = 2.0 * v[4 + k] - 0.5 * (v[EI[k]] + v[EJ[k]]);
  double mx = 0.0, mn = b[0], mxb = b[0];
#pragma unroll
  for (int k = 0; k < 10; ++k) {
    mx = fmax(mx, fabs(b[k]));
    mn = fmin(mn, b[k]);
    mxb = fmax(mxb, b[k]);
  }
  const double tol = 1e-12 * mx;
  uint8_t d = XG_UNDECIDED;
  if (mx > 0.0 && mn > tol) d = POS;
  else if (mx > 0.0 && mxb < -tol) d = NEG;
  dt_out[e] = d;
}

// one wave per group of 64 elements; the undecided ones get the lattice check
__global__ __launch_bounds__(BLK) void k_mark_p2_lattice(int ne, const double *__restrict__ lset_nodal, int L,
                                                         uint8_t *__restrict__ dt_out) {
  const int lane = threadIdx.x & 63;
  const int g = __builtin_amdgcn_readfirstlane((blockIdx.x * BLK + threadIdx.x) >> 6);
  const int e = g * 64 + lane;
  const bool und = e < ne && dt_out[e] == XG_UNDECIDED;
  unsigned long long m = __ballot(und);
  const int n = 1 << L;
  const int off = xg_lat_off_d[L], cnt = xg_lat_cnt_d[L];
  const double inv_n = 1.0 / n;
  while (m) {
    const int j = __builtin_ctzll(m);
    m &= m - 1;
    const int ej = g * 64 + j;
    double cl[10];
#pragma unroll
    for (int k = 0; k < 10; ++k) cl[k] = lset_nodal[(size_t)ej * 10 + k];
    bool hp = false, hn = false;
    for (int q = lane; q < cnt; q += 64) {
      const unsigned pk = xg_lat_d[off + q];
      const double x[3] = {(double)(pk & 255u) * inv_n, (double)((pk >> 8) & 255u) * inv_n,
                           (double)(pk >> 16) * inv_n};
      const double v = lagrange_eval3<2>(cl, x);
      if (v >= 0.0) hp = true;
      else hn = true;
    }
    const bool anyp = __ballot(hp) != 0ull, anyn = __ballot(hn) != 0ull;
    if (lane == 0) dt_out[ej] = (uint8_t)((anyp && anyn) ? IF : (anyp ? POS : NEG));
  }
}

// ---- boundary triangles of a tetrahedral mesh with a P2 level set
//
// CutInformation::Update sums the order-0 NEG and POS rules of each boundary
// element (cutinfo.cpp:80-124); with a P2 level set these are the
// subdivision rule on the triangle (xintegration.cpp:620-838 lattice check
// and four children, :1019-1061 uncut nodes, :1391-1545 cut leaves after
// Simplex::CheckIfCut, xdecompose.hpp:34-85).  For order 0 only areas count,
// and a node the lattice finds uncut has leaves whose vertices all carry its
// sign, so the adaptive tree gives the areas of the uniform level-L
// refinement: the triangles the Bernstein certificate leaves undecided, one
// wave each in turn, lanes over the n x n cells of the lattice (n = 2^L), each cell holding an upward (a,b),(a+1,b),(a,b+1) and a
// downward (a+1,b),(a,b+1),(a+1,b+1) leaf.  Coordinates are exact dyadics,
// so the values equal those of the nested evaluation.  Domain from the part
// areas (IF iff both > 0), cut ratio NEG / (NEG + POS).
XG_DEV double tri_area(const double *a, const double *b, const double *c) {
  const double u0 = b[0] - a[0], u1 = b[1] - a[1], w0 = c[0] - a[0], w1 = c[1] - a[1];
  return 0.5 * fabs(u0 * w1 - u1 * w0);
}

// areas of one leaf (vertices V, level-set values v) to NEG (an) and POS (ap)
XG_DEV void trig_leaf(const double (*V)[2], const double *v, double &an, double &ap) {
  bool hp = false, hn = false;
  double sp = 0.0, sn = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (v[i] >= 0.0) {
      sp += v[i];
      hp = true;
    } else {
      sn -= v[i];
      hn = true;
    }
  }
  // a leaf whose lattice check (its vertices) finds one sign goes whole to it
  if (!(hp && hn)) {
    (hp ? ap : an) += tri_area(V[0], V[1], V[2]);
    return;
  }
  const double s = sp + sn;
  if (sp / s < 1e-14) hp = false;
  if (sn / s < 1e-14) hn = false;
  if (!(hp && hn)) {
    (hp ? ap : an) += tri_area(V[0], V[1], V[2]);
    return;
  }
  // CutSimplex<2>::MakeQuad: zero counts as POS; an edge is cut where the
  // signs differ or a zero meets a negative value
  constexpr int EA[3] = {0, 0, 1}, EB[3] = {1, 2, 2};
  int neg[3], pos[3], nneg = 0, npos = 0, v2c[3] = {-1, -1, -1}, nc = 0;
  double cp[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    if (v[j] < 0.0) neg[nneg++] = j;
    else pos[npos++] = j;
  }
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int l = EA[j], r = EB[j];
    const bool cut = v[l] * v[r] < 0.0 || (v[l] == 0.0 && v[r] < 0.0) || (v[r] == 0.0 && v[l] < 0.0);
    if (cut && nc < 2) {
      const double t = v[l] / (v[l] - v[r]);
      cp[nc][0] = (1 - t) * V[l][0] + t * V[r][0];
      cp[nc][1] = (1 - t) * V[l][1] + t * V[r][1];
      if (v2c[l] == -1) v2c[l] = nc;
      if (v2c[r] == -1) v2c[r] = nc;
      ++nc;
    }
  }
  const bool minor_neg = nneg <= npos;
  const int mi = minor_neg ? neg[0] : pos[0];
  const int m0 = minor_neg ? pos[0] : neg[0], m1 = minor_neg ? pos[1] : neg[1];
  const double amin = tri_area(V[mi], cp[0], cp[1]);
  const double *q2 = cp[v2c[m0] < 0 ? 0 : v2c[m0]], *q3 = cp[v2c[m1] < 0 ? 0 : v2c[m1]];
  const double amaj = tri_area(V[m0], V[m1], q2) + tri_area(V[m1], q2, q3);
  (minor_neg ? an : ap) += amin;
  (minor_neg ? ap : an) += amaj;
}

// pass 1, one lane per triangle: the Bernstein-Bezier certificate of
// k_mark_p2_bern (coefficients: vertex values, 2 phi_ij - (phi_i + phi_j) / 2
// on the edges); certified triangles get their domain and cut ratio
__global__ __launch_bounds__(BLK) void k_mark_p2_trig_bern(int ne, const double *__restrict__ lset_nodal,
                                                           uint8_t *__restrict__ dt_out,
                                                           double *__restrict__ cut_ratio) {
  const int e = blockIdx.x * BLK + threadIdx.x;
  if (e >= ne) return;
  double c[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) c[k] = lset_nodal[(size_t)e * 6 + k];
  constexpr int EI[3] = {0, 0, 1}, EJ[3] = {1, 2, 2};
  double bb[6];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    bb[k] = c[k];
    bb[3 + k] = 2.0 * c[3 + k] - 0.5 * (c[EI[k]] + c[EJ[k]]);
  }
  double mx = 0.0, mn = bb[0], mxb = bb[0];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    mx = fmax(mx, fabs(bb[k]));
    mn = fmin(mn, bb[k]);
    mxb = fmax(mxb, bb[k]);
  }
  const double tol = 1e-12 * mx;
  uint8_t d = XG_UNDECIDED;
  if (mx > 0.0 && mn > tol) d = POS;
  else if (mx > 0.0 && mxb < -tol) d = NEG;
  dt_out[e] = d;
  if (d != XG_UNDECIDED) cut_ratio[e] = d == POS ? 0.0 : 1.0;
}

// pass 2, one wave per triangle (undecided ones are clustered along the
// cut, so a wave per group of them would serialise): waves of certified
// triangles return after one byte load; lattice check and leaves
// lane-parallel
__global__ __launch_bounds__(BLK) void k_mark_p2_trig(int ne, const double *__restrict__ lset_nodal, int L,
                                                      uint8_t *__restrict__ dt_out,
                                                      double *__restrict__ cut_ratio) {
  const int lane = threadIdx.x & 63;
  const int e = __builtin_amdgcn_readfirstlane((blockIdx.x * BLK + threadIdx.x) >> 6);
  if (e >= ne || dt_out[e] != XG_UNDECIDED) return;
  const int n = 1 << L;
  const double inv_n = 1.0 / n;
  {
    double c[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) c[k] = lset_nodal[(size_t)e * 6 + k];
    // lattice check of the whole triangle (xintegration.cpp:620-755)
    bool hp = false, hn = false;
    for (int q = lane; q < (n + 1) * (n + 1); q += 64) {
      const int a = q % (n + 1), b = q / (n + 1);
      if (a + b > n) continue;
      const double x[2] = {a * inv_n, b * inv_n};
      const double v = lagrange_eval2<2>(c, x);
      if (v >= 0.0) hp = true;
      else hn = true;
    }
    const bool anyp = __ballot(hp) != 0ull, anyn = __ballot(hn) != 0ull;
    if (!(anyp && anyn)) {
      if (lane == 0) {
        dt_out[e] = anyp ? POS : NEG;
        cut_ratio[e] = anyp ? 0.0 : 1.0;
      }
      return;
    }
    double an = 0.0, ap = 0.0;
    for (int q = lane; q < n * n; q += 64) {
      const int a = q % n, b = q / n;
      if (a + b > n - 1) continue;
      const double x00[2] = {a * inv_n, b * inv_n}, x10[2] = {(a + 1) * inv_n, b * inv_n},
                   x01[2] = {a * inv_n, (b + 1) * inv_n};
      const double v00 = lagrange_eval2<2>(c, x00), v10 = lagrange_eval2<2>(c, x10),
                   v01 = lagrange_eval2<2>(c, x01);
      {
        const double V[3][2] = {{x00[0], x00[1]}, {x10[0], x10[1]}, {x01[0], x01[1]}};
        const double v[3] = {v00, v10, v01};
        trig_leaf(V, v, an, ap);
      }
      if (a + b <= n - 2) {
        const double x11[2] = {(a + 1) * inv_n, (b + 1) * inv_n};
        const double V[3][2] = {{x10[0], x10[1]}, {x01[0], x01[1]}, {x11[0], x11[1]}};
        const double v[3] = {v10, v01, lagrange_eval2<2>(c, x11)};
        trig_leaf(V, v, an, ap);
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      an += __shfl_xor(an, o, 64);
      ap += __shfl_xor(ap, o, 64);
    }
    if (lane == 0) {
      dt_out[e] = an > 0.0 ? (ap > 0.0 ? IF : NEG) : POS;
      cut_ratio[e] = an / (an + ap);
    }
  }
}

// bit arrays and block IF counts from the domain types (as k_mark)
__global__ __launch_bounds__(BLK) void k_bits_from_dt(int ne, const uint8_t *__restrict__ dt,
                                                      uint32_t *__restrict__ bits, int nw,
                                                      int32_t *__restrict__ blk_if) {
  __shared__ int s_cnt[BLK / 64];
  const int e = blockIdx.x * BLK + threadIdx.x;
  const bool valid = e < ne;
  int cdt = 0;
  if (valid) {
    const int d = dt[e];
    cdt = d == IF ? 4 : (d == NEG ? 1 : 2);
  }
  const int lane = threadIdx.x & 63;
  const int wave_e0 = e - lane;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const bool on = valid && ((c == 7) || (c & cdt));
    const unsigned long long m = __ballot(on);
    if (lane == 0) {
      const int w = wave_e0 >> 5;
      uint32_t *row = bits + (size_t)c * nw;
      if (w < nw) row[w] = (uint32_t)(m & 0xffffffffull);
      if (w + 1 < nw) row[w + 1] = (uint32_t)(m >> 32);
    }
  }
  const unsigned long long mif = __ballot(valid && cdt == 4);
  if (lane == 0) s_cnt[threadIdx.x >> 6] = __popcll(mif);
  __syncthreads();
  if (threadIdx.x == 0 && blk_if) {
    int s = 0;
#pragma unroll
    for (int k = 0; k < BLK / 64; ++k) s += s_cnt[k];
    blk_if[blockIdx.x] = s;
  }
}

__global__ void k_perturb(int n, double *v, double eps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const double x = v[i];
    if (fabs(x) < eps) v[i] = eps;
  }
}

// ---------------------------------------------------------------- scan
constexpr int SCAN_BLK = 1024;
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_TILE = SCAN_BLK * SCAN_ITEMS;

__device__ __forceinline__ int block_exclusive_scan(int v, int *s_wave, int &total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_wave[wv] = x;
  __syncthreads();
  if (wv == 0) {
    const int nwv = blockDim.x >> 6;
    int t = lane < nwv ? s_wave[lane] : 0;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(t, o, 64);
      if (lane >= o) t += y;
    }
    if (lane < nwv) s_wave[lane] = t;
  }
  __syncthreads();
  total = s_wave[(blockDim.x >> 6) - 1];
  const int wave_off = wv > 0 ? s_wave[wv - 1] : 0;
  __syncthreads();
  return wave_off + x - v;
}

// pass 1: scan each tile locally, write tile sums
__global__ __launch_bounds__(SCAN_BLK) void k_scan_tiles(int32_t *data, int n, int32_t *sums) {
  __shared__ int s_wave[SCAN_BLK / 64];
  const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  int v[SCAN_ITEMS], s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    v[k] = (base + k < n) ? data[base + k] : 0;
    s += v[k];
  }
  int total;
  int off = block_exclusive_scan(s, s_wave, total);
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    if (base + k < n) data[base + k] = off;
    off += v[k];
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// pass 2 (single block): exclusive scan of the tile sums, serial over
// chunks of SCAN_BLK
__global__ __launch_bounds__(SCAN_BLK) void k_scan_sums(int32_t *sums, int nt, int32_t *total_out,
                                                        int32_t *total_out2 = nullptr) {
  __shared__ int s_wave[SCAN_BLK / 64];
  int carry = 0;
  for (int base = 0; base < nt; base += SCAN_BLK) {
    const int i = base + threadIdx.x;
    const int v = i < nt ? sums[i] : 0;
    int total;
    const int off = block_exclusive_scan(v, s_wave, total);
    if (i < nt) sums[i] = carry + off;
    carry += total;
  }
  if (threadIdx.x == 0 && total_out) *total_out = carry;
  if (threadIdx.x == 0 && total_out2) *total_out2 = carry;
}

__global__ __launch_bounds__(SCAN_BLK) void k_scan_add(int32_t *data, int n, const int32_t *sums) {
  const int add = sums[blockIdx.x];
  const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k)
    if (base + k < n) data[base + k] += add;
}

// Exclusive scan of the n per-block counts of a marking pass, in place, by a
// single block: thread t owns the run of ceil(n / SCAN_BLK) consecutive
// counts starting at t * ceil(n / SCAN_BLK), read 8 at a time so that the
// loads of a run are in flight together.  data[n] = total (and *total_out).
// One launch instead of the three of the tiled scan: for the block counts of
// a few million elements the tiled passes are launch-latency bound.
constexpr int SCAN_ONE_BLOCK_MAX = 8 * SCAN_BLK;

__global__ __launch_bounds__(SCAN_BLK) void k_scan_counts(int32_t *data, int n, int32_t *total_out) {
  __shared__ int s_wave[SCAN_BLK / 64];
  const int per = (n + SCAN_BLK - 1) / SCAN_BLK;
  const int b = min(n, (int)threadIdx.x * per), e = min(n, b + per);
  int s = 0;
  for (int i = b; i < e; i += 8) {
    int v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = (i + k < e) ? data[i + k] : 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) s += v[k];
  }
  int total;
  int off = block_exclusive_scan(s, s_wave, total);
  for (int i = b; i < e; i += 8) {
    int v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = (i + k < e) ? data[i + k] : 0;
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (i + k < e) {
        data[i + k] = off;
        off += v[k];
      }
  }
  if (threadIdx.x == 0) {
    data[n] = total;
    if (total_out) *total_out = total;
  }
}

// exclusive scan of the n block counts at blk (blk[n] = total, also stored
// to *n_cut when given); sums: (n + SCAN_TILE - 1) / SCAN_TILE ints of
// scratch, used by the tiled path only
void scan_block_counts(hipStream_t st, int32_t *blk, int n, int32_t *sums, int32_t *n_cut) {
  if (n <= SCAN_ONE_BLOCK_MAX) {
    k_scan_counts<<<1, SCAN_BLK, 0, st>>>(blk, n, n_cut);
    return;
  }
  const int nt = (n + SCAN_TILE - 1) / SCAN_TILE;
  k_scan_tiles<<<nt, SCAN_BLK, 0, st>>>(blk, n, sums);
  k_scan_sums<<<1, SCAN_BLK, 0, st>>>(sums, nt, blk + n, n_cut);
  k_scan_add<<<nt, SCAN_BLK, 0, st>>>(blk, n, sums);
}

// ------------------------------------------------- facets (cutinfo.cpp:476-538)
__device__ __forceinline__ bool bit_of(const uint32_t *b, int e) { return (b[e >> 5] >> (e & 31)) & 1u; }

__global__ void k_facets(int nf, const int32_t *__restrict__ f2el, const uint32_t *__restrict__ a,
                         const uint32_t *__restrict__ b, int bva, int bvb, int use_and,
                         uint32_t *__restrict__ out, int nwf) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  bool r = false;
  if (f < nf) {
    const int e0 = f2el[2 * f], e1 = f2el[2 * f + 1];
    bool a0, a1, b0, b1;
    if (e1 < 0) {
      // boundary facet: the missing neighbour takes the boundary values
      a0 = bit_of(a, e0);
      b0 = bit_of(b, e0);
      a1 = bva != 0;
      b1 = bvb != 0;
    } else {
      a0 = bit_of(a, e0);
      a1 = bit_of(a, e1);
      b0 = bit_of(b, e0);
      b1 = bit_of(b, e1);
    }
    if (use_and)
      r = (a0 && b1) || (a1 && b0);
    else
      r = (a0 || b1) || (a1 || b0);
  }
  const unsigned long long m = __ballot(r);
  const int lane = threadIdx.x & 63;
  if (lane == 0) {
    const int w = (f - lane) >> 5;
    if (w < nwf) out[w] = (uint32_t)(m & 0xffffffffull);
    if (w + 1 < nwf) out[w + 1] = (uint32_t)(m >> 32);
  }
}

}  // namespace

namespace xg {

int scan_exclusive_i32(xg_ctx *ctx, int32_t *data, int n, bool with_total) {
  if (n <= 0) {
    if (with_total) XG_CHECK(hipMemsetAsync(data, 0, sizeof(int32_t), ctx->stream));
    return XG_OK;
  }
  const int nt = (n + SCAN_TILE - 1) / SCAN_TILE;
  int r = ensure_ws(ctx, sizeof(int32_t) * (size_t)(nt + 1));
  if (r) return r;
  int32_t *sums = (int32_t *)ctx->ws;
  k_scan_tiles<<<nt, SCAN_BLK, 0, ctx->stream>>>(data, n, sums);
  XG_LAUNCH_CHECK();
  k_scan_sums<<<1, SCAN_BLK, 0, ctx->stream>>>(sums, nt, with_total ? data + n : nullptr);
  XG_LAUNCH_CHECK();
  k_scan_add<<<nt, SCAN_BLK, 0, ctx->stream>>>(data, n, sums);
  XG_LAUNCH_CHECK();
  return XG_OK;
}

}  // namespace xg

namespace xg {

size_t mark_scratch_ints(int ne) {
  const int nblk = (ne + BLK - 1) / BLK;
  return (size_t)(nblk + 1) + (size_t)((nblk + SCAN_TILE - 1) / SCAN_TILE + 1);
}

// marking + compaction on an explicit stream with caller-owned scratch
// (mark_scratch_ints(ne) int32); only kernel launches, so the sequence can be
// captured into a hipGraph.  n_cut (nullable) receives the cut count.
int mark_compact(hipStream_t st, int32_t *d_err, int D, int ne, const int32_t *el2v, const double *lset,
                 uint8_t *dt, double *cr, uint32_t *bits, int32_t *scratch, int32_t *cut_list, int32_t *n_cut) {
  const int nw = (ne + 31) / 32;
  const int nblk = (ne + MARK_BLK * MARK_EPT - 1) / (MARK_BLK * MARK_EPT);
  int32_t *blk = scratch, *sums = scratch + nblk + 1;
  switch (D) {
    case 1: k_mark<1, MARK_BLK, MARK_EPT><<<nblk, MARK_BLK, 0, st>>>(ne, el2v, lset, dt, cr, bits, nw, blk, d_err); break;
    case 2: k_mark<2, MARK_BLK, MARK_EPT><<<nblk, MARK_BLK, 0, st>>>(ne, el2v, lset, dt, cr, bits, nw, blk, d_err); break;
    default: k_mark<3, MARK_BLK, MARK_EPT><<<nblk, MARK_BLK, 0, st>>>(ne, el2v, lset, dt, cr, bits, nw, blk, d_err); break;
  }
  XG_LAUNCH_CHECK();
  scan_block_counts(st, blk, nblk, sums, n_cut);
  k_compact<MARK_BLK, MARK_EPT><<<nblk, MARK_BLK, 0, st>>>(ne, bits + (size_t)4 * nw, blk, cut_list);
  XG_LAUNCH_CHECK();
  return XG_OK;
}

}  // namespace xg

namespace xg {

// the 8 bit arrays, the cut list (IF elements, ascending) and its length
// from per-element domain types (k_bits_from_dt, block counts, scan,
// compaction); cut_list and n_cut nullable
int bits_and_cut_list_from_dt(xg_ctx *ctx, int ne, const uint8_t *dt, uint32_t *bits, int32_t *cut_list,
                              int32_t *n_cut) {
  const int nw = (ne + 31) / 32;
  const int nblk = (ne + BLK - 1) / BLK;
  const bool compact = cut_list || n_cut;
  int32_t *blk = nullptr;
  if (compact) {
    // block counts after the scan scratch, as in xg_mark
    const size_t scan_ws = sizeof(int32_t) * (size_t)((nblk + 1 + SCAN_TILE - 1) / SCAN_TILE + 1);
    int r = ensure_ws(ctx, scan_ws + sizeof(int32_t) * (size_t)(nblk + 1) + 256);
    if (r) return r;
    blk = (int32_t *)((char *)ctx->ws + ((scan_ws + 255) / 256) * 256);
  }
  k_bits_from_dt<<<nblk, BLK, 0, ctx->stream>>>(ne, dt, bits, nw, blk);
  XG_LAUNCH_CHECK();
  if (!compact) return XG_OK;
  scan_block_counts(ctx->stream, blk, nblk, (int32_t *)ctx->ws, n_cut);
  XG_LAUNCH_CHECK();
  if (cut_list) {
    k_compact<BLK, 1><<<nblk, BLK, 0, ctx->stream>>>(ne, bits + (size_t)4 * nw, blk, cut_list);
    XG_LAUNCH_CHECK();
  }
  return XG_OK;
}

}  // namespace xg

extern "C" int xg_p1_perturb(xg_ctx *ctx, int n, double *v, double eps) {
  if (!ctx || n < 0 || (n > 0 && !v)) return set_error(XG_ERR_ARG, "xg_p1_perturb: bad argument");
  if (n == 0) return XG_OK;
  k_perturb<<<(n + 255) / 256, 256, 0, ctx->stream>>>(n, v, eps);
  XG_LAUNCH_CHECK();
  return XG_OK;
}

extern "C" int xg_mark(xg_ctx *ctx, int D, int ne, const int32_t *el2v, const double *lset_v,
                       uint8_t *dt, double *cut_ratio, uint32_t *bits, int32_t *cut_list,
                       int32_t *n_cut) {
  if (!ctx || D < 1 || D > 3 || ne < 0 || !bits || (ne > 0 && (!el2v || !lset_v)))
    return set_error(XG_ERR_ARG, "xg_mark: bad argument");
  const int nw = (ne + 31) / 32;
  if (ne == 0) {
    if (n_cut) XG_CHECK(hipMemsetAsync(n_cut, 0, sizeof(int32_t), ctx->stream));
    return XG_OK;
  }
  const int nblk = (ne + MARK_BLK * MARK_EPT - 1) / (MARK_BLK * MARK_EPT);
  const bool compact = cut_list || n_cut;
  // block counts live after the scan scratch: [0, nblk] counts (+ total)
  int32_t *blk = nullptr;
  if (compact) {
    const size_t scan_ws = sizeof(int32_t) * (size_t)((nblk + 1 + SCAN_TILE - 1) / SCAN_TILE + 1);
    int r = ensure_ws(ctx, scan_ws + sizeof(int32_t) * (size_t)(nblk + 1) + 256);
    if (r) return r;
    blk = (int32_t *)((char *)ctx->ws + ((scan_ws + 255) / 256) * 256);
  }
  switch (D) {
    case 1: k_mark<1, MARK_BLK, MARK_EPT><<<nblk, MARK_BLK, 0, ctx->stream>>>(ne, el2v, lset_v, dt, cut_ratio, bits, nw, blk, ctx->d_err); break;
    case 2: k_mark<2, MARK_BLK, MARK_EPT><<<nblk, MARK_BLK, 0, ctx->stream>>>(ne, el2v, lset_v, dt, cut_ratio, bits, nw, blk, ctx->d_err); break;
    default: k_mark<3, MARK_BLK, MARK_EPT><<<nblk, MARK_BLK, 0, ctx->stream>>>(ne, el2v, lset_v, dt, cut_ratio, bits, nw, blk, ctx->d_err); break;
  }
  XG_LAUNCH_CHECK();
  if (!compact) return XG_OK;
  // scan of block counts: blk[nblk] = total.  The scan scratch is the start
  // of ctx->ws, disjoint from blk.
  scan_block_counts(ctx->stream, blk, nblk, (int32_t *)ctx->ws, n_cut);
  XG_LAUNCH_CHECK();
  if (cut_list) {
    k_compact<MARK_BLK, MARK_EPT><<<nblk, MARK_BLK, 0, ctx->stream>>>(ne, bits + (size_t)4 * nw, blk, cut_list);
    XG_LAUNCH_CHECK();
  }
  return XG_OK;
}

// ---------------------------------------------------------------- no-alloc
// xg_mark without the context workspace: the block counts and the scan
// partials live in one file-scope device array.  Every GPU that loads the
// library holds its own copy; its address is looked up once per device
// (hipGetSymbolAddress with that device current) and only that address is
// passed to kernels.  k_mark writes every block count it later reads and the
// scan writes its partials before reading them, so nothing has to start at
// zero.  Calls that overlap on one GPU share the array and race.
namespace {
constexpr int NOALLOC_MAX_NE = XG_MARK_NOALLOC_MAX_NE;
constexpr int NOALLOC_MAX_BLK = NOALLOC_MAX_NE / (MARK_BLK * MARK_EPT);
constexpr int NOALLOC_INTS = NOALLOC_MAX_BLK + 1 + (NOALLOC_MAX_BLK + 1 + SCAN_TILE - 1) / SCAN_TILE + 1;
static __device__ __align__(256) int32_t scratch[NOALLOC_INTS];

constexpr int NOALLOC_MAX_DEV = 64;
std::mutex g_scratch_mu;
int32_t *g_scratch_ptr[NOALLOC_MAX_DEV] = {};

__global__ void k_zero_i32(int32_t *p) { *p = 0; }
}  // namespace

namespace xg {
// the device address of `scratch` on `device` (looked up on first use, with
// that device current)
int noalloc_scratch(int device, int32_t **out) {
  if (device < 0 || device >= NOALLOC_MAX_DEV) return set_error(XG_ERR_ARG, "no-alloc scratch: device id too large");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = 0;
    XG_CHECK(hipGetDevice(&cur));
    if (cur != device) XG_CHECK(hipSetDevice(device));
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != device) XG_CHECK(hipSetDevice(cur));
    if (e != hipSuccess) return check_hip(e, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[device] = (int32_t *)ptr;
  }
  *out = g_scratch_ptr[device];
  return XG_OK;
}
}  // namespace xg

extern "C" int xg_mark_noalloc(xg_ctx *ctx, int D, int ne, const int32_t *el2v, const double *lset_v, uint8_t *dt,
                               double *cut_ratio, uint32_t *bits, int32_t *cut_list, int32_t *n_cut) {
  if (!ctx || D < 1 || D > 3 || ne < 0 || !bits || (ne > 0 && (!el2v || !lset_v)))
    return set_error(XG_ERR_ARG, "xg_mark_noalloc: bad argument");
  if (ne > NOALLOC_MAX_NE)
    return set_error(XG_ERR_CAPACITY, "xg_mark_noalloc: more than XG_MARK_NOALLOC_MAX_NE (2^28) elements; "
                                      "use xg_mark");
  int32_t *scr = ctx->noalloc_scratch;
  if (!scr) {
    const int r = noalloc_scratch(ctx->device, &scr);
    if (r != XG_OK) return r;
    ctx->noalloc_scratch = scr;
  }
  if (ne == 0) {
    if (n_cut) {
      k_zero_i32<<<1, 1, 0, ctx->stream>>>(n_cut);
      XG_LAUNCH_CHECK();
    }
    return XG_OK;
  }
  const int nw = (ne + 31) / 32;
  const int nblk = (ne + MARK_BLK * MARK_EPT - 1) / (MARK_BLK * MARK_EPT);
  const bool compact = cut_list || n_cut;
  int32_t *blk = compact ? scr : nullptr, *sums = scr + nblk + 1;
  switch (D) {
    case 1: k_mark<1, MARK_BLK, MARK_EPT><<<nblk, MARK_BLK, 0, ctx->stream>>>(ne, el2v, lset_v, dt, cut_ratio, bits, nw, blk, ctx->d_err); break;
    case 2: k_mark<2, MARK_BLK, MARK_EPT><<<nblk, MARK_BLK, 0, ctx->stream>>>(ne, el2v, lset_v, dt, cut_ratio, bits, nw, blk, ctx->d_err); break;
    default: k_mark<3, MARK_BLK, MARK_EPT><<<nblk, MARK_BLK, 0, ctx->stream>>>(ne, el2v, lset_v, dt, cut_ratio, bits, nw, blk, ctx->d_err); break;
  }
  XG_LAUNCH_CHECK();
  if (!compact) return XG_OK;
  scan_block_counts(ctx->stream, blk, nblk, sums, n_cut);
  XG_LAUNCH_CHECK();
  if (cut_list) {
    k_compact<MARK_BLK, MARK_EPT><<<nblk, MARK_BLK, 0, ctx->stream>>>(ne, bits + (size_t)4 * nw, blk, cut_list);
    XG_LAUNCH_CHECK();
  }
  return XG_OK;
}

extern "C" int xg_facets_with_neighbor_types(xg_ctx *ctx, int nf, const int32_t *f2el,
                                             const uint32_t *a, const uint32_t *b, int bnd_val_a,
                                             int bnd_val_b, int use_and, uint32_t *out) {
  if (!ctx || nf < 0 || (nf > 0 && (!f2el || !a || !b || !out)))
    return set_error(XG_ERR_ARG, "xg_facets_with_neighbor_types: bad argument");
  if (nf == 0) return XG_OK;
  const int nwf = (nf + 31) / 32;
  k_facets<<<(nf + 255) / 256, 256, 0, ctx->stream>>>(nf, f2el, a, b, bnd_val_a, bnd_val_b, use_and,
                                                      out, nwf);
  XG_LAUNCH_CHECK();
  return XG_OK;
}

extern "C" int xg_mark_p2_trig(xg_ctx *ctx, int ne, const double *lset_nodal, int L, uint8_t *dt,
                               double *cut_ratio, uint32_t *bits, int32_t *cut_list, int32_t *n_cut) {
  if (!ctx || ne < 0 || L < 0 || L > XG_LAT_LMAX || !dt || !cut_ratio || !bits || (ne > 0 && !lset_nodal))
    return set_error(XG_ERR_ARG, "xg_mark_p2_trig: bad argument (dt, cut_ratio and bits required, L <= 5)");
  if (ne == 0) {
    if (n_cut) XG_CHECK(hipMemsetAsync(n_cut, 0, sizeof(int32_t), ctx->stream));
    return XG_OK;
  }
  const int nblk = (ne + BLK - 1) / BLK;
  k_mark_p2_trig_bern<<<nblk, BLK, 0, ctx->stream>>>(ne, lset_nodal, dt, cut_ratio);
  k_mark_p2_trig<<<(unsigned)(((long long)ne * 64 + BLK - 1) / BLK), BLK, 0, ctx->stream>>>(ne, lset_nodal, L, dt,
                                                                                       cut_ratio);
  XG_LAUNCH_CHECK();
  return xg::bits_and_cut_list_from_dt(ctx, ne, dt, bits, cut_list, n_cut);
}

extern "C" int xg_mark_p2(xg_ctx *ctx, int ne, const double *lset_nodal, int L, uint8_t *dt, uint32_t *bits,
                          int32_t *cut_list, int32_t *n_cut) {
  if (!ctx || ne < 0 || L < 0 || L > XG_LAT_LMAX || !dt || !bits || (ne > 0 && !lset_nodal))
    return set_error(XG_ERR_ARG, "xg_mark_p2: bad argument (dt and bits required, L <= 5)");
  if (ne == 0) {
    if (n_cut) XG_CHECK(hipMemsetAsync(n_cut, 0, sizeof(int32_t), ctx->stream));
    return XG_OK;
  }
  const int nblk = (ne + BLK - 1) / BLK;
  k_mark_p2_bern<<<nblk, BLK, 0, ctx->stream>>>(ne, lset_nodal, dt);
  k_mark_p2_lattice<<<nblk, BLK, 0, ctx->stream>>>(ne, lset_nodal, L, dt);
  XG_LAUNCH_CHECK();
  return xg::bits_and_cut_list_from_dt(ctx, ne, dt, bits, cut_list, n_cut);
}
