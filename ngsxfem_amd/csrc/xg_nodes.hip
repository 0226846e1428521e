// xg_nodes.hip -- marking of mesh nodes and dofs from element or facet marks.
//
// Reference: xfem/cutinfo.cpp:135-224 (cut_neighboring_node, dom_of_node in
//            CutInformation::Update), :540-564 (GetElementsWithNeighborFacets),
//            :566-596 (GetElementsWithSharedVertex), :598-622
//            (GetDofsOfElements), :624-647 (GetDofsOfFacets).
//
// Every one of these is a scatter of one bit (or one byte) per (item, node)
// pair.  One lane per item; its k node numbers are contiguous in the item's
// table row; a bit is set with a 32-bit atomicOr on the word that holds it
// (the reference's SetBitAtomic), a byte with a plain byte store (all writers
// of one byte store the same value).  Bound by HBM: 4k bytes of table plus one
// mark bit per item; the bit-array words are few and stay in L2.
#include "xg_internal.hpp"

namespace {

constexpr int BLK = 256;
constexpr int ERR_RANGE = 8;  // a node number >= the node count

__device__ __forceinline__ bool bit_of(const uint32_t *b, int e) { return (b[e >> 5] >> (e & 31)) & 1u; }

__global__ void k_fill_u32(uint32_t *p, int n, uint32_t v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_u8(uint8_t *p, int n, uint8_t v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// out bit d for every node d of a marked item (mark == nullptr: every item).
// Negative node numbers are "no node" and skipped; d >= nn sets the error word.
__global__ __launch_bounds__(BLK) void k_scatter_bits(int n, int k, const int32_t *__restrict__ item2node,
                                                      const uint32_t *__restrict__ mark, int nn,
                                                      uint32_t *out, int32_t *err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (mark && !bit_of(mark, i)) return;
  const int32_t *row = item2node + (size_t)i * k;
  for (int j = 0; j < k; ++j) {
    const int d = row[j];
    if (d < 0) continue;
    if (d >= nn) {
      atomicOr(err, ERR_RANGE);
      continue;
    }
    atomicOr(&out[d >> 5], 1u << (d & 31));
  }
}

// node arrays of one node type from the element domain codes
// (cutinfo.cpp:135-210): nodes of IF elements -> cut-neighbour bit; nodes of
// NEG elements -> dom NEG, all other nodes stay POS.
__global__ __launch_bounds__(BLK) void k_node_domains(int ne, int k, const int32_t *__restrict__ el2node,
                                                      const uint8_t *__restrict__ dt, int nn, uint32_t *cutnb,
                                                      uint8_t *dom, int32_t *err) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int d = dt[e];
  if (d == 1) return;  // POS: no mark
  const int32_t *row = el2node + (size_t)e * k;
  for (int j = 0; j < k; ++j) {
    const int v = row[j];
    if (v < 0) continue;
    if (v >= nn) {
      atomicOr(err, ERR_RANGE);
      continue;
    }
    if (d == 2) {
      if (cutnb) atomicOr(&cutnb[v >> 5], 1u << (v & 31));
    } else if (dom) {
      dom[v] = 0;  // NEG
    }
  }
}

// the element level (the cell nodes): cut-neighbour bit |= IF, dom by the
// cut ratio (cutinfo.cpp:205-221).  One wave ORs 64 bits from one ballot
// into the two words it owns.
__global__ __launch_bounds__(BLK) void k_cell_domains(int ne, const uint8_t *__restrict__ dt,
                                                      const double *__restrict__ cr, uint32_t *cutnb,
                                                      uint8_t *dom) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  bool is_if = false;
  if (e < ne) {
    is_if = dt[e] == 2;
    if (dom) dom[e] = cr[e] > 0.5 ? 0 : 1;
  }
  const unsigned long long m = __ballot(is_if);
  if (cutnb && (threadIdx.x & 63) == 0) {
    const int w = (e - (int)(threadIdx.x & 63)) >> 5;
    const int nw = (ne + 31) >> 5;
    if (w < nw) cutnb[w] |= (uint32_t)(m & 0xffffffffull);
    if (w + 1 < nw) cutnb[w + 1] |= (uint32_t)(m >> 32);
  }
}

// GetElementsWithSharedVertex, second pass: element bit = any vertex marked
template <int NV>
__global__ __launch_bounds__(BLK) void k_elems_touching(int ne, const int32_t *__restrict__ el2v,
                                                        const uint32_t *__restrict__ vmark, int nv,
                                                        uint32_t *out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  bool hit = false;
  if (e < ne) {
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      const int v = el2v[(size_t)e * NV + j];
      hit |= (unsigned)v < (unsigned)nv && bit_of(vmark, v);
    }
  }
  const unsigned long long m = __ballot(hit);
  if ((threadIdx.x & 63) == 0) {
    const int w = (e - (int)(threadIdx.x & 63)) >> 5;
    const int nw = (ne + 31) >> 5;
    if (w < nw) out[w] = (uint32_t)(m & 0xffffffffull);
    if (w + 1 < nw) out[w + 1] = (uint32_t)(m >> 32);
  }
}

inline unsigned nblk(long long n) { return (unsigned)((n + BLK - 1) / BLK); }

int clear_bits(xg_ctx *ctx, uint32_t *bits, int nbits) {
  const int nw = (nbits + 31) / 32;
  if (nw) k_fill_u32<<<nblk(nw), BLK, 0, ctx->stream>>>(bits, nw, 0u);
  XG_LAUNCH_CHECK();
  return XG_OK;
}

}  // namespace

using xg::set_error;

extern "C" int xg_mark_from_items(xg_ctx *ctx, int n, int k, const int32_t *item2node, const uint32_t *mark,
                                  int nn, uint32_t *out) {
  if (!ctx || n < 0 || k < 1 || nn < 0 || (nn > 0 && !out) || (n > 0 && !item2node))
    return set_error(XG_ERR_ARG, "xg_mark_from_items: bad argument");
  int r = clear_bits(ctx, out, nn);
  if (r) return r;
  if (n == 0) return XG_OK;
  k_scatter_bits<<<nblk(n), BLK, 0, ctx->stream>>>(n, k, item2node, mark, nn, out, ctx->d_err);
  XG_LAUNCH_CHECK();
  return XG_OK;
}

extern "C" int xg_elements_with_neighbor_facets(xg_ctx *ctx, int nf, const int32_t *f2el,
                                                const uint32_t *facet_mark, int ne, uint32_t *out) {
  if (!ctx || nf < 0 || ne < 0 || (ne > 0 && !out) || (nf > 0 && (!f2el || !facet_mark)))
    return set_error(XG_ERR_ARG, "xg_elements_with_neighbor_facets: bad argument");
  return xg_mark_from_items(ctx, nf, 2, f2el, facet_mark, ne, out);
}

extern "C" int xg_elements_with_shared_vertex(xg_ctx *ctx, int k, int ne, const int32_t *el2v,
                                              const uint32_t *elem_mark, int nv, uint32_t *out) {
  if (!ctx || (k != 2 && k != 3 && k != 4 && k != 8) || ne < 0 || nv < 0 ||
      (ne > 0 && (!el2v || !elem_mark || !out)))
    return set_error(XG_ERR_ARG, "xg_elements_with_shared_vertex: bad argument (k in 2, 3, 4, 8)");
  if (ne == 0) return XG_OK;
  // vertex marks in the context scratch
  const int nwv = (nv + 31) / 32;
  int r = xg::ensure_ws(ctx, sizeof(uint32_t) * (size_t)(nwv > 0 ? nwv : 1));
  if (r) return r;
  uint32_t *vmark = (uint32_t *)ctx->ws;
  r = xg_mark_from_items(ctx, ne, k, el2v, elem_mark, nv, vmark);
  if (r) return r;
  switch (k) {
    case 2: k_elems_touching<2><<<nblk(ne), BLK, 0, ctx->stream>>>(ne, el2v, vmark, nv, out); break;
    case 3: k_elems_touching<3><<<nblk(ne), BLK, 0, ctx->stream>>>(ne, el2v, vmark, nv, out); break;
    case 4: k_elems_touching<4><<<nblk(ne), BLK, 0, ctx->stream>>>(ne, el2v, vmark, nv, out); break;
    default: k_elems_touching<8><<<nblk(ne), BLK, 0, ctx->stream>>>(ne, el2v, vmark, nv, out); break;
  }
  XG_LAUNCH_CHECK();
  return XG_OK;
}

extern "C" int xg_node_domains(xg_ctx *ctx, int ne, int k, const int32_t *el2node, const uint8_t *dt,
                               const double *cut_ratio, int nn, uint32_t *cutnb, uint8_t *dom) {
  if (!ctx || ne < 0 || nn < 0 || (ne > 0 && !dt))
    return set_error(XG_ERR_ARG, "xg_node_domains: bad argument");
  if (!el2node) {
    // the cells themselves
    if (nn != ne || (dom && !cut_ratio))
      return set_error(XG_ERR_ARG, "xg_node_domains: cells need nn == ne and the cut ratios");
    if (ne == 0) return XG_OK;
    k_cell_domains<<<nblk(ne), BLK, 0, ctx->stream>>>(ne, dt, cut_ratio, cutnb, dom);
    XG_LAUNCH_CHECK();
    return XG_OK;
  }
  if (k < 1) return set_error(XG_ERR_ARG, "xg_node_domains: k < 1");
  // cutnb is not cleared: the reference clears it only when the
  // CutInformation is constructed (cutinfo.cpp:28-34), so every Update adds
  // the nodes of its cut elements to those of the earlier ones
  if (dom && nn) k_fill_u8<<<nblk(nn), BLK, 0, ctx->stream>>>(dom, nn, (uint8_t)1);  // POS
  if (ne) k_node_domains<<<nblk(ne), BLK, 0, ctx->stream>>>(ne, k, el2node, dt, nn, cutnb, dom, ctx->d_err);
  XG_LAUNCH_CHECK();
  return XG_OK;
}
