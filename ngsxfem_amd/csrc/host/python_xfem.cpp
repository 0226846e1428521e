// python_xfem.cpp -- pybind11 module `xfem` of the MI355X build: the host
// objects of the cut-element path over the C ABI (include/xg_abi.h).
//
// Mirrors the reference bindings:
//   CutInfo                      xfem/python_xfem.cpp:47-199 (CutInformation,
//                                xfem/cutinfo.cpp:14-133, 263-284)
//   GetFacetsWithNeighborTypes   xfem/python_xfem.cpp:346-405 (cutinfo.cpp:476-538)
//   GetElementsWithNeighborFacets, GetElementsWithSharedVertex,
//   GetDofsOfElements, GetDofsOfFacets
//                                xfem/python_xfem.cpp:407-490 (cutinfo.cpp:540-647)
//   IntegrateX                   cutint/python_cutint.cpp:21-225
//   InterpolateToP1 (perturb)    utils/python_utils.cpp:134,166 (p1interpol.cpp:104-107)
//   enums                        python/python_ngsxfem.cpp:15-46
// plus CutElementMatrices, the element-matrix kernel behind
// SymbolicCutBFI / dCut bilinear forms (xfem/symboliccutbfi.cpp:37-486).
//
// NGSolve is not part of this build: meshes, grid functions and coefficient
// functions are the lightweight Python stand-ins of ngsxfem_amd.xfem; this
// module reads their numpy arrays.  All heavy work runs in HIP kernels;
// there is no CPU fallback.
#include <hip/hip_runtime_api.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <array>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/xg_abi.h"

namespace py = pybind11;

namespace {

enum DOMAIN_TYPE { NEG = 0, POS = 1, IF = 2 };
enum COMBINED_DOMAIN_TYPE { CDOM_NO = 0, CDOM_NEG = 1, CDOM_POS = 2, CDOM_UNCUT = 3, CDOM_IF = 4,
                            CDOM_HASNEG = 5, CDOM_HASPOS = 6, CDOM_ANY = 7 };
enum SWAP_DIMENSIONS_POLICY { FIRST_ALLOWED = 0, FIND_OPTIMAL = 1, ALWAYS_NONE = 2 };
enum VorB { VOL = 0, BND = 1 };
enum NODE_TYPE { NT_VERTEX = 0, NT_EDGE = 1, NT_FACE = 2, NT_CELL = 3, NT_ELEMENT = 4, NT_FACET = 5 };

int to_cdt(DOMAIN_TYPE a) { return a == NEG ? CDOM_NEG : (a == POS ? CDOM_POS : CDOM_IF); }

void check(int rc) {
  if (rc != XG_OK) throw std::runtime_error(std::string("xg: ") + xg_last_error());
}
void hipcheck(hipError_t e, const char *what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

xg_ctx *ctx() {
  static thread_local xg_ctx *c = nullptr;
  if (!c) check(xg_ctx_create(0, &c));
  return c;
}

template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t n_) { resize(n_); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() {
    if (p) hipFree(p);
  }
  void resize(size_t n_) {
    if (n_ <= n && p) return;
    if (p) hipFree(p);
    p = nullptr;
    n = n_;
    if (n_) hipcheck(hipMalloc(&p, sizeof(T) * n_), "hipMalloc");
  }
  void upload(const T *h, size_t cnt) {
    resize(cnt);
    if (cnt) hipcheck(hipMemcpy(p, h, sizeof(T) * cnt, hipMemcpyHostToDevice), "upload");
  }
  void download(T *h, size_t cnt) const {
    if (cnt) hipcheck(hipMemcpy(h, p, sizeof(T) * cnt, hipMemcpyDeviceToHost), "download");
  }
};

// ---------------------------------------------------------------- BitArray
// Same memory layout as NGSolve's BitArray: bit i of byte i/8 (LSB first);
// stored as uint32 words.
class BitArray {
 public:
  explicit BitArray(size_t n) : n_(n), w_((n + 31) / 32, 0u) {}
  size_t Size() const { return n_; }
  bool Test(size_t i) const { return (w_.at(i >> 5) >> (i & 31)) & 1u; }
  void SetBit(size_t i) { w_.at(i >> 5) |= 1u << (i & 31); }
  void ClearBit(size_t i) { w_.at(i >> 5) &= ~(1u << (i & 31)); }
  void Set() {
    for (auto &x : w_) x = ~0u;
    trim();
  }
  void Clear() {
    for (auto &x : w_) x = 0u;
  }
  size_t NumSet() const {
    size_t c = 0;
    for (auto x : w_) c += __builtin_popcount(x);
    return c;
  }
  std::vector<uint32_t> &words() { return w_; }
  const std::vector<uint32_t> &words() const { return w_; }
  void trim() {
    if (n_ & 31) w_.back() &= (1u << (n_ & 31)) - 1u;
  }

 private:
  size_t n_;
  std::vector<uint32_t> w_;
};
using BA = std::shared_ptr<BitArray>;

// ---------------------------------------------------------------- mesh
struct MeshAccess {
  py::object pymesh;
  int dim = 0, nv = 0, ne = 0, nse = 0;
  int nvel = 0, nvbnd = 0;  // vertices per element / boundary element
  bool quad = false;        // quadrilaterals / hexahedra (2^dim vertices)
  DevBuf<double> coords;
  DevBuf<int32_t> el2v, bnd2v, f2el, el2f, el2e;
  int nf = -1, nedges = -1, nf_el = 0, ne_el = 0;
  std::vector<int32_t> h_el2v, h_bnd2v;

  explicit MeshAccess(py::object m) : pymesh(m) {
    auto c = py::array_t<double, py::array::c_style | py::array::forcecast>(m.attr("coords"));
    auto e = py::array_t<int32_t, py::array::c_style | py::array::forcecast>(m.attr("el2v"));
    dim = (int)c.shape(1);
    nv = (int)c.shape(0);
    ne = (int)e.shape(0);
    nvel = (int)e.shape(1);
    quad = dim >= 2 && nvel == (1 << dim);
    if (!quad && nvel != dim + 1) throw std::runtime_error("mesh: simplices or quadrilaterals / hexahedra only");
    coords.upload(c.data(), (size_t)nv * dim);
    el2v.upload(e.data(), (size_t)ne * nvel);
    h_el2v.assign(e.data(), e.data() + (size_t)ne * nvel);
    auto b = py::array_t<int32_t, py::array::c_style | py::array::forcecast>(m.attr("bnd2v"));
    nse = (int)b.shape(0);
    nvbnd = b.ndim() == 2 ? (int)b.shape(1) : dim;
    if (nse) bnd2v.upload(b.data(), (size_t)nse * nvbnd);
    h_bnd2v.assign(b.data(), b.data() + (size_t)nse * nvbnd);
  }
  int GetNE(VorB vb) const { return vb == VOL ? ne : nse; }
  void ensure_facets() {
    if (nf >= 0) return;
    py::tuple t = pymesh.attr("facets")();
    auto f = py::array_t<int32_t, py::array::c_style | py::array::forcecast>(t[2]);
    nf = (int)f.shape(0);
    f2el.upload(f.data(), (size_t)nf * 2);
    auto ef = py::array_t<int32_t, py::array::c_style | py::array::forcecast>(t[1]);
    nf_el = (int)ef.shape(1);
    el2f.upload(ef.data(), (size_t)ef.size());
  }
  void ensure_edges() {
    if (nedges >= 0) return;
    py::tuple t = pymesh.attr("edges")();
    auto ev = py::array_t<int32_t, py::array::c_style | py::array::forcecast>(t[0]);
    auto ee = py::array_t<int32_t, py::array::c_style | py::array::forcecast>(t[1]);
    nedges = (int)ev.shape(0);
    ne_el = (int)ee.shape(1);
    el2e.upload(ee.data(), (size_t)ee.size());
  }
  // NGSolve's GetNNodes for NT_VERTEX, NT_EDGE, NT_FACE, NT_CELL
  int nnodes(int nt) {
    switch (nt) {
      case NT_VERTEX: return nv;
      case NT_EDGE:
        if (dim == 1) return ne;
        ensure_edges();
        return nedges;
      case NT_FACE:
        if (dim == 2) return ne;
        if (dim == 3) {
          ensure_facets();
          return nf;
        }
        return 0;
      case NT_CELL: return dim == 3 ? ne : 0;
    }
    return 0;
  }
  // element -> node table of node type nt, or nullptr where the nodes of
  // that type are the elements themselves
  const int32_t *el2node(int nt, int &k) {
    if (nt == NT_VERTEX) {
      k = nvel;
      return el2v.p;
    }
    if (nt == NT_EDGE && dim >= 2) {
      ensure_edges();
      k = ne_el;
      return el2e.p;
    }
    if (nt == NT_FACE && dim == 3) {
      ensure_facets();
      k = nf_el;
      return el2f.p;
    }
    k = 0;
    return nullptr;
  }
};
using MA = std::shared_ptr<MeshAccess>;

std::shared_ptr<MeshAccess> as_mesh(py::object m) {
  if (py::isinstance<MeshAccess>(m)) return m.cast<MA>();
  if (py::hasattr(m, "_xg_access")) {
    py::object a = m.attr("_xg_access");
    if (!a.is_none()) return a.cast<MA>();
  }
  auto ma = std::make_shared<MeshAccess>(m);
  try {
    m.attr("_xg_access") = py::cast(ma);
  } catch (...) {
  }
  return ma;
}

// vertex values of a P1 level set: a grid function stand-in (attribute
// `vec` and `order`) or any array of vertex values
// the paths built for simplices only (element matrices, subdivision and
// multi-level-set rules); quadrilateral / hexahedral meshes have the straight
// cut rule, marking and IntegrateX
void simplices_only(MA ma, const char *what) {
  if (ma->quad)
    throw std::runtime_error(std::string(what) + ": simplicial meshes only in this build "
                             "(quadrilaterals / hexahedra: CutInfo and Integrate with P1 level sets)");
}

int parse_policy(py::dict lsetdom) {
  if (!lsetdom.contains("quad_dir_policy")) return FIND_OPTIMAL;  // xintegration.hpp:39
  py::object o = lsetdom["quad_dir_policy"];
  const int v = py::isinstance<py::int_>(o) ? o.cast<int>() : (int)o.cast<SWAP_DIMENSIONS_POLICY>();
  if (v < 0 || v > 2) throw std::runtime_error("quad_dir_policy: FIRST, OPTIMAL or FALLBACK");
  return v;
}

std::vector<double> p1_values(py::object lset, int nv, bool &is_p1) {
  py::object vec = py::hasattr(lset, "vec") ? lset.attr("vec") : lset;
  int order = py::hasattr(lset, "order") ? lset.attr("order").cast<int>() : 1;
  auto a = py::array_t<double, py::array::c_style | py::array::forcecast>(vec);
  is_p1 = order <= 1;
  if (is_p1 && a.size() != nv) throw std::runtime_error("level set: expected one value per vertex");
  return std::vector<double>(a.data(), a.data() + a.size());
}

// Level sets that need the subdivision rule (xintegration.cpp:1626-1754):
// P2 grid functions, or any level set with subdivlvl > 0.  Their element
// nodal values in this library's P2 dof order (vertices, then the edges
// (0,1),(0,2),(0,3),(1,2),(1,3),(2,3)) for the listed elements; a P1 level
// set is lifted to P2 by its edge-midpoint averages.
bool needs_subdivision(py::object lset, int subdivlvl) {
  const int order = py::hasattr(lset, "order") ? lset.attr("order").cast<int>() : 1;
  if (order > 2) throw std::runtime_error("level sets of order > 2 are not supported");
  return order == 2 || subdivlvl > 0;
}

std::vector<double> p2_nodal_values(MA ma, py::object lset, const std::vector<int32_t> &el) {
  if (ma->dim != 3) throw std::runtime_error("subdivision rules (P2 level sets, subdivlvl > 0): tetrahedra only");
  const int order = py::hasattr(lset, "order") ? lset.attr("order").cast<int>() : 1;
  py::object vec = py::hasattr(lset, "vec") ? lset.attr("vec") : lset;
  auto a = py::array_t<double, py::array::c_style | py::array::forcecast>(vec);
  const double *v = a.data();
  std::vector<double> out(el.size() * 10);
  static const int ed[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
  if (order == 2) {
    py::tuple t = ma->pymesh.attr("dofs")(2);
    auto dofs = py::array_t<int32_t, py::array::c_style | py::array::forcecast>(t[0]);
    const int ndof = t[1].cast<int>();
    if (a.size() != ndof) throw std::runtime_error("P2 level set: expected one value per P2 dof");
    const int32_t *d = dofs.data();
    for (size_t i = 0; i < el.size(); ++i)
      for (int k = 0; k < 10; ++k) out[i * 10 + k] = v[d[(size_t)el[i] * 10 + k]];
  } else {
    if (a.size() != ma->nv) throw std::runtime_error("level set: expected one value per vertex");
    for (size_t i = 0; i < el.size(); ++i) {
      const int32_t *vv = &ma->h_el2v[(size_t)el[i] * 4];
      for (int k = 0; k < 4; ++k) out[i * 10 + k] = v[vv[k]];
      for (int k = 0; k < 6; ++k) out[i * 10 + 4 + k] = 0.5 * (v[vv[ed[k][0]]] + v[vv[ed[k][1]]]);
    }
  }
  return out;
}

// F^{-1} (row-major, 3 x 3) of the listed tetrahedra, for the interface
// weights of the subdivision rule
std::vector<double> tet_finv(MA ma, const std::vector<int32_t> &el) {
  auto c = py::array_t<double, py::array::c_style | py::array::forcecast>(ma->pymesh.attr("coords"));
  const double *X = c.data();
  std::vector<double> out(el.size() * 9);
  for (size_t i = 0; i < el.size(); ++i) {
    const int32_t *vv = &ma->h_el2v[(size_t)el[i] * 4];
    double F[9];
    for (int r = 0; r < 3; ++r)
      for (int q = 0; q < 3; ++q) F[r * 3 + q] = X[(size_t)vv[q] * 3 + r] - X[(size_t)vv[3] * 3 + r];
    const double A = F[4] * F[8] - F[5] * F[7], B = -(F[3] * F[8] - F[5] * F[6]), C = F[3] * F[7] - F[4] * F[6];
    const double det = F[0] * A + F[1] * B + F[2] * C;
    double *o = &out[i * 9];
    o[0] = A / det; o[1] = -(F[1] * F[8] - F[2] * F[7]) / det; o[2] = (F[1] * F[5] - F[2] * F[4]) / det;
    o[3] = B / det; o[4] = (F[0] * F[8] - F[2] * F[6]) / det; o[5] = -(F[0] * F[5] - F[2] * F[3]) / det;
    o[6] = C / det; o[7] = -(F[0] * F[7] - F[1] * F[6]) / det; o[8] = (F[0] * F[4] - F[1] * F[3]) / det;
  }
  return out;
}

// ---------------------------------------------------------------- CutInfo
class CutInformation {
 public:
  MA ma;
  std::array<BA, 8> elems, selems, facets;
  std::vector<double> cut_ratio[2];
  DevBuf<double> d_lset;
  // cut_neighboring_node / dom_of_node of NT_VERTEX .. NT_CELL
  // (cutinfo.hpp:30-33); NT_ELEMENT and NT_FACET alias them by dimension
  std::array<BA, 4> cutnb;
  std::array<std::vector<uint8_t>, 4> dom;

  explicit CutInformation(MA m) : ma(m) {
    // cutinfo.cpp:28-34: cleared cut-neighbour arrays, every node NEG
    for (int nt = 0; nt < 4; ++nt) {
      const int nn = ma->nnodes(nt);
      cutnb[nt] = std::make_shared<BitArray>(nn);
      dom[nt].assign(nn, (uint8_t)NEG);
    }
    for (int c = 0; c < 8; ++c) {
      elems[c] = std::make_shared<BitArray>(ma->ne);
      selems[c] = std::make_shared<BitArray>(ma->nse);
    }
    // cutinfo.cpp:22-26: facet arrays are only initialised here, indexed by
    // DOMAIN_TYPE values (NEG = 0 -> all set, POS = 1 and IF = 2 cleared)
    int nf = 0;
    try {
      ma->ensure_facets();
      nf = ma->nf;
    } catch (...) {
      nf = 0;
    }
    for (int c = 0; c < 8; ++c) facets[c] = std::make_shared<BitArray>(nf);
    facets[NEG]->Set();
    facets[POS]->Clear();
    facets[IF]->Clear();
    cut_ratio[VOL].assign(ma->ne, 0.0);
    cut_ratio[BND].assign(ma->nse, 0.0);
  }

  void Update(py::object lset, int subdivlvl, int time_order) {
    if (time_order > 0)
      py::print("Warning: you used CutInfo.Update with time_order > 0. time_order = 0 suffices to detect the "
                "cut topology and should be prefered.");
    if (needs_subdivision(lset, subdivlvl)) {
      UpdateSubdivision(lset, subdivlvl);
      return;
    }
    bool is_p1 = true;
    std::vector<double> v = p1_values(lset, ma->nv, is_p1);
    d_lset.upload(v.data(), v.size());
    for (int vb = 0; vb < 2; ++vb) {
      const int ne = ma->GetNE((VorB)vb);
      auto &arr = vb == VOL ? elems : selems;
      if (ne == 0) {
        for (auto &b : arr) b->Clear();
        continue;
      }
      const int D = vb == VOL ? ma->dim : ma->dim - 1;
      const int nw = (ne + 31) / 32;
      DevBuf<uint32_t> bits((size_t)8 * nw);
      DevBuf<double> cr(ne);
      DevBuf<uint8_t> dt(ne);
      if (D < 1) continue;  // point boundaries of 1D meshes
      const bool quad_el = vb == VOL ? ma->quad : (ma->quad && D == 2);  // hexes have quadrilateral faces
      if (quad_el)
        check(xg_quad_mark(ctx(), D, ne, vb == VOL ? ma->el2v.p : ma->bnd2v.p, d_lset.p, FIND_OPTIMAL, dt.p, cr.p,
                           bits.p, nullptr, nullptr));
      else
        check(xg_mark(ctx(), D, ne, vb == VOL ? ma->el2v.p : ma->bnd2v.p, d_lset.p, dt.p, cr.p, bits.p, nullptr,
                      nullptr));
      if (vb == VOL) UpdateNodes(dt.p, cr.p);
      check(xg_sync(ctx()));
      std::vector<uint32_t> hb((size_t)8 * nw);
      bits.download(hb.data(), hb.size());
      for (int c = 0; c < 8; ++c) {
        auto &w = arr[c]->words();
        std::copy(hb.begin() + (size_t)c * nw, hb.begin() + (size_t)(c + 1) * nw, w.begin());
      }
      cut_ratio[vb].resize(ne);
      cr.download(cut_ratio[vb].data(), ne);
    }
  }

  // cutinfo.cpp:135-224.  The cut-neighbour arrays are not cleared here (the
  // reference clears them only in the constructor): they collect the nodes
  // of the cut elements of every Update so far.
  void UpdateNodes(const uint8_t *d_dt, const double *d_cr) {
    for (int nt = 0; nt < 4; ++nt) {
      const int nn = ma->nnodes(nt);
      if (nn == 0) continue;
      int k = 0;
      const int32_t *tab = ma->el2node(nt, k);
      DevBuf<uint32_t> dc;
      DevBuf<uint8_t> dd(nn);
      dc.upload(cutnb[nt]->words().data(), cutnb[nt]->words().size());
      check(xg_node_domains(ctx(), ma->ne, k, tab, d_dt, d_cr, nn, dc.p, dd.p));
      check(xg_sync(ctx()));
      dc.download(cutnb[nt]->words().data(), cutnb[nt]->words().size());
      dd.download(dom[nt].data(), nn);
    }
  }
  int node_index(int nt) const {
    if (nt == NT_ELEMENT) return ma->dim == 1 ? NT_EDGE : (ma->dim == 2 ? NT_FACE : NT_CELL);
    if (nt == NT_FACET) return ma->dim == 1 ? NT_VERTEX : (ma->dim == 2 ? NT_EDGE : NT_FACE);
    if (nt < 0 || nt > 3) throw std::runtime_error("node type: NT_VERTEX .. NT_FACET");
    return nt;
  }

  // P2 level sets / subdivlvl > 0 (cutinfo.cpp:80-126 with the subdivision
  // rule): element domains from the lattice check of the level-L rule
  // (xg_mark_p2), cut ratios part_vol[NEG] / (part_vol[NEG] + part_vol[POS])
  // from the rule's NEG and POS parts on the cut elements
  // (xg_subdiv_moments_mesh); boundary triangles by xg_mark_p2_trig.
  void UpdateSubdivision(py::object lset, int L) {
    simplices_only(ma, "CutInfo with a P2 level set / subdivlvl > 0");
    if (L > 5) throw std::runtime_error("CutInfo.Update: subdivlvl <= 5 in this build");
    const int ne = ma->ne;
    std::vector<int32_t> all(ne);
    for (int e = 0; e < ne; ++e) all[e] = e;
    std::vector<double> nod = p2_nodal_values(ma, lset, all);
    const int nw = (ne + 31) / 32;
    DevBuf<double> d_nod;
    d_nod.upload(nod.data(), nod.size());
    DevBuf<uint8_t> d_dt(std::max(ne, 1));
    DevBuf<uint32_t> bits((size_t)8 * std::max(nw, 1));
    DevBuf<int32_t> cl(std::max(ne, 1)), nc(1);
    check(xg_mark_p2(ctx(), ne, d_nod.p, L, d_dt.p, bits.p, cl.p, nc.p));
    check(xg_sync(ctx()));
    int32_t ncut = 0;
    nc.download(&ncut, 1);
    std::vector<uint32_t> hb((size_t)8 * nw);
    bits.download(hb.data(), hb.size());
    for (int c = 0; c < 8; ++c) {
      auto &w = elems[c]->words();
      std::copy(hb.begin() + (size_t)c * nw, hb.begin() + (size_t)(c + 1) * nw, w.begin());
    }
    std::vector<uint8_t> dt(ne);
    d_dt.download(dt.data(), ne);
    auto &cr = cut_ratio[VOL];
    cr.assign(ne, 0.0);
    for (int e = 0; e < ne; ++e) cr[e] = dt[e] == NEG ? 1.0 : 0.0;
    if (ncut > 0) {
      std::vector<int32_t> cut(ncut);
      cl.download(cut.data(), ncut);
      // NEG and POS part volumes of the subdivision rule (exact base rules of
      // degree 2: the degree-2 moments of each side; zero level-set values
      // count as POS, xintegration.cpp:1159-1164); volume = sum_alpha
      // (2 / alpha!) m_alpha = the integral of (lam_0 + .. + lam_3)^2
      DevBuf<int32_t> d_cut;
      d_cut.upload(cut.data(), cut.size());
      DevBuf<double> mom((size_t)ncut * 2 * 10);
      const int sides[2] = {NEG, POS};
      check(xg_subdiv_moments_mesh(ctx(), 1, L, 2, ncut, d_cut.p, ma->el2v.p, ma->coords.p, d_nod.p, 2, sides, mom.p,
                                   nullptr));
      check(xg_sync(ctx()));
      std::vector<double> hm((size_t)ncut * 2 * 10);
      mom.download(hm.data(), hm.size());
      static const double mult[10] = {1, 2, 2, 2, 1, 2, 2, 1, 2, 1};
      // The element domain is decided on the part volumes, not on the
      // lattice (cutinfo.cpp:104-124): a lattice-cut element whose leaves
      // all fall to one side (exact zeros, or values below the leaves'
      // relative 1e-14 threshold, xdecompose.hpp:34-85) has an empty rule on
      // the other side and is marked uncut.
      bool moved = false;
      for (int i = 0; i < ncut; ++i) {
        double a = 0.0, b = 0.0;
        for (int k = 0; k < 10; ++k) {
          a += mult[k] * hm[((size_t)i * 2 + 0) * 10 + k];
          b += mult[k] * hm[((size_t)i * 2 + 1) * 10 + k];
        }
        cr[cut[i]] = a / (a + b);
        const uint8_t d = a > 0.0 ? (b > 0.0 ? (uint8_t)IF : (uint8_t)NEG) : (uint8_t)POS;
        if (d != IF) {
          const int e = cut[i];
          dt[e] = d;
          const uint32_t m = 1u << (e & 31);
          for (int c = 0; c < 8; ++c) {
            const bool on = c == 7 || (c & (d == NEG ? 1 : 2));
            uint32_t &w = elems[c]->words()[e >> 5];
            w = on ? (w | m) : (w & ~m);
          }
          moved = true;
        }
      }
      if (moved) d_dt.upload(dt.data(), ne);
    }
    DevBuf<double> d_cr(std::max(ne, 1));
    d_cr.upload(cr.data(), ne);
    if (ne) UpdateNodes(d_dt.p, d_cr.p);
    UpdateSubdivisionBnd(lset, L);
  }

  // Boundary triangles (cutinfo.cpp:80-124, vb = BND): the subdivision rule
  // on each triangle with the level set's restriction (6 nodal values).
  void UpdateSubdivisionBnd(py::object lset, int L) {
    const int nse = ma->nse;
    cut_ratio[BND].assign(nse, 0.0);
    if (nse == 0) {
      for (auto &b : selems) b->Clear();
      return;
    }
    const int order = py::hasattr(lset, "order") ? lset.attr("order").cast<int>() : 1;
    py::object vec = py::hasattr(lset, "vec") ? lset.attr("vec") : lset;
    auto a = py::array_t<double, py::array::c_style | py::array::forcecast>(vec);
    const double *v = a.data();
    std::vector<double> nod((size_t)nse * 6);
    if (order == 2) {
      py::tuple t = ma->pymesh.attr("bnd_dofs")(2);
      auto dofs = py::array_t<int32_t, py::array::c_style | py::array::forcecast>(t[0]);
      const int32_t *d = dofs.data();
      for (size_t i = 0; i < (size_t)nse * 6; ++i) nod[i] = v[d[i]];
    } else {
      static const int ed[3][2] = {{0, 1}, {0, 2}, {1, 2}};
      for (int i = 0; i < nse; ++i) {
        const int32_t *vv = &ma->h_bnd2v[(size_t)i * 3];
        for (int k = 0; k < 3; ++k) nod[(size_t)i * 6 + k] = v[vv[k]];
        for (int k = 0; k < 3; ++k)
          nod[(size_t)i * 6 + 3 + k] = 0.5 * (v[vv[ed[k][0]]] + v[vv[ed[k][1]]]);
      }
    }
    const int nw = (nse + 31) / 32;
    DevBuf<double> d_nod, d_cr(nse);
    d_nod.upload(nod.data(), nod.size());
    DevBuf<uint8_t> d_dt(nse);
    DevBuf<uint32_t> bits((size_t)8 * nw);
    check(xg_mark_p2_trig(ctx(), nse, d_nod.p, L, d_dt.p, d_cr.p, bits.p, nullptr, nullptr));
    check(xg_sync(ctx()));
    std::vector<uint32_t> hb((size_t)8 * nw);
    bits.download(hb.data(), hb.size());
    for (int c = 0; c < 8; ++c) {
      auto &w = selems[c]->words();
      std::copy(hb.begin() + (size_t)c * nw, hb.begin() + (size_t)(c + 1) * nw, w.begin());
    }
    d_cr.download(cut_ratio[BND].data(), nse);
  }

  BA GetElementsOfDomainType(int cdt, VorB vb) const {
    return vb == VOL ? elems.at(cdt) : selems.at(cdt);
  }
  BA GetFacetsOfDomainType(int cdt) const { return facets.at(cdt); }

  // cutinfo.cpp:263-284
  BA GetElementsWithThresholdContribution(DOMAIN_TYPE dt, double threshold, VorB vb) const {
    const auto &cr = cut_ratio[vb];
    auto ret = std::make_shared<BitArray>(cr.size());
    if (dt == POS) threshold = 1 - threshold;
    for (size_t e = 0; e < cr.size(); ++e) {
      if (dt == NEG && cr[e] >= threshold) ret->SetBit(e);
      else if (dt == POS && cr[e] <= threshold) ret->SetBit(e);
    }
    return ret;
  }
};

int parse_cdt(py::object dt) {
  if (py::isinstance<py::int_>(dt)) return dt.cast<int>();
  try {
    return (int)dt.cast<COMBINED_DOMAIN_TYPE>();
  } catch (...) {
  }
  try {
    return to_cdt(dt.cast<DOMAIN_TYPE>());
  } catch (...) {
  }
  throw std::runtime_error(" unknown type for dt ");
}

int parse_dt(py::object dt) {
  try {
    return (int)dt.cast<DOMAIN_TYPE>();
  } catch (...) {
  }
  if (py::isinstance<py::int_>(dt)) return dt.cast<int>();
  throw std::runtime_error("domain_type must be NEG, POS or IF");
}

// ---------------------------------------------------------------- rules
struct DeviceRule {
  int nel = 0;
  DevBuf<int32_t> elems, offsets;
  DevBuf<double> pts;
  int total = 0;
};

void build_rule(MA ma, const double *d_lset, int dt, int order, const std::vector<int32_t> *elems, DeviceRule &r) {
  simplices_only(ma, "straight cut rule on simplices");
  r.nel = elems ? (int)elems->size() : ma->ne;
  if (elems) r.elems.upload(elems->data(), elems->size());
  r.offsets.resize(r.nel + 1);
  check(xg_cut_rule_count(ctx(), ma->dim, r.nel, elems ? r.elems.p : nullptr, ma->el2v.p, d_lset, dt, order,
                          r.offsets.p, nullptr));
  check(xg_sync(ctx()));
  int32_t tot = 0;
  hipcheck(hipMemcpy(&tot, r.offsets.p + r.nel, sizeof(int32_t), hipMemcpyDeviceToHost), "rule total");
  r.total = tot;
  r.pts.resize((size_t)std::max(tot, 1) * 4);
  check(xg_cut_rule_fill(ctx(), ma->dim, r.nel, elems ? r.elems.p : nullptr, ma->el2v.p, ma->coords.p, d_lset, dt,
                         order, r.offsets.p, r.pts.p));
  check(xg_sync(ctx()));
}

// the straight cut rule of a quadrilateral / hexahedral mesh
// (LevelsetCutQuadrilateral, straightcutrule.cpp:206-501)
void build_quad_rule(MA ma, const double *d_lset, int dt, int order, int policy, DeviceRule &r,
                     const std::vector<int32_t> *elems = nullptr) {
  r.nel = elems ? (int)elems->size() : ma->ne;
  if (elems) r.elems.upload(elems->data(), elems->size());
  const int32_t *el = elems ? r.elems.p : nullptr;
  r.offsets.resize(r.nel + 1);
  check(xg_quad_rule_count(ctx(), ma->dim, r.nel, el, ma->el2v.p, d_lset, dt, order, policy, r.offsets.p));
  check(xg_sync(ctx()));
  int32_t tot = 0;
  hipcheck(hipMemcpy(&tot, r.offsets.p + r.nel, sizeof(int32_t), hipMemcpyDeviceToHost), "rule total");
  r.total = tot;
  r.pts.resize((size_t)std::max(tot, 1) * 4);
  check(xg_quad_rule_fill(ctx(), ma->dim, r.nel, el, ma->el2v.p, ma->coords.p, d_lset, dt, order, policy,
                          r.offsets.p, r.pts.p));
  check(xg_sync(ctx()));
}

// ---------------------------------------------------------------- coefficient functions
// A coefficient's postfix program (CoefficientFunction.program(): code,
// constants, P1 grid functions) with the grid functions' vertex values on
// the device; cf points into the vectors here.
struct DeviceCf {
  std::vector<int32_t> code;
  std::vector<double> consts;
  std::vector<std::unique_ptr<DevBuf<double>>> gfbuf;
  std::vector<const double *> gfp;
  xg_cf cf{};
};

// vmap (nullable): the mesh's vertices as vertices of the coefficient's
// grid functions (the boundary submeshes' vertex copies); ne: the number of
// elements a piecewise constant grid function must have (< 0: none allowed)
void make_device_cf(py::object cfobj, int nv, DeviceCf &d, const std::vector<int32_t> *vmap = nullptr,
                    int ne = -1) {
  if (!py::hasattr(cfobj, "program")) throw std::runtime_error("coefficient: a CoefficientFunction expected");
  py::tuple t = cfobj.attr("program")();
  d.code = t[0].cast<std::vector<int32_t>>();
  d.consts = t[1].cast<std::vector<double>>();
  for (auto g : t[2].cast<py::list>()) {
    auto a = py::array_t<double, py::array::c_style | py::array::forcecast>(g.attr("vec"));
    std::vector<double> vals;
    const bool per_element = py::hasattr(g.attr("space"), "is_l2");
    if (per_element) {
      if (ne < 0 || vmap)
        throw std::runtime_error("coefficient: piecewise constant grid functions only on the mesh's elements");
      if ((int)a.size() != ne) throw std::runtime_error("coefficient: L2 order-0 grid functions (one value per element)");
      vals.assign(a.data(), a.data() + a.size());
    } else if (vmap) {
      if ((int)vmap->size() != nv) throw std::runtime_error("coefficient: vertex map size");
      vals.resize(nv);
      for (int k = 0; k < nv; ++k) {
        const int32_t v = (*vmap)[k];
        if (v < 0 || v >= (int)a.size()) throw std::runtime_error("coefficient: P1 grid functions (one value per vertex)");
        vals[k] = a.data()[v];
      }
    } else {
      if ((int)a.size() != nv) throw std::runtime_error("coefficient: P1 grid functions (one value per vertex)");
      vals.assign(a.data(), a.data() + a.size());
    }
    d.gfbuf.emplace_back(new DevBuf<double>());
    d.gfbuf.back()->upload(vals.data(), vals.size());
    d.gfp.push_back(d.gfbuf.back()->p);
  }
  d.cf.ncode = (int32_t)d.code.size();
  d.cf.code = d.code.data();
  d.cf.nconst = (int32_t)d.consts.size();
  d.cf.consts = d.consts.data();
  d.cf.ngf = (int32_t)d.gfp.size();
  d.cf.gf = d.gfp.data();
}

// ---------------------------------------------------------------- several level sets
// A set of P1 level sets on the device (one array of vertex values each) and
// a list of domain tuples (LevelsetIntegrationDomain with several level sets,
// cutint/lsetintdom.cpp:178-301).
struct MlsetDomain {
  int M = 0, K = 0;
  std::vector<std::unique_ptr<DevBuf<double>>> lsets;
  std::vector<const double *> ptrs;
  std::vector<int8_t> dtts;  // K x M
};

bool is_level_set_tuple(py::object o) { return py::isinstance<py::tuple>(o) || py::isinstance<py::list>(o); }

// domain tuples: a tuple of domain types, a list of such tuples, or any
// object with `as_list` (DomainTypeArray)
std::vector<std::vector<int>> parse_dtts(py::object o) {
  if (py::hasattr(o, "as_list")) o = o.attr("as_list");
  std::vector<std::vector<int>> out;
  auto one = [](py::object t) {
    std::vector<int> v;
    for (auto x : t) v.push_back(parse_dt(py::reinterpret_borrow<py::object>(x)));
    return v;
  };
  if (py::isinstance<py::tuple>(o)) {
    out.push_back(one(o));
  } else if (py::isinstance<py::list>(o)) {
    for (auto t : o) out.push_back(one(py::reinterpret_borrow<py::object>(t)));
  } else {
    throw std::runtime_error("domain_type of several level sets: a tuple of domain types or a list of tuples");
  }
  return out;
}

std::vector<std::vector<double>> level_set_values(MA ma, py::object lsets) {
  std::vector<std::vector<double>> out;
  for (auto l : lsets) {
    bool is_p1 = true;
    out.push_back(p1_values(py::reinterpret_borrow<py::object>(l), ma->nv, is_p1));
    if (!is_p1) throw std::runtime_error("several level sets: P1 level sets only");
  }
  if (out.empty()) throw std::runtime_error("several level sets: empty level set tuple");
  return out;
}

void make_mlset(const std::vector<std::vector<double>> &vals, py::object dts, MlsetDomain &md) {
  md.M = (int)vals.size();
  for (const auto &v : vals) {
    md.lsets.emplace_back(new DevBuf<double>());
    md.lsets.back()->upload(v.data(), v.size());
    md.ptrs.push_back(md.lsets.back()->p);
  }
  auto t = parse_dtts(dts);
  md.K = (int)t.size();
  for (auto &tt : t) {
    if ((int)tt.size() != md.M)
      throw std::runtime_error("several level sets: the domain tuples must have one entry per level set");
    for (int v : tt) md.dtts.push_back((int8_t)v);
  }
}

void build_mlset_rule(MA ma, const MlsetDomain &md, int order, const std::vector<int32_t> *elems, DeviceRule &r) {
  simplices_only(ma, "several level sets");
  r.nel = elems ? (int)elems->size() : ma->ne;
  if (elems) r.elems.upload(elems->data(), elems->size());
  r.offsets.resize(r.nel + 1);
  check(xg_mlset_rule_count(ctx(), ma->dim, r.nel, elems ? r.elems.p : nullptr, ma->el2v.p, md.M, md.ptrs.data(),
                            md.K, md.dtts.data(), order, r.offsets.p));
  check(xg_sync(ctx()));
  int32_t tot = 0;
  hipcheck(hipMemcpy(&tot, r.offsets.p + r.nel, sizeof(int32_t), hipMemcpyDeviceToHost), "rule total");
  r.total = tot;
  r.pts.resize((size_t)std::max(tot, 1) * 4);
  check(xg_mlset_rule_fill(ctx(), ma->dim, r.nel, elems ? r.elems.p : nullptr, ma->el2v.p, ma->coords.p, md.M,
                           md.ptrs.data(), md.K, md.dtts.data(), order, r.offsets.p, r.pts.p));
  check(xg_sync(ctx()));
}

// MultiLevelsetCutInformation (xfem/cutinfo.cpp:286-415,
// python_xfem.cpp:600-720).  As in the reference, the level set values are
// copied at construction and by Update only, and the element marks of each
// requested list of domain tuples are kept: asking again for the same list
// (as a multiset of tuples) returns the same BitArray, and Update recomputes
// every kept array in place.
class MultiLevelsetCutInformation {
 public:
  MA ma;
  std::vector<std::vector<double>> vals;
  struct Kept {
    BA bits;
    std::vector<std::vector<int>> dtts;  // sorted
    bool contrib;
    py::object dts;
  };
  std::vector<Kept> kept;

  MultiLevelsetCutInformation(MA m, py::object l) : ma(m) {
    simplices_only(ma, "MultiLevelsetCutInfo");
    check_lsets(l);
    vals = level_set_values(ma, l);
  }
  static void check_lsets(py::object l) {
    if (!is_level_set_tuple(l)) throw std::runtime_error("MultiLevelsetCutInfo: a tuple of level sets");
    for (auto x : l)
      if (!py::hasattr(x, "vec"))
        throw std::runtime_error("MultiLevelsetCutInfo: the level sets must be (P1) grid functions");
  }
  void compute(py::object dts, bool contrib, BitArray &out) const {
    MlsetDomain md;
    make_mlset(vals, dts, md);
    if (ma->ne == 0) return;
    DevBuf<uint32_t> d(out.words().size());
    check(xg_mlset_mark(ctx(), ma->dim, ma->ne, ma->el2v.p, ma->coords.p, md.M, md.ptrs.data(), md.K,
                        md.dtts.data(), contrib ? nullptr : d.p, contrib ? d.p : nullptr));
    check(xg_sync(ctx()));
    d.download(out.words().data(), out.words().size());
  }
  void Update(py::object l) {
    check_lsets(l);
    if (py::len(l) != vals.size()) throw std::runtime_error("MultiLevelsetCutInfo.Update: number of level sets");
    vals = level_set_values(ma, l);
    for (auto &k : kept) compute(k.dts, k.contrib, *k.bits);
  }
  BA mark(py::object dts, bool contrib, VorB vb) {
    if (vb != VOL) throw std::runtime_error("MultiLevelsetCutInfo: VOL elements in this build");
    auto key = parse_dtts(dts);
    for (auto &t : key)
      if (t.size() != vals.size())
        throw std::runtime_error("MultiLevelsetCutInfo: the domain tuples must have one entry per level set");
    std::sort(key.begin(), key.end());
    for (auto &k : kept)
      if (k.contrib == contrib && k.dtts == key) return k.bits;
    auto ret = std::make_shared<BitArray>(ma->ne);
    compute(dts, contrib, *ret);
    py::list copy;
    py::object src = py::hasattr(dts, "as_list") ? dts.attr("as_list") : dts;
    if (py::isinstance<py::list>(src)) {
      for (auto t : src) copy.append(t);
      kept.push_back({ret, key, contrib, copy});
    } else {
      kept.push_back({ret, key, contrib, src});
    }
    return ret;
  }
};

// polynomial coefficient function -> (coef, exponent triples); false for a
// general coefficient function
bool poly_terms_of(py::object cf, std::vector<double> &coef, std::vector<int32_t> &ex) {
  const bool poly = cf.is_none() || py::isinstance<py::float_>(cf) || py::isinstance<py::int_>(cf) ||
                    (py::hasattr(cf, "poly_terms") && (!py::hasattr(cf, "is_poly") || cf.attr("is_poly")().cast<bool>()));
  if (!poly) return false;
  coef.clear();
  ex.clear();
  if (cf.is_none()) {
    coef = {0.0};
    ex = {0, 0, 0};
  } else if (py::isinstance<py::float_>(cf) || py::isinstance<py::int_>(cf)) {
    coef = {cf.cast<double>()};
    ex = {0, 0, 0};
  } else {
    for (auto t : cf.attr("poly_terms")()) {
      py::tuple tt = t.cast<py::tuple>();
      coef.push_back(tt[0].cast<double>());
      py::tuple e = tt[1].cast<py::tuple>();
      for (int d = 0; d < 3; ++d) ex.push_back(d < (int)e.size() ? e[d].cast<int>() : 0);
    }
  }
  if (coef.size() > 32) throw std::runtime_error("IntegrateX: polynomial with more than 32 terms");
  return true;
}

// IntegrateX on the subdivision rule (P2 level sets, subdivlvl > 0): the rule
// of each element contracted to weights at the 35 nodes of the degree-4
// Lagrange lattice (xg_subdiv_weights_mesh), which reproduce the rule's point
// sum for every polynomial of degree <= 4; below order 4 (inexact base rules)
// xg_subdiv_weights with the POS part as the NEG part of -phi
py::object IntegrateSubdivision(MA ma, py::object lset, int dt, int L, int order, py::object cf, bool element_wise) {
  simplices_only(ma, "IntegrateX with a P2 level set / subdivlvl > 0");
  std::vector<double> coef;
  std::vector<int32_t> ex;
  if (!poly_terms_of(cf, coef, ex))
    throw std::runtime_error("IntegrateX with a P2 level set / subdivlvl > 0: polynomial integrands only");
  for (size_t t = 0; t < coef.size(); ++t)
    if (ex[3 * t] + ex[3 * t + 1] + ex[3 * t + 2] > 4)
      throw std::runtime_error("IntegrateX with a P2 level set / subdivlvl > 0: polynomials of degree <= 4");
  if (L > 5) throw std::runtime_error("IntegrateX: subdivlvl <= 5 in this build");
  const int ne = ma->ne;
  std::vector<int32_t> all(ne);
  for (int e = 0; e < ne; ++e) all[e] = e;
  std::vector<double> nod = p2_nodal_values(ma, lset, all);
  const bool exact = order >= 4;  // exact base rules: the mesh entry, POS kept as such
  if (dt == POS && !exact)
    for (auto &v : nod) v = -v;  // lower orders: the POS part as the NEG part of -phi
  DevBuf<double> d_nod, d_fi, W((size_t)std::max(ne, 1) * 35), d_pe(std::max(ne, 1)), d_tot(1);
  d_nod.upload(nod.data(), nod.size());
  if (ne && exact) {
    DevBuf<int32_t> d_all;
    d_all.upload(all.data(), all.size());
    const int dts[1] = {dt};
    check(xg_subdiv_weights_mesh(ctx(), L, order, ne, d_all.p, ma->el2v.p, ma->coords.p, d_nod.p, 1, dts, W.p,
                                 nullptr));
  } else if (ne) {
    if (dt == IF) {
      std::vector<double> fi = tet_finv(ma, all);
      d_fi.upload(fi.data(), fi.size());
    }
    check(xg_subdiv_weights(ctx(), ne, d_nod.p, 2, L, order, 4, dt == IF ? d_fi.p : nullptr,
                            dt == IF ? nullptr : W.p, dt == IF ? W.p : nullptr, nullptr));
  }
  check(xg_integrate_nodes(ctx(), ne, nullptr, ma->el2v.p, ma->coords.p, 4, W.p, (int)coef.size(), coef.data(),
                           ex.data(), d_pe.p, d_tot.p));
  check(xg_sync(ctx()));
  std::vector<double> pe(ne);
  double total = 0.0;
  d_pe.download(pe.data(), ne);
  d_tot.download(&total, 1);
  if (element_wise) return py::array_t<double>((py::ssize_t)pe.size(), pe.data());
  return py::float_(total);
}

// element matrices on the subdivision rule (dCut with a P2 level set or
// subdivlvl > 0): with exact base rules (order >= 2p) the fused mesh entry
// xg_subdiv_pipeline; below, the degree-2p moments of the rule
// (xg_subdiv_moments) on the listed tetrahedra, then the moment-form
// identities (xg_elmat_moments)
py::tuple SubdivElementMatrices(py::object mesh, py::object lset, int dt, int L, int order, int p, int mask,
                                py::object elements, int precision) {
  MA ma = as_mesh(mesh);
  simplices_only(ma, "SubdivElementMatrices");
  if (p < 1 || p > 3) throw std::runtime_error("SubdivElementMatrices: p = 1, 2, 3");
  if (L > 5) throw std::runtime_error("SubdivElementMatrices: subdivlvl <= 5 in this build");
  std::vector<int32_t> el;
  if (elements.is_none()) {
    el.resize(ma->ne);
    for (int e = 0; e < ma->ne; ++e) el[e] = e;
  } else if (py::isinstance<BitArray>(elements)) {
    const BitArray &b = elements.cast<const BitArray &>();
    for (size_t e = 0; e < b.Size(); ++e)
      if (b.Test(e)) el.push_back((int32_t)e);
  } else {
    auto a = py::array_t<int32_t, py::array::c_style | py::array::forcecast>(elements);
    el.assign(a.data(), a.data() + a.size());
  }
  const int nel = (int)el.size();
  const int nm = (2 * p + 3) * (2 * p + 2) * (2 * p + 1) / 6;
  const int nd = (p + 1) * (p + 2) * (p + 3) / 6;
  if (order >= 2 * p) {
    // exact base rules: the fused mesh entry (nodal values of all elements,
    // indexed by element id; the POS side keeps zero values as the
    // reference's leaves do)
    std::vector<int32_t> all(ma->ne);
    for (int e = 0; e < ma->ne; ++e) all[e] = e;
    std::vector<double> nod_all = p2_nodal_values(ma, lset, all);
    DevBuf<double> d_all;
    d_all.upload(nod_all.data(), nod_all.size());
    DevBuf<int32_t> d_el;
    d_el.upload(el.data(), el.size());
    py::array_t<int32_t> elarr((py::ssize_t)nel, el.data());
    const int dts[1] = {dt};
    const size_t cnt = (size_t)nel * nd * nd;
    if (precision == XG_PREC_FP64) {
      DevBuf<double> o(std::max(cnt, (size_t)1));
      if (nel) check(xg_subdiv_pipeline(ctx(), p, L, order, nel, d_el.p, ma->el2v.p, ma->coords.p, d_all.p, 1, dts,
                                        mask, precision, o.p, nullptr));
      check(xg_sync(ctx()));
      py::array_t<double> h({(py::ssize_t)nel, (py::ssize_t)nd, (py::ssize_t)nd});
      o.download(h.mutable_data(), cnt);
      return py::make_tuple(elarr, h);
    }
    DevBuf<float> o(std::max(cnt, (size_t)1));
    if (nel) check(xg_subdiv_pipeline(ctx(), p, L, order, nel, d_el.p, ma->el2v.p, ma->coords.p, d_all.p, 1, dts,
                                      mask, XG_PREC_FP32, o.p, nullptr));
    check(xg_sync(ctx()));
    py::array_t<float> h({(py::ssize_t)nel, (py::ssize_t)nd, (py::ssize_t)nd});
    o.download(h.mutable_data(), cnt);
    return py::make_tuple(elarr, h);
  }
  std::vector<double> nod = p2_nodal_values(ma, lset, el);
  if (dt == POS)
    for (auto &v : nod) v = -v;  // inexact base rules: the POS part as the NEG part of -phi
  std::vector<double> fi = tet_finv(ma, el);
  DevBuf<double> d_nod, d_fi, d_m((size_t)std::max(nel, 1) * nm);
  DevBuf<int32_t> d_el;
  d_nod.upload(nod.data(), nod.size());
  d_fi.upload(fi.data(), fi.size());
  d_el.upload(el.data(), el.size());
  py::array_t<int32_t> elarr((py::ssize_t)nel, el.data());
  const bool f64 = precision == XG_PREC_FP64;
  DevBuf<double> o64;
  DevBuf<float> o32;
  if (f64) o64.resize((size_t)std::max(nel, 1) * nd * nd);
  else o32.resize((size_t)std::max(nel, 1) * nd * nd);
  if (nel) {
    check(xg_subdiv_moments(ctx(), nel, d_nod.p, 2, L, order, p, d_fi.p, dt == IF ? nullptr : d_m.p,
                            dt == IF ? d_m.p : nullptr, nullptr));
    check(xg_elmat_moments(ctx(), 3, p, nel, d_el.p, ma->el2v.p, ma->coords.p, d_m.p, mask, precision,
                           f64 ? (void *)o64.p : (void *)o32.p));
  }
  check(xg_sync(ctx()));
  if (f64) {
    py::array_t<double> h({(py::ssize_t)nel, (py::ssize_t)nd, (py::ssize_t)nd});
    o64.download(h.mutable_data(), (size_t)nel * nd * nd);
    return py::make_tuple(elarr, h);
  }
  py::array_t<float> h({(py::ssize_t)nel, (py::ssize_t)nd, (py::ssize_t)nd});
  o32.download(h.mutable_data(), (size_t)nel * nd * nd);
  return py::make_tuple(elarr, h);
}

// IntegrateX (python_cutint.cpp:21-225) for straight-cut (P1) level sets.
// cf: an object with poly_terms() -> [(coef, (a, b, c)), ...] runs on the GPU
// reduction kernel; any other cf must provide evaluate(xyz ndarray) -> values
// and is evaluated on the host at the GPU-built rule's physical points.
py::object IntegrateX(py::dict lsetdom, py::object mesh, py::object cf, py::object deformation,
                      py::object ip_container, bool element_wise, int heapsize) {
  if (!deformation.is_none()) throw std::runtime_error("IntegrateX: deformation is not supported in this build");
  MA ma = as_mesh(mesh);
  py::object lset = lsetdom["levelset"];
  const int order = lsetdom.contains("order") ? lsetdom["order"].cast<int>() : 5;
  const int subdiv = lsetdom.contains("subdivlvl") ? lsetdom["subdivlvl"].cast<int>() : 0;
  DeviceRule r;
  if (is_level_set_tuple(lset)) {
    // several level sets (mlsetintegration.cpp)
    if (subdiv != 0) throw std::runtime_error("IntegrateX: several level sets need subdivlvl = 0");
    MlsetDomain md;
    make_mlset(level_set_values(ma, lset), lsetdom["domain_type"], md);
    build_mlset_rule(ma, md, order, nullptr, r);
  } else if (needs_subdivision(lset, subdiv)) {
    if (!ip_container.is_none())
      throw std::runtime_error("IntegrateX: ip_container with subdivision rules is not supported");
    return IntegrateSubdivision(ma, lset, parse_dt(lsetdom["domain_type"]), subdiv, order, cf, element_wise);
  } else {
    const int dt = parse_dt(lsetdom["domain_type"]);
    bool is_p1 = true;
    std::vector<double> v = p1_values(lset, ma->nv, is_p1);
    DevBuf<double> d_l;
    d_l.upload(v.data(), v.size());
    if (ma->quad)
      build_quad_rule(ma, d_l.p, dt, order, parse_policy(lsetdom), r);
    else
      build_rule(ma, d_l.p, dt, order, nullptr, r);
  }
  const int D = ma->dim;
  std::vector<double> pe(ma->ne, 0.0);
  double total = 0.0;
  std::vector<double> coef;
  std::vector<int32_t> ex;
  const bool poly = poly_terms_of(cf, coef, ex);
  if (ma->quad && (!poly || !ip_container.is_none()))
    throw std::runtime_error("IntegrateX on quadrilaterals / hexahedra: polynomial integrands, no ip_container");
  if (ma->quad) {
    DevBuf<double> d_pe(ma->ne), d_tot(1);
    check(xg_quad_integrate(ctx(), D, ma->ne, nullptr, ma->el2v.p, ma->coords.p, r.offsets.p, r.pts.p,
                            (int)coef.size(), coef.data(), ex.data(), d_pe.p, d_tot.p));
    check(xg_sync(ctx()));
    d_pe.download(pe.data(), ma->ne);
    d_tot.download(&total, 1);
  } else if (poly) {
    DevBuf<double> d_pe(ma->ne), d_tot(1);
    check(xg_integrate(ctx(), D, ma->ne, nullptr, ma->el2v.p, ma->coords.p, r.offsets.p, r.pts.p, (int)coef.size(),
                       coef.data(), ex.data(), d_pe.p, d_tot.p));
    check(xg_sync(ctx()));
    d_pe.download(pe.data(), ma->ne);
    d_tot.download(&total, 1);
  } else {
    // any other coefficient function: its postfix program scales the rule's
    // weights on the device, then the integral of f = 1
    if (!ip_container.is_none()) {
      // the rule's physical points and weights as the reference lists them
      std::vector<int32_t> off(ma->ne + 1);
      r.offsets.download(off.data(), off.size());
      std::vector<double> p((size_t)r.total * 4);
      r.pts.download(p.data(), p.size());
      auto c = py::array_t<double, py::array::c_style | py::array::forcecast>(mesh.attr("coords"));
      const double *X = c.data();
      py::list l = ip_container.cast<py::list>();
      for (int e = 0; e < ma->ne; ++e) {
        const int32_t *vv = &ma->h_el2v[(size_t)e * (D + 1)];
        double F[9], det;
        for (int rr = 0; rr < D; ++rr)
          for (int cc = 0; cc < D; ++cc) F[rr * D + cc] = X[(size_t)vv[cc] * D + rr] - X[(size_t)vv[D] * D + rr];
        if (D == 1) det = F[0];
        else if (D == 2) det = F[0] * F[3] - F[1] * F[2];
        else
          det = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
                F[2] * (F[3] * F[7] - F[4] * F[6]);
        for (int q = off[e]; q < off[e + 1]; ++q) {
          double xq[3] = {0.0, 0.0, 0.0};
          for (int rr = 0; rr < D; ++rr) {
            double a = X[(size_t)vv[D] * D + rr];
            for (int cc = 0; cc < D; ++cc) a += F[rr * D + cc] * p[(size_t)4 * q + cc];
            xq[rr] = a;
          }
          l.append(py::make_tuple(xq[0], xq[1], xq[2], std::abs(det) * p[(size_t)4 * q + 3]));
        }
      }
    }
    DeviceCf dcf;
    make_device_cf(cf, ma->nv, dcf, nullptr, ma->ne);
    check(xg_rule_scale_cf(ctx(), D, D, ma->ne, nullptr, ma->el2v.p, ma->coords.p, r.offsets.p, r.pts.p, &dcf.cf));
    const double one = 1.0;
    const int32_t ex0[3] = {0, 0, 0};
    DevBuf<double> d_pe(ma->ne), d_tot(1);
    check(xg_integrate(ctx(), D, ma->ne, nullptr, ma->el2v.p, ma->coords.p, r.offsets.p, r.pts.p, 1, &one, ex0,
                       d_pe.p, d_tot.p));
    check(xg_sync(ctx()));
    d_pe.download(pe.data(), ma->ne);
    d_tot.download(&total, 1);
  }
  if (element_wise) return py::array_t<double>((py::ssize_t)pe.size(), pe.data());
  return py::float_(total);
}

// element matrices of (mass | laplace) on the cut rule of dt for the listed
// elements (all if None): the kernel behind SymbolicCutBFI / dCut forms
py::tuple CutElementMatrices(py::object mesh, py::object lset, int dt, int order, int p, int mask,
                             py::object elements, int precision) {
  MA ma = as_mesh(mesh);
  simplices_only(ma, "CutElementMatrices");
  bool is_p1 = true;
  std::vector<double> v = p1_values(lset, ma->nv, is_p1);
  if (!is_p1) throw std::runtime_error("CutElementMatrices: P1 level sets only");
  DevBuf<double> d_l;
  d_l.upload(v.data(), v.size());
  std::vector<int32_t> el;
  if (elements.is_none()) {
    el.resize(ma->ne);
    for (int e = 0; e < ma->ne; ++e) el[e] = e;
  } else if (py::isinstance<BitArray>(elements)) {
    const BitArray &b = elements.cast<const BitArray &>();
    for (size_t e = 0; e < b.Size(); ++e)
      if (b.Test(e)) el.push_back((int32_t)e);
  } else {
    auto a = py::array_t<int32_t, py::array::c_style | py::array::forcecast>(elements);
    el.assign(a.data(), a.data() + a.size());
  }
  const int D = ma->dim;
  int nd = 1;
  for (int k = 1; k <= D; ++k) nd = nd * (p + k) / k;
  const int nel = (int)el.size();
  DevBuf<int32_t> d_el;
  d_el.upload(el.data(), el.size());
  const int dts[1] = {dt};
  py::array_t<int32_t> elarr((py::ssize_t)nel, el.data());
  if (precision == XG_PREC_FP64) {
    DevBuf<double> out((size_t)nel * nd * nd);
    if (nel) check(xg_cut_pipeline(ctx(), D, p, order, nel, d_el.p, ma->el2v.p, ma->coords.p, d_l.p, 1, dts, mask,
                                   precision, out.p));
    check(xg_sync(ctx()));
    py::array_t<double> h({(py::ssize_t)nel, (py::ssize_t)nd, (py::ssize_t)nd});
    out.download(h.mutable_data(), (size_t)nel * nd * nd);
    return py::make_tuple(elarr, h);
  }
  DevBuf<float> out((size_t)nel * nd * nd);
  if (nel) check(xg_cut_pipeline(ctx(), D, p, order, nel, d_el.p, ma->el2v.p, ma->coords.p, d_l.p, 1, dts, mask,
                                 precision, out.p));
  check(xg_sync(ctx()));
  py::array_t<float> h({(py::ssize_t)nel, (py::ssize_t)nd, (py::ssize_t)nd});
  out.download(h.mutable_data(), (size_t)nel * nd * nd);
  return py::make_tuple(elarr, h);
}

// element matrices on the rule of several level sets (the dCut forms of a
// LevelsetIntegrationDomain with several level sets): GPU mlset rule, then
// the element-matrix kernel on the materialised rule
std::vector<int32_t> element_list(py::object elements, int ne);

// element matrices of the Q1 space on a quad / hex mesh: the straight cut
// rule of the elements (xg_quad_rule_*), then mass / stiffness on it
// (xg_quad_elmat)
py::tuple QuadElementMatrices(py::object mesh, py::object lset, int dt, int order, int policy, int mask,
                              py::object elements, int precision) {
  MA ma = as_mesh(mesh);
  if (!ma->quad) throw std::runtime_error("QuadElementMatrices: a quadrilateral / hexahedral mesh");
  bool is_p1 = true;
  std::vector<double> v = p1_values(lset, ma->nv, is_p1);
  if (!is_p1) throw std::runtime_error("QuadElementMatrices: P1 (Q1) level sets only");
  DevBuf<double> d_l;
  d_l.upload(v.data(), v.size());
  std::vector<int32_t> el = element_list(elements, ma->ne);
  DeviceRule r;
  build_quad_rule(ma, d_l.p, dt, order, policy, r, &el);
  const int nel = (int)el.size(), nd = ma->nvel;
  py::array_t<int32_t> elarr((py::ssize_t)nel, el.data());
  if (precision == XG_PREC_FP64) {
    DevBuf<double> out((size_t)std::max(nel, 1) * nd * nd);
    check(xg_quad_elmat(ctx(), ma->dim, nel, r.elems.p, ma->el2v.p, ma->coords.p, r.offsets.p, r.pts.p, mask,
                        precision, out.p));
    check(xg_sync(ctx()));
    py::array_t<double> h({(py::ssize_t)nel, (py::ssize_t)nd, (py::ssize_t)nd});
    out.download(h.mutable_data(), (size_t)nel * nd * nd);
    return py::make_tuple(elarr, h);
  }
  DevBuf<float> out((size_t)std::max(nel, 1) * nd * nd);
  check(xg_quad_elmat(ctx(), ma->dim, nel, r.elems.p, ma->el2v.p, ma->coords.p, r.offsets.p, r.pts.p, mask,
                      XG_PREC_FP32, out.p));
  check(xg_sync(ctx()));
  py::array_t<float> h({(py::ssize_t)nel, (py::ssize_t)nd, (py::ssize_t)nd});
  out.download(h.mutable_data(), (size_t)nel * nd * nd);
  return py::make_tuple(elarr, h);
}

py::tuple MlsetElementMatrices(py::object mesh, py::object lsets, py::object dts, int order, int p, int mask,
                               py::object elements, int precision, py::object cf) {
  MA ma = as_mesh(mesh);
  simplices_only(ma, "MlsetElementMatrices");
  MlsetDomain md;
  make_mlset(level_set_values(ma, lsets), dts, md);
  std::vector<int32_t> el;
  if (elements.is_none()) {
    el.resize(ma->ne);
    for (int e = 0; e < ma->ne; ++e) el[e] = e;
  } else if (py::isinstance<BitArray>(elements)) {
    const BitArray &b = elements.cast<const BitArray &>();
    for (size_t e = 0; e < b.Size(); ++e)
      if (b.Test(e)) el.push_back((int32_t)e);
  } else {
    auto a = py::array_t<int32_t, py::array::c_style | py::array::forcecast>(elements);
    el.assign(a.data(), a.data() + a.size());
  }
  const int D = ma->dim;
  int nd = 1;
  for (int k = 1; k <= D; ++k) nd = nd * (p + k) / k;
  const int nel = (int)el.size();
  py::array_t<int32_t> elarr((py::ssize_t)nel, el.data());
  DeviceRule r;
  build_mlset_rule(ma, md, order, &el, r);
  if (!cf.is_none() && nel) {
    // a coefficient function: the rule's weights scaled by it
    DeviceCf dcf;
    make_device_cf(cf, ma->nv, dcf, nullptr, ma->ne);
    check(xg_rule_scale_cf(ctx(), D, D, nel, r.elems.p, ma->el2v.p, ma->coords.p, r.offsets.p, r.pts.p, &dcf.cf));
  }
  const bool f64 = precision == XG_PREC_FP64;
  DevBuf<double> o64;
  DevBuf<float> o32;
  if (f64) o64.resize((size_t)std::max(nel, 1) * nd * nd);
  else o32.resize((size_t)std::max(nel, 1) * nd * nd);
  if (nel)
    check(xg_elmat(ctx(), D, p, nel, r.elems.p, ma->el2v.p, ma->coords.p, r.offsets.p, r.pts.p, mask, precision,
                   f64 ? (void *)o64.p : (void *)o32.p));
  check(xg_sync(ctx()));
  if (f64) {
    py::array_t<double> h({(py::ssize_t)nel, (py::ssize_t)nd, (py::ssize_t)nd});
    o64.download(h.mutable_data(), (size_t)nel * nd * nd);
    return py::make_tuple(elarr, h);
  }
  py::array_t<float> h({(py::ssize_t)nel, (py::ssize_t)nd, (py::ssize_t)nd});
  o32.download(h.mutable_data(), (size_t)nel * nd * nd);
  return py::make_tuple(elarr, h);
}

// element list argument: None (all), BitArray or integer array
std::vector<int32_t> element_list(py::object elements, int ne) {
  std::vector<int32_t> el;
  if (elements.is_none()) {
    el.resize(ne);
    for (int e = 0; e < ne; ++e) el[e] = e;
  } else if (py::isinstance<BitArray>(elements)) {
    const BitArray &b = elements.cast<const BitArray &>();
    for (size_t e = 0; e < b.Size(); ++e)
      if (b.Test(e)) el.push_back((int32_t)e);
  } else {
    auto a = py::array_t<int32_t, py::array::c_style | py::array::forcecast>(elements);
    el.assign(a.data(), a.data() + a.size());
  }
  return el;
}

// SymbolicCutLinearFormIntegrator::T_CalcElementVector (xfem/symboliccutlfi.cpp:
// 57-140) for f = 1 on the straight cut rule (moment form, FP64): returns
// (elements, nel x ndof element vectors)
py::tuple CutElementVectors(py::object mesh, py::object lset, int dt, int order, int p, py::object elements) {
  MA ma = as_mesh(mesh);
  if (ma->quad) {
    // the Q1 space on the quad / hex cut rule (FIND_OPTIMAL)
    if (p != 1) throw std::runtime_error("CutElementVectors on quads / hexes: order 1");
    bool is_p1q = true;
    std::vector<double> vq = p1_values(lset, ma->nv, is_p1q);
    if (!is_p1q) throw std::runtime_error("CutElementVectors: P1 level sets only");
    DevBuf<double> d_lq;
    d_lq.upload(vq.data(), vq.size());
    std::vector<int32_t> elq = element_list(elements, ma->ne);
    DeviceRule r;
    build_quad_rule(ma, d_lq.p, dt, order, FIND_OPTIMAL, r, &elq);
    const int nel = (int)elq.size(), nd = ma->nvel;
    DevBuf<double> out((size_t)std::max(nel, 1) * nd);
    check(xg_quad_elvec(ctx(), ma->dim, nel, r.elems.p, ma->el2v.p, ma->coords.p, r.offsets.p, r.pts.p, out.p));
    check(xg_sync(ctx()));
    py::array_t<double> h({(py::ssize_t)nel, (py::ssize_t)nd});
    out.download(h.mutable_data(), (size_t)nel * nd);
    return py::make_tuple(py::array_t<int32_t>((py::ssize_t)nel, elq.data()), h);
  }
  simplices_only(ma, "CutElementVectors");
  bool is_p1 = true;
  std::vector<double> v = p1_values(lset, ma->nv, is_p1);
  if (!is_p1) throw std::runtime_error("CutElementVectors: P1 level sets only");
  DevBuf<double> d_l;
  d_l.upload(v.data(), v.size());
  std::vector<int32_t> el = element_list(elements, ma->ne);
  const int D = ma->dim;
  int nd = 1;
  for (int k = 1; k <= D; ++k) nd = nd * (p + k) / k;
  const int nel = (int)el.size();
  DevBuf<int32_t> d_el;
  d_el.upload(el.data(), el.size());
  const int dts[1] = {dt};
  DevBuf<double> out((size_t)nel * nd);
  if (nel)
    check(xg_cut_elvec(ctx(), D, p, order, nel, d_el.p, ma->el2v.p, ma->coords.p, d_l.p, 1, dts, XG_PREC_FP64, out.p));
  check(xg_sync(ctx()));
  py::array_t<double> h({(py::ssize_t)nel, (py::ssize_t)nd});
  out.download(h.mutable_data(), (size_t)nel * nd);
  py::array_t<int32_t> elarr((py::ssize_t)nel, el.data());
  return py::make_tuple(elarr, h);
}

// element matrices of c u v (mask 1) / c grad u . grad v (mask 2) with a
// coefficient function c: the straight cut rule of the elements, its weights
// scaled by c at the physical points (xg_rule_scale_cf), then the element
// matrices of the scaled rule (xg_elmat)
// ambient (nullable): the vertices' coordinates where the coefficient is
// evaluated (nv x A, A = D + 1 for boundary submeshes), vmap (nullable): the
// grid functions' vertex of each vertex
struct AmbientCoords {
  int A = 0;
  DevBuf<double> buf;
  std::vector<int32_t> vmap;
  bool has_vmap = false;
};

void make_ambient(MA ma, py::object ambient, py::object vmap, AmbientCoords &a) {
  a.A = ma->dim;
  if (!ambient.is_none()) {
    auto x = py::array_t<double, py::array::c_style | py::array::forcecast>(ambient);
    if (x.ndim() != 2 || x.shape(0) != ma->nv || (x.shape(1) != ma->dim && x.shape(1) != ma->dim + 1))
      throw std::runtime_error("coefficient coordinates: nv x D or nv x (D + 1)");
    a.A = (int)x.shape(1);
    a.buf.upload(x.data(), (size_t)x.size());
  }
  if (!vmap.is_none()) {
    auto m = py::array_t<int32_t, py::array::c_style | py::array::forcecast>(vmap);
    a.vmap.assign(m.data(), m.data() + m.size());
    a.has_vmap = true;
  }
}

py::tuple CutElementMatricesCF(py::object mesh, py::object lset, int dt, int order, int p, int mask,
                               py::object elements, int precision, py::object cf, py::object ambient,
                               py::object vmap) {
  MA ma = as_mesh(mesh);
  simplices_only(ma, "CutElementMatricesCF");
  bool is_p1 = true;
  std::vector<double> v = p1_values(lset, ma->nv, is_p1);
  if (!is_p1) throw std::runtime_error("coefficient functions: straight cut rules of P1 level sets");
  DevBuf<double> d_l;
  d_l.upload(v.data(), v.size());
  std::vector<int32_t> el = element_list(elements, ma->ne);
  DeviceRule r;
  build_rule(ma, d_l.p, dt, order, &el, r);
  AmbientCoords amb;
  make_ambient(ma, ambient, vmap, amb);
  DeviceCf dcf;
  make_device_cf(cf, ma->nv, dcf, amb.has_vmap ? &amb.vmap : nullptr, ma->ne);
  const int D = ma->dim, nel = (int)el.size();
  int nd = 1;
  for (int k = 1; k <= D; ++k) nd = nd * (p + k) / k;
  py::array_t<int32_t> elarr((py::ssize_t)nel, el.data());
  if (nel)
    check(xg_rule_scale_cf(ctx(), D, amb.A, nel, r.elems.p, ma->el2v.p, amb.buf.p ? amb.buf.p : ma->coords.p,
                           r.offsets.p, r.pts.p, &dcf.cf));
  if (precision == XG_PREC_FP64) {
    DevBuf<double> out((size_t)std::max(nel, 1) * nd * nd);
    if (nel) check(xg_elmat(ctx(), D, p, nel, r.elems.p, ma->el2v.p, ma->coords.p, r.offsets.p, r.pts.p, mask,
                            precision, out.p));
    check(xg_sync(ctx()));
    py::array_t<double> h({(py::ssize_t)nel, (py::ssize_t)nd, (py::ssize_t)nd});
    out.download(h.mutable_data(), (size_t)nel * nd * nd);
    return py::make_tuple(elarr, h);
  }
  DevBuf<float> out((size_t)std::max(nel, 1) * nd * nd);
  if (nel) check(xg_elmat(ctx(), D, p, nel, r.elems.p, ma->el2v.p, ma->coords.p, r.offsets.p, r.pts.p, mask,
                          precision, out.p));
  check(xg_sync(ctx()));
  py::array_t<float> h({(py::ssize_t)nel, (py::ssize_t)nd, (py::ssize_t)nd});
  out.download(h.mutable_data(), (size_t)nel * nd * nd);
  return py::make_tuple(elarr, h);
}

// element vectors of f v with a coefficient function f (FP64)
py::tuple CutElementVectorsCF(py::object mesh, py::object lset, int dt, int order, int p, py::object elements,
                              py::object cf, py::object ambient, py::object vmap) {
  MA ma = as_mesh(mesh);
  simplices_only(ma, "CutElementVectorsCF");
  bool is_p1 = true;
  std::vector<double> v = p1_values(lset, ma->nv, is_p1);
  if (!is_p1) throw std::runtime_error("coefficient functions: straight cut rules of P1 level sets");
  DevBuf<double> d_l;
  d_l.upload(v.data(), v.size());
  std::vector<int32_t> el = element_list(elements, ma->ne);
  DeviceRule r;
  build_rule(ma, d_l.p, dt, order, &el, r);
  AmbientCoords amb;
  make_ambient(ma, ambient, vmap, amb);
  DeviceCf dcf;
  make_device_cf(cf, ma->nv, dcf, amb.has_vmap ? &amb.vmap : nullptr, ma->ne);
  const int D = ma->dim, nel = (int)el.size();
  int nd = 1;
  for (int k = 1; k <= D; ++k) nd = nd * (p + k) / k;
  DevBuf<double> out((size_t)std::max(nel, 1) * nd);
  if (nel) {
    check(xg_rule_scale_cf(ctx(), D, amb.A, nel, r.elems.p, ma->el2v.p, amb.buf.p ? amb.buf.p : ma->coords.p,
                           r.offsets.p, r.pts.p, &dcf.cf));
    check(xg_elvec(ctx(), D, p, nel, r.elems.p, ma->el2v.p, ma->coords.p, r.offsets.p, r.pts.p, out.p));
  }
  check(xg_sync(ctx()));
  py::array_t<double> h({(py::ssize_t)nel, (py::ssize_t)nd});
  out.download(h.mutable_data(), (size_t)nel * nd);
  return py::make_tuple(py::array_t<int32_t>((py::ssize_t)nel, el.data()), h);
}

// integral of cf over the straight cut rule of a P1 level set on every
// element (per element when element_wise), cf evaluated on the GPU at the
// points given by ambient (nullable, see CutElementMatricesCF)
py::object IntegrateCF(py::object mesh, py::object lset, int dt, int order, py::object cf, bool element_wise,
                       py::object ambient, py::object vmap) {
  MA ma = as_mesh(mesh);
  simplices_only(ma, "IntegrateCF");
  bool is_p1 = true;
  std::vector<double> v = p1_values(lset, ma->nv, is_p1);
  if (!is_p1) throw std::runtime_error("coefficient functions: straight cut rules of P1 level sets");
  DevBuf<double> d_l;
  d_l.upload(v.data(), v.size());
  DeviceRule r;
  build_rule(ma, d_l.p, dt, order, nullptr, r);
  AmbientCoords amb;
  make_ambient(ma, ambient, vmap, amb);
  DeviceCf dcf;
  make_device_cf(cf, ma->nv, dcf, amb.has_vmap ? &amb.vmap : nullptr, ma->ne);
  const int D = ma->dim;
  std::vector<double> pe(ma->ne, 0.0);
  double total = 0.0;
  if (ma->ne) {
    check(xg_rule_scale_cf(ctx(), D, amb.A, ma->ne, nullptr, ma->el2v.p, amb.buf.p ? amb.buf.p : ma->coords.p,
                           r.offsets.p, r.pts.p, &dcf.cf));
    const double one = 1.0;
    const int32_t ex0[3] = {0, 0, 0};
    DevBuf<double> d_pe(ma->ne), d_tot(1);
    check(xg_integrate(ctx(), D, ma->ne, nullptr, ma->el2v.p, ma->coords.p, r.offsets.p, r.pts.p, 1, &one, ex0,
                       d_pe.p, d_tot.p));
    check(xg_sync(ctx()));
    d_pe.download(pe.data(), ma->ne);
    d_tot.download(&total, 1);
  }
  if (element_wise) return py::array_t<double>((py::ssize_t)pe.size(), pe.data());
  return py::float_(total);
}

// element-boundary element matrices (SymbolicCutBFI with element_boundary,
// symboliccutbfi.cpp:488-640): the cut rules of the listed elements' facets
// on the facet submesh fmesh (facet k of element i = its element i (D + 1) +
// k, level set values flset per vertex copy), mapped into the elements
// (xg_facet_rule_to_element), times cf when given, then xg_elmat with the
// volume basis
py::tuple ElementBoundaryMatrices(py::object mesh, py::object fmesh, py::object flset, int dt, int order, int p,
                                  int mask, py::array_t<int32_t, py::array::c_style | py::array::forcecast> elements,
                                  int precision, py::object cf) {
  MA ma = as_mesh(mesh);
  simplices_only(ma, "ElementBoundaryMatrices");
  MA fm = as_mesh(fmesh);
  const int D = ma->dim, nel = (int)elements.size();
  if (D < 2 || fm->dim != D - 1 || fm->ne != nel * (D + 1))
    throw std::runtime_error("ElementBoundaryMatrices: D + 1 facets per element, of dimension D - 1");
  bool is_p1 = true;
  std::vector<double> v = p1_values(flset, fm->nv, is_p1);
  DevBuf<double> d_l;
  d_l.upload(v.data(), v.size());
  DeviceRule fr;
  build_rule(fm, d_l.p, dt, order, nullptr, fr);
  std::vector<int32_t> el(elements.data(), elements.data() + nel);
  for (int32_t e : el)
    if (e < 0 || e >= ma->ne) throw std::runtime_error("ElementBoundaryMatrices: element out of range");
  DevBuf<int32_t> d_el, eoff((size_t)nel + 1);
  d_el.upload(el.data(), el.size());
  check(xg_facet_rule_to_element(ctx(), D, nel, d_el.p, ma->el2v.p, ma->coords.p, fr.offsets.p, fr.pts.p, eoff.p));
  if (!cf.is_none() && nel) {
    DeviceCf dcf;
    make_device_cf(cf, ma->nv, dcf, nullptr, ma->ne);
    check(xg_rule_scale_cf(ctx(), D, D, nel, d_el.p, ma->el2v.p, ma->coords.p, eoff.p, fr.pts.p, &dcf.cf));
  }
  int nd = 1;
  for (int k = 1; k <= D; ++k) nd = nd * (p + k) / k;
  py::array_t<int32_t> elarr((py::ssize_t)nel, el.data());
  const bool f64 = precision == XG_PREC_FP64;
  DevBuf<double> o64;
  DevBuf<float> o32;
  if (f64) o64.resize((size_t)std::max(nel, 1) * nd * nd);
  else o32.resize((size_t)std::max(nel, 1) * nd * nd);
  if (nel)
    check(xg_elmat(ctx(), D, p, nel, d_el.p, ma->el2v.p, ma->coords.p, eoff.p, fr.pts.p, mask, precision,
                   f64 ? (void *)o64.p : (void *)o32.p));
  check(xg_sync(ctx()));
  if (f64) {
    py::array_t<double> h({(py::ssize_t)nel, (py::ssize_t)nd, (py::ssize_t)nd});
    o64.download(h.mutable_data(), (size_t)nel * nd * nd);
    return py::make_tuple(elarr, h);
  }
  py::array_t<float> h({(py::ssize_t)nel, (py::ssize_t)nd, (py::ssize_t)nd});
  o32.download(h.mutable_data(), (size_t)nel * nd * nd);
  return py::make_tuple(elarr, h);
}

// device scatter of element matrices into a CSR pattern (the assembly loop of
// NGSolve's BilinearForm::Assemble); returns the FP64 values array
py::array_t<double> AssembleCSR(py::array_t<int32_t, py::array::c_style | py::array::forcecast> dofs,
                                py::array_t<double, py::array::c_style | py::array::forcecast> elmats,
                                py::array_t<int32_t, py::array::c_style | py::array::forcecast> row_ptr,
                                py::array_t<int32_t, py::array::c_style | py::array::forcecast> col_idx) {
  if (dofs.ndim() != 2) throw std::runtime_error("AssembleCSR: dofs must be nel x ndof");
  const int nel = (int)dofs.shape(0), nd = (int)dofs.shape(1);
  if ((size_t)elmats.size() != (size_t)nel * nd * nd) throw std::runtime_error("AssembleCSR: elmats size");
  const size_t nnz = (size_t)col_idx.size();
  DevBuf<int32_t> d_dofs, d_rp, d_ci;
  DevBuf<double> d_m, d_v(nnz ? nnz : 1);
  d_dofs.upload(dofs.data(), dofs.size());
  d_m.upload(elmats.data(), elmats.size());
  d_rp.upload(row_ptr.data(), row_ptr.size());
  d_ci.upload(col_idx.data(), col_idx.size());
  if (nnz) hipcheck(hipMemset(d_v.p, 0, sizeof(double) * nnz), "hipMemset");
  if (nel) check(xg_assemble_csr(ctx(), nel, nd, d_dofs.p, d_m.p, XG_PREC_FP64, d_rp.p, d_ci.p, d_v.p));
  check(xg_sync(ctx()));
  py::array_t<double> h((py::ssize_t)nnz);
  d_v.download(h.mutable_data(), nnz);
  return h;
}

// BilinearForm.Assemble with everything on the device.  NGSolve's
// BilinearForm::Assemble builds the matrix graph of the space and adds every
// element's CalcElementMatrix (symboliccutbfi.cpp:37-46) into it; here the
// graph of the whole element dof table is built on the GPU
// (xg_csr_pattern), each cut integral's element matrices stay in device
// memory (xg_cut_pipeline for P1 level sets, xg_subdiv_pipeline for P2 ones
// or subdivlvl > 0 with exact base rules) and are scattered with their
// coefficient (xg_assemble_csr_scaled).  Only the CSR arrays return.
// terms: [(lset, dt, order, p, mask, elements, coef, subdivlvl), ...];
// precision: the element-matrix arithmetic of xg_cut_pipeline (P1 terms).
py::tuple AssembleCutFormsDevice(py::object mesh, py::array_t<int32_t, py::array::c_style | py::array::forcecast> dofs,
                                 int n, py::list terms, int precision, py::object pattern_elements) {
  MA ma = as_mesh(mesh);
  simplices_only(ma, "AssembleCutFormsDevice");
  if (dofs.ndim() != 2) throw std::runtime_error("AssembleCutFormsDevice: dofs must be ne x ndof");
  const int ne = (int)dofs.shape(0), nd = (int)dofs.shape(1);
  if (ne != ma->ne) throw std::runtime_error("AssembleCutFormsDevice: one dof row per element");
  // matrix graph: the dofs of all elements, or of the active elements of a
  // restricted bilinear form (MeshEntityToDofTable with active_elems,
  // utils/restrictedblf.cpp:8-64)
  std::vector<int32_t> prow;
  if (!pattern_elements.is_none()) {
    std::vector<int32_t> pe = element_list(pattern_elements, ne);
    prow.resize(pe.size() * (size_t)nd);
    for (size_t i = 0; i < pe.size(); ++i) {
      if (pe[i] < 0 || pe[i] >= ne) throw std::runtime_error("AssembleCutFormsDevice: element out of range");
      std::copy(dofs.data() + (size_t)pe[i] * nd, dofs.data() + (size_t)(pe[i] + 1) * nd, prow.begin() + i * nd);
    }
  }
  const int npat = pattern_elements.is_none() ? ne : (int)(prow.size() / nd);
  DevBuf<int32_t> d_dofs, rp((size_t)n + 1), nnz(1);
  if (pattern_elements.is_none()) d_dofs.upload(dofs.data(), dofs.size());
  else d_dofs.upload(prow.data(), prow.size());
  const long long cap = std::max((long long)npat * nd * nd, 1LL);
  DevBuf<int32_t> ci((size_t)cap);
  check(xg_csr_pattern(ctx(), npat, nd, d_dofs.p, n, rp.p, ci.p, cap, nnz.p));
  check(xg_sync(ctx()));
  int32_t h_nnz = 0;
  nnz.download(&h_nnz, 1);
  DevBuf<double> vals((size_t)std::max(h_nnz, 1));
  if (h_nnz) hipcheck(hipMemset(vals.p, 0, sizeof(double) * h_nnz), "hipMemset");
  const int32_t *hd = dofs.data();
  for (py::handle th : terms) {
    py::tuple t = th.cast<py::tuple>();
    py::object lset = t[0];
    const int dt = t[1].cast<int>(), order = t[2].cast<int>(), p = t[3].cast<int>(), mask = t[4].cast<int>();
    const double coef = t[6].cast<double>();
    const int L = t[7].cast<int>();
    std::vector<int32_t> el = element_list(t[5], ma->ne);
    const int nel = (int)el.size();
    if (nel == 0) continue;
    int pd = 1;
    for (int k = 1; k <= ma->dim; ++k) pd = pd * (p + k) / k;
    if (pd != nd) throw std::runtime_error("AssembleCutFormsDevice: the form's order does not match the dof table");
    std::vector<int32_t> edofs((size_t)nel * nd);
    for (int i = 0; i < nel; ++i)
      std::copy(hd + (size_t)el[i] * nd, hd + (size_t)(el[i] + 1) * nd, edofs.begin() + (size_t)i * nd);
    DevBuf<int32_t> d_el, d_edofs;
    d_el.upload(el.data(), el.size());
    d_edofs.upload(edofs.data(), edofs.size());
    const int dts[1] = {dt};
    const bool subdiv = needs_subdivision(lset, L);
    int mprec = precision & ~XG_METHOD_MOMENTS;
    if (subdiv) {
      if (order < 2 * p)
        throw std::runtime_error("AssembleCutFormsDevice: subdivision rules with order >= 2p (exact base rules)");
      mprec = mprec == XG_PREC_FP64 ? XG_PREC_FP64 : XG_PREC_FP32;
    }
    const bool f64 = mprec == XG_PREC_FP64;
    DevBuf<double> o64;
    DevBuf<float> o32;
    void *out = nullptr;
    if (f64) {
      o64.resize((size_t)nel * nd * nd);
      out = o64.p;
    } else {
      o32.resize((size_t)nel * nd * nd);
      out = o32.p;
    }
    if (subdiv) {
      std::vector<int32_t> all(ma->ne);
      for (int e = 0; e < ma->ne; ++e) all[e] = e;
      std::vector<double> nod = p2_nodal_values(ma, lset, all);
      DevBuf<double> d_nod;
      d_nod.upload(nod.data(), nod.size());
      check(xg_subdiv_pipeline(ctx(), p, L, order, nel, d_el.p, ma->el2v.p, ma->coords.p, d_nod.p, 1, dts, mask, mprec,
                               out, nullptr));
      check(xg_assemble_csr_scaled(ctx(), nel, nd, d_edofs.p, out, mprec, coef, rp.p, ci.p, vals.p));
      check(xg_sync(ctx()));
    } else {
      bool is_p1 = true;
      std::vector<double> v = p1_values(lset, ma->nv, is_p1);
      DevBuf<double> d_l;
      d_l.upload(v.data(), v.size());
      check(xg_cut_pipeline(ctx(), ma->dim, p, order, nel, d_el.p, ma->el2v.p, ma->coords.p, d_l.p, 1, dts, mask,
                            precision, out));
      check(xg_assemble_csr_scaled(ctx(), nel, nd, d_edofs.p, out, f64 ? XG_PREC_FP64 : XG_PREC_FP32, coef, rp.p,
                                   ci.p, vals.p));
      check(xg_sync(ctx()));
    }
  }
  check(xg_sync(ctx()));
  py::array_t<int32_t> h_rp((py::ssize_t)n + 1), h_ci((py::ssize_t)h_nnz);
  py::array_t<double> h_v((py::ssize_t)h_nnz);
  rp.download(h_rp.mutable_data(), (size_t)n + 1);
  ci.download(h_ci.mutable_data(), (size_t)h_nnz);
  vals.download(h_v.mutable_data(), (size_t)h_nnz);
  return py::make_tuple(h_rp, h_ci, h_v);
}

py::array_t<double> AssembleVector(py::array_t<int32_t, py::array::c_style | py::array::forcecast> dofs,
                                   py::array_t<double, py::array::c_style | py::array::forcecast> elvecs, int n) {
  const int nel = (int)dofs.shape(0), nd = (int)dofs.shape(1);
  if ((size_t)elvecs.size() != (size_t)nel * nd) throw std::runtime_error("AssembleVector: elvecs size");
  DevBuf<int32_t> d_dofs;
  DevBuf<double> d_e, d_o((size_t)(n ? n : 1));
  d_dofs.upload(dofs.data(), dofs.size());
  d_e.upload(elvecs.data(), elvecs.size());
  if (n) hipcheck(hipMemset(d_o.p, 0, sizeof(double) * n), "hipMemset");
  if (nel) check(xg_assemble_vec(ctx(), nel, nd, d_dofs.p, d_e.p, XG_PREC_FP64, d_o.p));
  check(xg_sync(ctx()));
  py::array_t<double> h((py::ssize_t)n);
  d_o.download(h.mutable_data(), (size_t)n);
  return h;
}

}  // namespace

PYBIND11_MODULE(xfem, m) {
  m.doc() = "MI355X cut-element hot path (HIP kernels behind the xg C ABI)";

  py::enum_<DOMAIN_TYPE>(m, "DOMAIN_TYPE").value("POS", POS).value("NEG", NEG).value("IF", IF).export_values();
  py::enum_<COMBINED_DOMAIN_TYPE>(m, "COMBINED_DOMAIN_TYPE")
      .value("NO", CDOM_NO)
      .value("CDOM_NEG", CDOM_NEG)
      .value("CDOM_POS", CDOM_POS)
      .value("UNCUT", CDOM_UNCUT)
      .value("CDOM_IF", CDOM_IF)
      .value("HASNEG", CDOM_HASNEG)
      .value("HASPOS", CDOM_HASPOS)
      .value("ANY", CDOM_ANY)
      .export_values();
  py::enum_<SWAP_DIMENSIONS_POLICY>(m, "QUAD_DIRECTION_POLICY")
      .value("FIRST", FIRST_ALLOWED)
      .value("OPTIMAL", FIND_OPTIMAL)
      .value("FALLBACK", ALWAYS_NONE)
      .export_values();
  py::enum_<VorB>(m, "VorB").value("VOL", VOL).value("BND", BND).export_values();
  py::enum_<NODE_TYPE>(m, "NODE_TYPE")
      .value("VERTEX", NT_VERTEX)
      .value("EDGE", NT_EDGE)
      .value("FACE", NT_FACE)
      .value("CELL", NT_CELL)
      .value("ELEMENT", NT_ELEMENT)
      .value("FACET", NT_FACET);
  m.attr("NT_VERTEX") = NT_VERTEX;
  m.attr("NT_EDGE") = NT_EDGE;
  m.attr("NT_FACE") = NT_FACE;
  m.attr("NT_CELL") = NT_CELL;
  m.attr("NT_ELEMENT") = NT_ELEMENT;
  m.attr("NT_FACET") = NT_FACET;

  py::class_<BitArray, BA>(m, "BitArray")
      .def(py::init<size_t>())
      .def("__len__", &BitArray::Size)
      .def("__getitem__", [](const BitArray &b, long i) {
        if (i < 0) i += (long)b.Size();
        if (i < 0 || (size_t)i >= b.Size()) throw py::index_error();
        return b.Test((size_t)i);
      })
      .def("__setitem__", [](BitArray &b, long i, bool v) {
        if (i < 0 || (size_t)i >= b.Size()) throw py::index_error();
        if (v) b.SetBit(i);
        else b.ClearBit(i);
      })
      .def("NumSet", &BitArray::NumSet)
      .def("Set", [](BitArray &b) { b.Set(); })
      .def("Clear", [](BitArray &b) { b.Clear(); })
      .def("__or__", [](const BitArray &a, const BitArray &b) {
        if (a.Size() != b.Size()) throw std::runtime_error("BitArray size mismatch");
        auto r = std::make_shared<BitArray>(a.Size());
        for (size_t k = 0; k < r->words().size(); ++k) r->words()[k] = a.words()[k] | b.words()[k];
        return r;
      })
      .def("__and__", [](const BitArray &a, const BitArray &b) {
        if (a.Size() != b.Size()) throw std::runtime_error("BitArray size mismatch");
        auto r = std::make_shared<BitArray>(a.Size());
        for (size_t k = 0; k < r->words().size(); ++k) r->words()[k] = a.words()[k] & b.words()[k];
        return r;
      })
      .def("__invert__", [](const BitArray &a) {
        auto r = std::make_shared<BitArray>(a.Size());
        for (size_t k = 0; k < r->words().size(); ++k) r->words()[k] = ~a.words()[k];
        r->trim();
        return r;
      })
      .def("words", [](const BitArray &b) {
        return py::array_t<uint32_t>((py::ssize_t)b.words().size(), b.words().data());
      })
      .def("numpy", [](const BitArray &b) {
        py::array_t<bool> r((py::ssize_t)b.Size());
        auto v = r.mutable_unchecked<1>();
        for (size_t i = 0; i < b.Size(); ++i) v(i) = b.Test(i);
        return r;
      });

  py::class_<MeshAccess, MA>(m, "MeshAccess")
      .def(py::init([](py::object mesh) { return as_mesh(mesh); }))
      .def_readonly("dim", &MeshAccess::dim)
      .def_readonly("ne", &MeshAccess::ne)
      .def_readonly("nv", &MeshAccess::nv);

  py::class_<CutInformation, std::shared_ptr<CutInformation>>(m, "CutInfo")
      .def(py::init([](py::object mesh, py::object levelset, int subdivlvl, int time_order, int heapsize) {
             auto ci = std::make_shared<CutInformation>(as_mesh(mesh));
             if (!levelset.is_none()) ci->Update(levelset, subdivlvl, time_order);
             return ci;
           }),
           py::arg("mesh"), py::arg("levelset") = py::none(), py::arg("subdivlvl") = 0, py::arg("time_order") = -1,
           py::arg("heapsize") = 1000000)
      .def("Update",
           [](CutInformation &self, py::object lset, int subdivlvl, int time_order, int heapsize) {
             self.Update(lset, subdivlvl, time_order);
           },
           py::arg("levelset"), py::arg("subdivlvl") = 0, py::arg("time_order") = -1, py::arg("heapsize") = 1000000)
      .def("Mesh", [](CutInformation &self) { return self.ma->pymesh; })
      .def("GetElementsOfType",
           [](CutInformation &self, py::object dt, VorB vb) { return self.GetElementsOfDomainType(parse_cdt(dt), vb); },
           py::arg("domain_type") = IF, py::arg("VOL_or_BND") = VOL)
      .def("GetFacetsOfType", [](CutInformation &self, py::object dt) { return self.GetFacetsOfDomainType(parse_cdt(dt)); },
           py::arg("domain_type") = IF)
      .def("GetCutNeighboringNodes",
           [](CutInformation &self, int nt) { return self.cutnb.at(self.node_index(nt)); }, py::arg("node_type"),
           "cut_neighboring_node[node_type] (cutinfo.hpp:30): the nodes of cut elements")
      .def("GetDomainOfNodes",
           [](CutInformation &self, int nt) {
             const auto &d = self.dom.at(self.node_index(nt));
             return py::array_t<uint8_t>((py::ssize_t)d.size(), d.data());
           },
           py::arg("node_type"), "dom_of_node[node_type] (cutinfo.hpp:32) as DOMAIN_TYPE codes (uint8)")
      .def("GetCutRatios",
           [](CutInformation &self, VorB vb) {
             const auto &c = self.cut_ratio[vb];
             return py::array_t<double>((py::ssize_t)c.size(), c.data());
           },
           py::arg("VOL_or_BND") = VOL)
      .def("GetElementsWithThresholdContribution",
           [](CutInformation &self, py::object dt, double threshold, VorB vb) {
             int d = -1;
             try {
               d = (int)dt.cast<DOMAIN_TYPE>();
             } catch (...) {
             }
             if (d != NEG && d != POS)
               throw std::runtime_error("Unknown/Invalid type for dt: Only POS, NEG are implemented a.t.m.");
             return self.GetElementsWithThresholdContribution((DOMAIN_TYPE)d, threshold, vb);
           },
           py::arg("domain_type") = NEG, py::arg("threshold") = 1.0, py::arg("VOL_or_BND") = VOL);

  m.def("GetFacetsWithNeighborTypes",
        [](py::object mesh, BA a, bool bv_a, bool bv_b, bool use_and, py::object bb, int heapsize) {
          MA ma = as_mesh(mesh);
          BA b = bb.is_none() ? a : bb.cast<BA>();
          ma->ensure_facets();
          auto ret = std::make_shared<BitArray>(ma->nf);
          if (ma->nf == 0) return ret;
          DevBuf<uint32_t> da, db, dout;
          da.upload(a->words().data(), a->words().size());
          db.upload(b->words().data(), b->words().size());
          dout.resize(ret->words().size());
          check(xg_facets_with_neighbor_types(ctx(), ma->nf, ma->f2el.p, da.p, db.p, bv_a, bv_b, use_and, dout.p));
          check(xg_sync(ctx()));
          dout.download(ret->words().data(), ret->words().size());
          return ret;
        },
        py::arg("mesh"), py::arg("a"), py::arg("bnd_val_a") = true, py::arg("bnd_val_b") = true,
        py::arg("use_and") = true, py::arg("b") = py::none(), py::arg("heapsize") = 1000000);

  // cutinfo.cpp:540-647 through xg_nodes.hip
  m.def("GetElementsWithNeighborFacets",
        [](py::object mesh, BA a, int heapsize) {
          MA ma = as_mesh(mesh);
          ma->ensure_facets();
          if ((int)a->Size() != ma->nf) throw std::runtime_error("GetElementsWithNeighborFacets: a must mark facets");
          auto ret = std::make_shared<BitArray>(ma->ne);
          DevBuf<uint32_t> da, dout(ret->words().size());
          da.upload(a->words().data(), a->words().size());
          check(xg_elements_with_neighbor_facets(ctx(), ma->nf, ma->f2el.p, da.p, ma->ne, dout.p));
          check(xg_sync(ctx()));
          dout.download(ret->words().data(), ret->words().size());
          return ret;
        },
        py::arg("mesh"), py::arg("a"), py::arg("heapsize") = 1000000);
  m.def("GetElementsWithSharedVertex",
        [](py::object mesh, BA a, int heapsize) {
          MA ma = as_mesh(mesh);
          if ((int)a->Size() != ma->ne) throw std::runtime_error("GetElementsWithSharedVertex: a must mark elements");
          auto ret = std::make_shared<BitArray>(ma->ne);
          DevBuf<uint32_t> da, dout(ret->words().size());
          da.upload(a->words().data(), a->words().size());
          check(xg_elements_with_shared_vertex(ctx(), ma->nvel, ma->ne, ma->el2v.p, da.p, ma->nv, dout.p));
          check(xg_sync(ctx()));
          dout.download(ret->words().data(), ret->words().size());
          return ret;
        },
        py::arg("mesh"), py::arg("a"), py::arg("heapsize") = 1000000);
  // the space stand-in gives its dof tables: dofs() -> (ne x nd, ndof) and
  // mesh.facet_dofs(order) -> nf x nfd
  auto dofs_of = [](py::object space, BA a, bool facets) {
    py::tuple t = space.attr("dofs")();
    const int ndof = t[1].cast<int>();
    py::object tab_obj = facets ? space.attr("mesh").attr("facet_dofs")(space.attr("order")) : t[0];
    auto tab = py::array_t<int32_t, py::array::c_style | py::array::forcecast>(tab_obj);
    const int n = (int)tab.shape(0), k = (int)tab.shape(1);
    if ((int)a->Size() != n)
      throw std::runtime_error(facets ? "GetDofsOfFacets: a must mark facets" : "GetDofsOfElements: a must mark elements");
    auto ret = std::make_shared<BitArray>(ndof);
    DevBuf<int32_t> dtab;
    DevBuf<uint32_t> da, dout(ret->words().size());
    dtab.upload(tab.data(), (size_t)n * k);
    da.upload(a->words().data(), a->words().size());
    check(xg_mark_from_items(ctx(), n, k, dtab.p, da.p, ndof, dout.p));
    check(xg_sync(ctx()));
    dout.download(ret->words().data(), ret->words().size());
    return ret;
  };
  m.def("GetDofsOfElements", [dofs_of](py::object space, BA a, int heapsize) { return dofs_of(space, a, false); },
        py::arg("space"), py::arg("a"), py::arg("heapsize") = 1000000);
  m.def("GetDofsOfFacets", [dofs_of](py::object space, BA a, int heapsize) { return dofs_of(space, a, true); },
        py::arg("space"), py::arg("a"), py::arg("heapsize") = 1000000);

  py::class_<MultiLevelsetCutInformation, std::shared_ptr<MultiLevelsetCutInformation>>(m, "MultiLevelsetCutInfo")
      .def(py::init([](py::object mesh, py::object lsets) {
             return std::make_shared<MultiLevelsetCutInformation>(as_mesh(mesh), lsets);
           }),
           py::arg("mesh"), py::arg("levelset"))
      .def("Update", [](MultiLevelsetCutInformation &self, py::object l, int heapsize) { self.Update(l); },
           py::arg("levelset"), py::arg("heapsize") = 1000000)
      .def("Mesh", [](MultiLevelsetCutInformation &self) { return self.ma->pymesh; })
      .def("GetElementsOfType",
           [](MultiLevelsetCutInformation &self, py::object dts, VorB vb, int heapsize) {
             return self.mark(dts, false, vb);
           },
           py::arg("domain_type"), py::arg("VOL_or_BND") = VOL, py::arg("heapsize") = 1000000)
      .def("GetElementsWithContribution",
           [](MultiLevelsetCutInformation &self, py::object dts, VorB vb, int heapsize) {
             return self.mark(dts, true, vb);
           },
           py::arg("domain_type"), py::arg("VOL_or_BND") = VOL, py::arg("heapsize") = 1000000);

  m.def("IntegrateX", &IntegrateX, py::arg("levelset_domain"), py::arg("mesh"), py::arg("cf") = py::none(),
        py::arg("deformation") = py::none(), py::arg("ip_container") = py::none(), py::arg("element_wise") = false,
        py::arg("heapsize") = 1000000);

  m.def("InterpolateCFToP1",
        [](py::object mesh, py::object cf, double eps) {
          MA ma = as_mesh(mesh);
          DeviceCf dcf;
          make_device_cf(cf, ma->nv, dcf);
          DevBuf<double> out((size_t)std::max(ma->nv, 1));
          check(xg_p1_interpolate_cf(ctx(), ma->dim, ma->nv, ma->coords.p, &dcf.cf, eps, out.p));
          check(xg_sync(ctx()));
          py::array_t<double> r(ma->nv);
          out.download(r.mutable_data(), (size_t)ma->nv);
          return r;
        },
        py::arg("mesh"), py::arg("cf"), py::arg("eps_perturbation") = 1e-14,
        "vertex values of a coefficient function on the GPU with the P1 perturbation (InterpolateP1)");
  m.def("PerturbP1",
        [](py::array_t<double, py::array::c_style | py::array::forcecast> v, double eps) {
          DevBuf<double> d;
          d.upload(v.data(), v.size());
          check(xg_p1_perturb(ctx(), (int)v.size(), d.p, eps));
          check(xg_sync(ctx()));
          py::array_t<double> r(v.size());
          d.download(r.mutable_data(), v.size());
          return r;
        },
        py::arg("values"), py::arg("eps_perturbation") = 1e-14);

  m.def("CutElementVectors", &CutElementVectors, py::arg("mesh"), py::arg("levelset"), py::arg("domain_type"),
        py::arg("order"), py::arg("p"), py::arg("elements") = py::none(),
        "element vectors of the source term f = 1 on the straight cut rule (FP64)");
  m.def("AssembleCutFormsDevice", &AssembleCutFormsDevice, py::arg("mesh"), py::arg("dofs"), py::arg("n"),
        py::arg("terms"), py::arg("precision"), py::arg("pattern_elements") = py::none(),
        "BilinearForm.Assemble on the device: pattern, element matrices and scatter; returns CSR arrays");
  m.def("AssembleCSR", &AssembleCSR, py::arg("dofs"), py::arg("elmats"), py::arg("row_ptr"), py::arg("col_idx"),
        "device scatter of element matrices into a CSR pattern (FP64 values)");
  m.def("AssembleVector", &AssembleVector, py::arg("dofs"), py::arg("elvecs"), py::arg("n"),
        "device scatter of element vectors");
  m.def("SubdivElementMatrices", &SubdivElementMatrices, py::arg("mesh"), py::arg("levelset"),
        py::arg("domain_type"), py::arg("subdivlvl"), py::arg("order"), py::arg("p"), py::arg("mask") = 3,
        py::arg("elements") = py::none(), py::arg("precision") = XG_PREC_FP32,
        "element matrices of (mass | laplace) on the subdivision rule of a P2 level set (tetrahedra)");
  m.def("MlsetElementMatrices", &MlsetElementMatrices, py::arg("mesh"), py::arg("levelsets"),
        py::arg("domain_type"), py::arg("order"), py::arg("p"), py::arg("mask") = 3,
        py::arg("elements") = py::none(), py::arg("precision") = XG_PREC_FP32, py::arg("cf") = py::none(),
        "element matrices of (mass | laplace) on the rule of several level sets (times cf when given)");
  m.def("CutElementMatricesCF", &CutElementMatricesCF, py::arg("mesh"), py::arg("levelset"),
        py::arg("domain_type"), py::arg("order"), py::arg("p"), py::arg("mask"), py::arg("elements"),
        py::arg("precision"), py::arg("cf"), py::arg("ambient") = py::none(), py::arg("vmap") = py::none(),
        "element matrices of cf u v (mask 1) / cf grad u . grad v (mask 2), cf evaluated on the GPU");
  m.def("ElementBoundaryMatrices", &ElementBoundaryMatrices, py::arg("mesh"), py::arg("facet_mesh"),
        py::arg("facet_levelset"), py::arg("domain_type"), py::arg("order"), py::arg("p"), py::arg("mask"),
        py::arg("elements"), py::arg("precision"), py::arg("cf") = py::none(),
        "element matrices on the cut rules of the elements' boundaries (facets), volume basis");
  m.def("IntegrateCF", &IntegrateCF, py::arg("mesh"), py::arg("levelset"), py::arg("domain_type"), py::arg("order"),
        py::arg("cf"), py::arg("element_wise") = false, py::arg("ambient") = py::none(), py::arg("vmap") = py::none(),
        "integral of cf on the straight cut rule, cf evaluated on the GPU");
  m.def("CutElementVectorsCF", &CutElementVectorsCF, py::arg("mesh"), py::arg("levelset"), py::arg("domain_type"),
        py::arg("order"), py::arg("p"), py::arg("elements"), py::arg("cf"), py::arg("ambient") = py::none(),
        py::arg("vmap") = py::none(), "element vectors of cf v on the straight cut rule, cf evaluated on the GPU (FP64)");
  m.def("CutElementMatrices", &CutElementMatrices, py::arg("mesh"), py::arg("levelset"), py::arg("domain_type"),
        py::arg("order"), py::arg("p"), py::arg("mask") = 3, py::arg("elements") = py::none(),
        py::arg("precision") = XG_PREC_FP32);
  m.def("QuadElementMatrices", &QuadElementMatrices, py::arg("mesh"), py::arg("lset"), py::arg("domain_type"),
        py::arg("order"), py::arg("policy") = (int)FIND_OPTIMAL, py::arg("mask") = 3,
        py::arg("elements") = py::none(), py::arg("precision") = (int)XG_PREC_FP64);
}
