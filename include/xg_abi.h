/*
 * xg_abi.h -- C ABI of the MI355X (gfx950) cut-element hot path.
 *
 * Plain pointers and sizes only.  Unless stated otherwise every array
 * argument is a DEVICE pointer (hipMalloc'ed or a torch CUDA tensor's
 * data_ptr); outputs are caller-allocated.  All calls are asynchronous on the
 * context's stream; results are valid after xg_sync() (or any later work on
 * the same stream).  Every function returns XG_OK (0) or a negative XG_ERR_*
 * code; xg_last_error() gives the message of the last failure on the calling
 * host thread.  One context per host thread.
 *
 * Conventions (shared with the reference):
 *   DOMAIN_TYPE      NEG = 0, POS = 1, IF = 2          (utils/ngsxstd.hpp:16-20)
 *   COMBINED_DOMAIN  NO=0 NEG=1 POS=2 UNCUT=3 IF=4 HASNEG=5 HASPOS=6 ANY=7
 *                                                       (utils/ngsxstd.hpp:23-31)
 *   bit arrays       uint32 words, element e is bit (e & 31) of word e >> 5;
 *                    byte-for-byte the layout of an NGSolve BitArray.
 *   mesh             coords: nv x D FP64, el2v: ne x (D+1) int32.
 *   reference simplex vertices (1,0,0), (0,1,0), (0,0,1), (0,0,0); the element
 *                    map sends reference vertex k to el2v[e][k].
 *   rules            CSR: offsets int32[n+1]; points double4 (x, y, z, w) in
 *                    reference coordinates; the physical weight is w * |det F|
 *                    (straightcutrule.cpp:580-581: "a scaling with
 *                    mip.GetMeasure() gives the correct weight").
 */
#ifndef XG_ABI_H
#define XG_ABI_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define XG_OK 0
#define XG_ERR_ARG (-1)        /* invalid argument / size */
#define XG_ERR_HIP (-2)        /* HIP runtime failure */
#define XG_ERR_DEGENERATE (-3) /* straightcutrule.cpp:45-49,148,193: cut of a vertex value 0 */
#define XG_ERR_CAPACITY (-4)   /* output capacity exceeded */
#define XG_ERR_SUBDIV (-5)     /* xintegration.cpp:1386 "did not expect this.. -3-" */
#define XG_ERR_UNSUPPORTED (-6)

#define XG_NEG 0
#define XG_POS 1
#define XG_IF 2

/* element-matrix arithmetic */
#define XG_PREC_FP64 0   /* FP64 MFMA (v_mfma_f64_16x16x4f64), reference mode */
#define XG_PREC_FP32 1   /* FP32 MFMA (v_mfma_f32_16x16x4f32), FP32 accumulate */
#define XG_PREC_BF16X3 2 /* 3-way split BF16 MFMA (6 products), FP32 accumulate */
/* OR-ed into the precision of xg_cut_pipeline: moment form.  Per (element,
   domain type) the FP64 barycentric moments sum_i w_i lam(x_i)^alpha,
   |alpha| = 2p, of the same cut rule are accumulated on the vector ALU and
   the mass / stiffness entries formed from them by exact polynomial
   identities (the same quadrature sum, reassociated).  FP64 arithmetic
   throughout; the output is FP32 (XG_PREC_FP32) or FP64 (XG_PREC_FP64). */
#define XG_METHOD_MOMENTS 16

/* element-matrix form mask */
#define XG_FORM_MASS 1      /* u*v */
#define XG_FORM_LAPLACE 2   /* grad(u)*grad(v) */

typedef struct xg_ctx xg_ctx;

/* ---- context --------------------------------------------------------- */
/* No reference counterpart (NGSolve's LocalHeap + TaskManager play this role,
   utils/ngsxstd.cpp:3-31): binds a device and a stream. */
int xg_ctx_create(int device, xg_ctx **out);
int xg_ctx_destroy(xg_ctx *ctx);
/* stream: a hipStream_t (0 = the null stream) */
int xg_ctx_set_stream(xg_ctx *ctx, void *stream);
int xg_sync(xg_ctx *ctx);
const char *xg_last_error(void);
/* number of quadrature points of the base rule of dimension D, order */
int xg_rule_npts(int D, int order);

/* ---- a2: InterpolateP1 perturbation (utils/p1interpol.cpp:104-107) ----- */
/* v[i] = eps if |v[i]| < eps, in place (n values, device). */
int xg_p1_perturb(xg_ctx *ctx, int n, double *v, double eps);
/* struct xg_cf: a coefficient function's postfix program (declared below) */
struct xg_cf;
/* InterpolateP1::Do (utils/p1interpol.cpp:27-113) for a coefficient
   function: out[v] = cf(x_v) at the nv vertices (coords nv x D, device),
   then v = eps where |v| < eps.  Grid function opcodes read the vertex's own
   value. */
int xg_p1_interpolate_cf(xg_ctx *ctx, int D, int nv, const double *coords, const struct xg_cf *cf, double eps,
                         double *out);

/* ---- a5, a12: CutInformation::Update marking (xfem/cutinfo.cpp:61-133) */
/* D = element dimension (1, 2, 3); el2v: ne x (D+1).
   dt (nullable): uint8[ne] DOMAIN_TYPE per element.
   cut_ratio (nullable): FP64[ne] = part_vol[NEG] / (part_vol[NEG] + part_vol[POS]).
   bits: uint32[8][nw], nw = ceil(ne/32): the 8 COMBINED_DOMAIN_TYPE arrays.
   cut_list (nullable): int32[ne], receives the IF elements in ascending order;
   n_cut (nullable, device int32[1]) receives their number.
   Bit-exact with the reference algorithm (FP64, no contraction). */
int xg_mark(xg_ctx *ctx, int D, int ne, const int32_t *el2v, const double *lset_v,
            uint8_t *dt, double *cut_ratio, uint32_t *bits, int32_t *cut_list,
            int32_t *n_cut);

/* xg_mark for callers that cannot allocate device memory (code under HIP
   graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
   counterpart in the reference.  Same arguments and outputs as xg_mark, bit
   for bit; the block counts and scan partials that xg_mark keeps in the
   context workspace live in a fixed file-scope device array instead
   (1 MiB per GPU, held until the process ends), so this call never
   allocates.  ne <= XG_MARK_NOALLOC_MAX_NE, else XG_ERR_CAPACITY.  Calls
   that overlap on one GPU (several streams or contexts) share that array
   and race: serialise them. */
#define XG_MARK_NOALLOC_MAX_NE (1 << 28)
int xg_mark_noalloc(xg_ctx *ctx, int D, int ne, const int32_t *el2v, const double *lset_v,
                    uint8_t *dt, double *cut_ratio, uint32_t *bits, int32_t *cut_list,
                    int32_t *n_cut);

/* ---- a12 for P2 level sets (cutinfo.cpp:61-133 with the subdivision rule of
   xintegration.cpp:617-752): the element domain is the lattice check of the
   level-L subdivision rule (IF iff the level set takes values >= 0 and < 0 on
   the C(2^L+3,3) lattice points).  lset_nodal: ne x 10 FP64 nodal values
   (Lattice<3,2> dof order); dt: uint8[ne] (required); bits: the 8
   COMBINED_DOMAIN_TYPE arrays as in xg_mark (required); cut_list / n_cut as
   in xg_mark (nullable).  No cut ratio (it needs the full rules). */
int xg_mark_p2(xg_ctx *ctx, int ne, const double *lset_nodal, int L, uint8_t *dt, uint32_t *bits,
               int32_t *cut_list, int32_t *n_cut);

/* a12 on the boundary triangles of a tetrahedral mesh with a P2 level set
   (cutinfo.cpp:80-124, BND elements, with the subdivision rule on triangles,
   xintegration.cpp:620-838,1019-1061 and CutSimplex<2>::MakeQuad
   :1391-1545): the order-0 NEG and POS areas of the level-L rule decide the
   domain (IF iff both > 0, else the nonzero side) and the cut ratio
   NEG / (NEG + POS).  lset_nodal: ne x 6 FP64 nodal values (vertices, then
   the edges (0,1), (0,2), (1,2)); dt, cut_ratio, bits required; cut_list /
   n_cut as in xg_mark.  L <= 5. */
int xg_mark_p2_trig(xg_ctx *ctx, int ne, const double *lset_nodal, int L, uint8_t *dt,
                    double *cut_ratio, uint32_t *bits, int32_t *cut_list, int32_t *n_cut);

/* ---- a13: GetFacetsWithNeighborTypes (xfem/cutinfo.cpp:476-538) ------- */
/* facet f has neighbours f2el[f][0], f2el[f][1] (-1 = none).  a, b: element
   bit arrays; result bit f = use_and ? (a(L)&b(R) | a(R)&b(L))
   : (a(L)|b(R) ... as the reference); boundary sides use bnd_val_a/b. */
int xg_facets_with_neighbor_types(xg_ctx *ctx, int nf, const int32_t *f2el,
                                  const uint32_t *a, const uint32_t *b, int bnd_val_a,
                                  int bnd_val_b, int use_and, uint32_t *out);

/* ---- node and dof marking (xfem/cutinfo.cpp:135-224, 540-647) -------- */
/* Bit arrays have the BitArray layout above (bit i of word i/32).
   out (nn bits) is cleared, then bit d is set for every node d of every item
   whose bit is set in mark (mark == NULL: every item).  item2node: n x k
   int32 node numbers; negative entries are skipped; an entry >= nn is
   reported by xg_sync (XG_ERR_ARG).  Items = elements with their dof table:
   GetDofsOfElements (cutinfo.cpp:598-622); items = facets with their dof
   table: GetDofsOfFacets (cutinfo.cpp:624-647). */
int xg_mark_from_items(xg_ctx *ctx, int n, int k, const int32_t *item2node, const uint32_t *mark,
                       int nn, uint32_t *out);
/* GetElementsWithNeighborFacets (cutinfo.cpp:540-564): f2el as in
   xg_facets_with_neighbor_types; out: ne bits. */
int xg_elements_with_neighbor_facets(xg_ctx *ctx, int nf, const int32_t *f2el,
                                     const uint32_t *facet_mark, int ne, uint32_t *out);
/* GetElementsWithSharedVertex (cutinfo.cpp:566-596): elements with a vertex
   of a marked element (the marked elements included); el2v: ne x k vertex
   numbers (k = D+1 for simplices, 4 / 8 for quadrilaterals / hexahedra);
   out: ne bits.  Uses ceil(nv/32) words of the context scratch. */
int xg_elements_with_shared_vertex(xg_ctx *ctx, int k, int ne, const int32_t *el2v,
                                   const uint32_t *elem_mark, int nv, uint32_t *out);
/* cut_neighboring_node / dom_of_node of one node type (cutinfo.cpp:135-224)
   from the element domains dt (uint8, NEG 0 / POS 1 / IF 2, as written by
   xg_mark).  el2node: ne x k node numbers of the type (vertices, edges or
   faces); cutnb (nullable, nn bits): the nodes of IF elements are OR-ed in
   (not cleared: the reference clears these arrays only in the CutInformation
   constructor, cutinfo.cpp:28-34); dom (nullable, uint8[nn]): NEG for nodes
   of NEG elements, else POS.  el2node == NULL selects the element level
   (nn == ne): cutnb |= the IF elements, dom = NEG where cut_ratio > 0.5,
   else POS. */
int xg_node_domains(xg_ctx *ctx, int ne, int k, const int32_t *el2node, const uint8_t *dt,
                    const double *cut_ratio, int nn, uint32_t *cutnb, uint8_t *dom);

/* ---- a8-a11: StraightCutIntegrationRule (cutint/straightcutrule.cpp:582-677) */
/* Pass 1: counts.  elems (nullable = 0..nel-1).  offsets: int32[nel+1]
   (exclusive prefix sum, offsets[nel] = total points); nsub (nullable):
   uint8[nel] sub-simplex count (0 = uncut). */
int xg_cut_rule_count(xg_ctx *ctx, int D, int nel, const int32_t *elems,
                      const int32_t *el2v, const double *lset_v, int dt, int order,
                      int32_t *offsets, uint8_t *nsub);
/* Pass 2: points (x, y, z, w) per element at offsets[i] (double4 array of
   offsets[nel] entries).  Interface weights include |F^{-T} n|. */
int xg_cut_rule_fill(xg_ctx *ctx, int D, int nel, const int32_t *elems,
                     const int32_t *el2v, const double *coords, const double *lset_v,
                     int dt, int order, const int32_t *offsets, double *pts);

/* ---- quadrilaterals and hexahedra ---------------------------------------
   StraightCutIntegrationRule, quad / hex branch (cutint/straightcutrule.cpp:
   582-677) with LevelsetCutQuadrilateral (:206-501): tensor product along
   the last coordinate, dimension swaps, simplex fallback.  D = 2 (ET_QUAD)
   or 3 (ET_HEX); el2v: nel x 2^D vertices in the reference element's order
   ((0,0),(1,0),(1,1),(0,1), for hexes that square at z = 0 then at z = 1);
   policy: SWAP_DIMENSIONS_POLICY (0 FIRST_ALLOWED, 1 FIND_OPTIMAL,
   2 ALWAYS_NONE).  Points on the unit element; interface weights include
   |F^{-T} n_ref| of the multilinear element map at the point.  Cases where
   the reference throws ("if_scale_factor larger than bound", a NaN bound, a
   fallback simplex it cannot cut) -> XG_ERR_DEGENERATE at xg_sync. */
int xg_quad_rule_count(xg_ctx *ctx, int D, int nel, const int32_t *elems, const int32_t *el2v,
                       const double *lset_v, int dt, int order, int policy, int32_t *offsets);
int xg_quad_rule_fill(xg_ctx *ctx, int D, int nel, const int32_t *elems, const int32_t *el2v,
                      const double *coords, const double *lset_v, int dt, int order, int policy,
                      const int32_t *offsets, double *pts);
/* CutInformation::Update on a quad / hex mesh (xfem/cutinfo.cpp:61-133):
   order-0 part volumes of that rule (FIND_OPTIMAL in the reference's
   CutInfo, any policy here).  dt and bits required; outputs as xg_mark. */
int xg_quad_mark(xg_ctx *ctx, int D, int ne, const int32_t *el2v, const double *lset_v, int policy,
                 uint8_t *dt, double *cut_ratio, uint32_t *bits, int32_t *cut_list, int32_t *n_cut);
/* IntegrateX sums (cutint/python_cutint.cpp:66-98) on a quad / hex mesh:
   sum |det F(xi)| w f(x(xi)) with the multilinear element map; arguments as
   xg_integrate. */
int xg_quad_integrate(xg_ctx *ctx, int D, int nel, const int32_t *elems, const int32_t *el2v,
                      const double *coords, const int32_t *offsets, const double *pts,
                      int nterms, const double *coef, const int32_t *expo, double *per_elem,
                      double *total);

/* CalcElementMatrix (xfem/symboliccutbfi.cpp:76-304) of the Q1 space of a
   quad / hex mesh on a rule from xg_quad_rule_fill (offsets, pts): mask 1 mass,
   2 stiffness, 3 their sum; out: nel x 2^D x 2^D (FP64 for XG_PREC_FP64, FP32
   for XG_PREC_FP32), dofs in the element's vertex order. */
int xg_quad_elmat(xg_ctx *ctx, int D, int nel, const int32_t *elems, const int32_t *el2v,
                  const double *coords, const int32_t *offsets, const double *pts, int mask,
                  int precision, void *out);

/* element vectors of f = 1 (SymbolicCutLinearFormIntegrator, xfem/
   symboliccutlfi.cpp:57-140) for the Q1 space on such a rule; out: nel x 2^D
   FP64. */
int xg_quad_elvec(xg_ctx *ctx, int D, int nel, const int32_t *elems, const int32_t *el2v,
                  const double *coords, const int32_t *offsets, const double *pts, double *out);

/* ---- several level sets on simplices ------------------------------------
   CreateMultiLevelsetCutIntegrationRule (cutint/mlsetintegration.cpp:8-158)
   with StraightCutsIntegrationRule (cutint/straightcutrule.cpp:679-800).
   lsets: HOST array of M DEVICE pointers to P1 vertex values (1 <= M <= 8);
   dtts: HOST array of K x M domain types (1 <= K <= 64), the domain tuples
   of the LevelsetIntegrationDomain.  Per element, the rules of the tuples it
   is compatible with are concatenated in tuple order (uncut: the element's
   base rule; cut: the element cut by each cutting level set in turn, the
   base rule of the final sub-simplices, codimension 1 / 2 / 3 weights as the
   reference).  Rule layout and weights as xg_cut_rule_count / _fill.  More
   than 256 sub-simplices per element -> XG_ERR_CAPACITY at xg_sync. */
int xg_mlset_rule_count(xg_ctx *ctx, int D, int nel, const int32_t *elems, const int32_t *el2v, int M,
                        const double *const *lsets, int K, const int8_t *dtts, int order, int32_t *offsets);
int xg_mlset_rule_fill(xg_ctx *ctx, int D, int nel, const int32_t *elems, const int32_t *el2v,
                       const double *coords, int M, const double *const *lsets, int K, const int8_t *dtts,
                       int order, const int32_t *offsets, double *pts);
/* MultiLevelsetCutInformation (xfem/cutinfo.cpp:286-415): of_type
   (nullable, ne bits): the element's tuple of CheckIfStraightCut values is
   one of dtts (GetElementsOfType); contrib (nullable, ne bits): the order-0
   rule of dtts has a positive weight sum (GetElementsWithContribution; uses
   8 ne bytes of the context scratch). */
int xg_mlset_mark(xg_ctx *ctx, int D, int ne, const int32_t *el2v, const double *coords, int M,
                  const double *const *lsets, int K, const int8_t *dtts, uint32_t *of_type, uint32_t *contrib);

/* ---- a16: IntegrateX (cutint/python_cutint.cpp:21-110) ----------------- */
/* f = sum_t coef[t] x^expo[3t] y^expo[3t+1] z^expo[3t+2] (host arrays,
   nterms <= 32).  per_elem (nullable, device FP64[nel]); total: device FP64[1]
   (overwritten). */
int xg_integrate(xg_ctx *ctx, int D, int nel, const int32_t *elems, const int32_t *el2v,
                 const double *coords, const int32_t *offsets, const double *pts,
                 int nterms, const double *coef, const int32_t *expo, double *per_elem,
                 double *total);

/* ---- coefficient functions at the points of a rule --------------------
   NGSolve evaluates a CoefficientFunction at the mapped points of the rule
   (symboliccutbfi.cpp:229-287 scales the proxy products by it,
   symboliccutlfi.cpp:57-140 the test functions, python_cutint.cpp:66-98 the
   weights).  Here a coefficient is a postfix program over the physical point
   (host arrays, at most 192 code words, 48 constants, stack depth 16,
   checked by the entry point: XG_ERR_ARG otherwise).  Opcodes, with the
   immediate operand that follows some of them:
     1 CONST k  push consts[k]          2 COORD d  push x_d (physical)
     3 ADD  4 SUB  5 MUL  6 DIV  (a b -> a op b)   7 NEG
     8 POWI n  a -> a^n (integer n)      9 POW  a b -> pow(a, b)
     10 SQRT 11 SIN 12 COS 13 TAN 14 EXP 15 LOG 16 ATAN 17 TANH 18 ABS
     19 GF1 s  push the P1 grid function gf[s] (device FP64 vertex values,
               nv of them) at the point: sum_k lam_k gf[s][v_k]
     20 IFPOS  a b c -> (a > 0 ? b : c)   21 MAX  22 MIN  (a b -> ...)
     23 EL0 s  push gf[s][e], a piecewise constant grid function (one value
               per element of the mesh, e the element's number; not for
               xg_p1_interpolate_cf)
   The program must leave exactly one value. */
typedef struct xg_cf {
  int32_t ncode;
  const int32_t *code;
  int32_t nconst;
  const double *consts;
  int32_t ngf;                /* <= 4 */
  const double *const *gf;    /* host array of device pointers */
} xg_cf;

/* w_q <- w_q * cf(x_q) for the points of a straight cut / several-level-set
   rule (layout of xg_cut_rule_fill, in place); coords holds A = D or D + 1
   values per vertex (D + 1: boundary elements given in the ambient space,
   x_q = sum_k lam_k X_k).  The element matrices
   (xg_elmat), element vectors (xg_elvec) and integrals (xg_integrate with
   f = 1) of the scaled rule are those of c u v / c grad u . grad v, f v and
   f.  No counterpart in the reference's interface: the reference applies the
   coefficient inside CalcElementMatrix / CalcElementVector /
   IntegrateX; this splits it out so the three kernels can share it. */
int xg_rule_scale_cf(xg_ctx *ctx, int D, int A, int nel, const int32_t *elems, const int32_t *el2v,
                     const double *coords, const int32_t *offsets, double *pts, const xg_cf *cf);

/* Element-boundary rules (SymbolicCutBilinearFormIntegrator::
   T_CalcElementMatrixEBAdd, xfem/symboliccutbfi.cpp:488-640): given the cut
   rules of the facets of nel elements (D = 2, 3), computed as rules of a
   (D-1)-dimensional mesh whose element i (D + 1) + k is facet k of element i
   (its vertices without vertex k, ascending; foffsets: nel (D + 1) + 1
   entries, weights in that mesh's reference measure), maps the points in
   place into the element's reference coordinates and rescales the weights to
   the element's measure (w |det F_facet| / |det F_element|); eoffsets
   (nel + 1) = foffsets[i (D + 1)].  xg_elmat / xg_elvec / xg_integrate then
   give the element-boundary integrals with the volume basis. */
int xg_facet_rule_to_element(xg_ctx *ctx, int D, int nel, const int32_t *elems, const int32_t *el2v,
                             const double *coords, const int32_t *foffsets, double *pts, int32_t *eoffsets);

/* ---- linear forms on a materialised rule: SymbolicCutLinearFormIntegrator::
   T_CalcElementVector (xfem/symboliccutlfi.cpp:57-140) with f = 1 in the
   weights: out[i][a] = |det F| sum_q w_q phi_a(xi_q), FP64 nel x ndof. */
int xg_elvec(xg_ctx *ctx, int D, int p, int nel, const int32_t *elems, const int32_t *el2v,
             const double *coords, const int32_t *offsets, const double *pts, double *out);

/* ---- a14: SymbolicCutBilinearFormIntegrator::CalcElementMatrix ---------
   (xfem/symboliccutbfi.cpp:37-46, 76-304) on a materialised rule.
   out: nel x ndof x ndof (FP64 for XG_PREC_FP64, FP32 otherwise), row = test
   dof, overwritten (CalcElementMatrix semantics). */
int xg_elmat(xg_ctx *ctx, int D, int p, int nel, const int32_t *elems, const int32_t *el2v,
             const double *coords, const int32_t *offsets, const double *pts, int mask,
             int precision, void *out);

/* ---- fused: rule + element matrix, points never leave the chip ---------
   For each element i and each dts[t] (host array, ndt <= 3): the straight cut
   rule of order `order` and CalcElementMatrix on it.
   out: nel x ndt x ndof x ndof (FP64 / FP32 by precision).
   precision | XG_METHOD_MOMENTS selects the moment form (any rule order). */
int xg_cut_pipeline(xg_ctx *ctx, int D, int p, int order, int nel, const int32_t *elems,
                    const int32_t *el2v, const double *coords, const double *lset_v,
                    int ndt, const int *dts, int mask, int precision, void *out);

/* ---- linear forms: SymbolicCutLinearFormIntegrator::T_CalcElementVector
   (xfem/symboliccutlfi.cpp:57-140) for the source term f = 1 on the straight
   cut rule, moment form (FP64 arithmetic): out[i][t][a] = sum_q w_q |det F|
   phi_a(x_q).  out: nel x ndt x ndof (FP32 / FP64 by precision). */
int xg_cut_elvec(xg_ctx *ctx, int D, int p, int order, int nel, const int32_t *elems, const int32_t *el2v,
                 const double *coords, const double *lset_v, int ndt, const int *dts, int precision,
                 void *out);

/* ---- global assembly: the scatter of NGSolve's BilinearForm::Assemble /
   LinearForm::Assemble (the callers of CalcElementMatrix / CalcElementVector)
   on the device.  dofs: nel x ndof global dof numbers (< 0: skipped); CSR
   pattern row_ptr / col_idx (columns sorted per row) must contain every
   (dofs[e][a], dofs[e][b]); values / out: FP64, accumulated with atomic adds
   (summation order not fixed: results agree to rounding, not bitwise).  A
   missing pattern entry is reported by xg_sync (XG_ERR_CAPACITY). */
int xg_assemble_csr(xg_ctx *ctx, int nel, int ndof, const int32_t *dofs, const void *elmat, int precision,
                    const int32_t *row_ptr, const int32_t *col_idx, double *values);
/* the same scatter with every element-matrix entry multiplied by scale
   (the coefficient of the form, e.g. 2 * u * v * dCut(...)) */
int xg_assemble_csr_scaled(xg_ctx *ctx, int nel, int ndof, const int32_t *dofs, const void *elmat, int precision,
                           double scale, const int32_t *row_ptr, const int32_t *col_idx, double *values);
/* Sparsity pattern of the element dof couplings on the device (the matrix
   graph NGSolve's BilinearForm::Assemble builds before adding the element
   matrices): every (dofs[e][a], dofs[e][b]) of the nel x ndof table (entries
   < 0 skipped), sorted and made unique.  row_ptr: int32[n + 1]; col_idx:
   int32[cap] (columns ascending per row); nnz: device int32[1].  Uses the
   context workspace (about 32 B per coupling); nnz > cap -> XG_ERR_CAPACITY
   and a dof >= n -> XG_ERR_ARG at xg_sync.  nel * ndof^2 < 2^31. */
int xg_csr_pattern(xg_ctx *ctx, int nel, int ndof, const int32_t *dofs, int n, int32_t *row_ptr,
                   int32_t *col_idx, long long cap, int32_t *nnz);
int xg_assemble_vec(xg_ctx *ctx, int nel, int ndof, const int32_t *dofs, const void *elvec, int precision,
                    double *out);

/* ---- a17: subdivision rule for non-P1 level sets ----------------------
   (cutint/xintegration.cpp:1626-1754 CutIntegrationRule with subdivlvl = L,
   NumericalIntegrationStrategy :617-1063, CutSimplex<3>::MakeQuad :1084-1388)
   on tetrahedra whose level set is given by its Lagrange nodal values of
   degree lset_order (1 or 2; lset_nodal: nel x 4 or nel x 10, this
   library's dof order).  The rule (thousands of points at L = 4) is
   contracted on the fly into fixed-node weights
       W_k = sum_points w psi_k(x),  psi = degree-q Lagrange basis (q == 4),
   which integrate every polynomial of degree <= q exactly like the rule.
   out_neg / out_if (nullable): nel x 35 FP64; finv (nullable): nel x 9
   element Jacobian inverses for the interface weights (identity if null);
   dom (nullable): uint8[nel] element domain (NEG/POS/IF).  L <= 5. */
int xg_subdiv_weights(xg_ctx *ctx, int nel, const double *lset_nodal, int lset_order, int L,
                      int order, int q, const double *finv, double *out_neg, double *out_if,
                      uint8_t *dom);
/* The same subdivision rule contracted into FP64 barycentric moments
   m_alpha = sum_points w lam(x)^alpha, |alpha| = 2p (the monomial order of
   the moment identities; 10 / 35 / 84 values per element for p = 1 / 2 / 3):
   every element matrix of degree-p elements on the rule is a linear function
   of them (xg_elmat_moments).  out_neg / out_if (nullable): nel x NM FP64. */
int xg_subdiv_moments(xg_ctx *ctx, int nel, const double *lset_nodal, int lset_order, int L,
                      int order, int p, const double *finv, double *out_neg, double *out_if,
                      uint8_t *dom);

/* The same subdivision rule contracted into FP64 barycentric moments of
   degree 2p (sum_points w lam(x)^alpha, as xg_subdiv_moments) for listed
   mesh elements and any of NEG / POS / IF (dts, host array, ndt <= 3):
   moments: nel x ndt x NM FP64, element-major.  Arguments, order >= 2p and
   stats as xg_subdiv_pipeline.  POS keeps the zero-valued side as the
   reference's leaves do (xintegration.cpp:1159-1164). */
int xg_subdiv_moments_mesh(xg_ctx *ctx, int p, int L, int order, int nel, const int32_t *elems,
                           const int32_t *el2v, const double *coords, const double *lset_nodal, int ndt,
                           const int *dts, double *moments, int32_t *stats);
/* The same rule as fixed-node weights on the degree-4 Lagrange lattice of
   each element (as xg_subdiv_weights, q = 4: they integrate every
   polynomial of degree <= 4 exactly like the rule) for listed mesh elements
   and NEG / POS / IF: weights nel x ndt x 35 FP64.  order >= 4; arguments as
   xg_subdiv_moments_mesh. */
int xg_subdiv_weights_mesh(xg_ctx *ctx, int L, int order, int nel, const int32_t *elems,
                           const int32_t *el2v, const double *coords, const double *lset_nodal, int ndt,
                           const int *dts, double *weights, int32_t *stats);

/* Fused subdivision rule + element matrices on a tetrahedral mesh with a P2
   level set: dCut(lset_p2, dt, subdivlvl=L) forms assembled element by
   element (CutIntegrationRule, xintegration.cpp:1626-1754, with
   NumericalIntegrationStrategy :617-1063 and CutSimplex<3>::MakeQuad
   :1084-1388; CalcElementMatrix, symboliccutbfi.cpp:76-304).  For each listed
   element elems[i] (required) and each dts[t] (NEG, POS or IF; host array,
   ndt <= 3) the level-L rule is contracted into FP64 degree-2p moments and
   the degree-p mass (mask 1) / stiffness (mask 2) matrix formed from them.
   lset_nodal: ne x 10 FP64 nodal values indexed by element id (Lattice<3,2>
   dof order); el2v / coords: the mesh (interface weights and the element
   map).  order >= 2p (the reference's intorder = 2p, symboliccutbfi.cpp:
   107-111: the base rules are then exact for the moments, uncut tree nodes
   come from a table).  out: nel x ndt x ndof x ndof, FP32 or FP64 by
   precision.  stats (nullable): int32 nel x ndt x 4 -- table rows, volume
   tets and interface triangles integrated point by point, cut leaves.
   Uses the context workspace (moments of up to 2^21 elements at a time).
   With order >= 2p, xg_subdiv_moments, xg_subdiv_weights and the three mesh
   entry points run in rounds of up to 2^19 elements over per-element node
   lists (8 KB per element in a grow-only context buffer, up to 4 GB; the
   elements whose lists overflow take the one-kernel path, whose launch reads
   their count from device memory).  The calls stay asynchronous; growing the
   buffer allocates, so only a call whose buffer is already large enough could
   be captured into a HIP graph (not tested).  Tuning knobs
   read from the environment at each call: XG_SDX_CAP (list capacity, 16 ..
   4096, default 4096), XG_SDX_QUEUE=0 (one block per element instead of
   persistent waves), XG_SDX_NOFUSE=1 (p = 3: volume side and interface in
   separate launches). */
int xg_subdiv_pipeline(xg_ctx *ctx, int p, int L, int order, int nel, const int32_t *elems,
                       const int32_t *el2v, const double *coords, const double *lset_nodal, int ndt,
                       const int *dts, int mask, int precision, void *out, int32_t *stats);

/* Integral of a polynomial (as xg_integrate) on fixed-node rules: weights
   (nel x 35 FP64, from xg_subdiv_weights with q = 4) at the nodes of the
   degree-4 Lagrange lattice of each listed tetrahedron, mapped to the
   physical element: sum_k |det F| W_k f(x_k), equal to the subdivision
   rule's point sum for polynomials of degree <= 4 (IntegrateX with a P2
   level set, python_cutint.cpp:21-110).  per_elem (nullable): FP64[nel];
   total: FP64[1] (overwritten). */
int xg_integrate_nodes(xg_ctx *ctx, int nel, const int32_t *elems, const int32_t *el2v, const double *coords,
                       int q, const double *weights, int nterms, const double *coef, const int32_t *expo,
                       double *per_elem, double *total);

/* ---- element matrices on a moment-form rule ----------------------------
   (symboliccutbfi.cpp:76-304 on any rule, given by its degree-2p barycentric
   moments, e.g. from xg_subdiv_moments): the mass / stiffness matrices of
   degree-p H1 elements from the exact polynomial identities of the moment
   form, FP64 arithmetic.  coords == NULL: reference element (identity map).
   moments: nel x NM FP64; out: nel x ndof x ndof (FP32 / FP64). */
int xg_elmat_moments(xg_ctx *ctx, int D, int p, int nel, const int32_t *elems, const int32_t *el2v,
                     const double *coords, const double *moments, int mask, int precision, void *out);

/* ---- element matrices on a fixed-node rule ----------------------------
   (symboliccutbfi.cpp:76-304 on the rule {(xi_k, W_k)}: xi_k = the 35 nodes of
   the degree-4 Lagrange lattice of the reference tet, W_k per element from
   xg_subdiv_weights or the learned rule).  coords == NULL: reference tets
   (identity map).  p <= 2 (P2 mass + stiffness integrands have degree <= 4).
   out: nel x ndof x ndof. */
int xg_elmat_weights(xg_ctx *ctx, int p, int q, int nel, const int32_t *elems, const int32_t *el2v,
                     const double *coords, const double *weights, int mask, int precision, void *out);

/* ---- batched multi-mesh pipeline ---------------------------------------
   No single reference counterpart: many independent meshes (each one
   CutInformation::Update + BilinearForm::Assemble of a dCut form in the
   reference) concatenated into one device-side sequence -- marking and
   compaction over all elements, a per-mesh plan, and one persistent fused
   rule + element-matrix kernel per H1 degree with dynamic work distribution,
   captured once into a hipGraph and replayed.
   Mesh m owns elements [elem_off[m], elem_off[m+1]) of the concatenated
   el2v (global vertex ids) and uses degree mesh_p[m] (1..3) with rule order
   2 p (host arrays, copied).  el2v / coords / lset are device arrays that must
   stay valid; D == 3, nmesh <= 1024.  Output (FP32): per mesh, its cut
   elements in ascending order, ndt x ndof(p)^2 floats each, at out_off[m]. */
typedef struct xg_batch xg_batch;
int xg_batch_create(xg_ctx *ctx, int D, int nmesh, const int32_t *elem_off, const int32_t *mesh_p, int ndt,
                    const int *dts, int mask, const int32_t *el2v, const double *coords, const double *lset,
                    xg_batch **out);
int xg_batch_destroy(xg_batch *b);
/* launch shape of the degree-p kernel: waves per block (1|2), resident
   blocks per CU, elements per work grab (per wave for the MFMA kernels, per
   lane for the moment form) */
int xg_batch_tune(xg_batch *b, int p, int waves_per_block, int blocks_per_cu, int grab);
/* element-matrix method of the degree-p kernel (p = 0: all degrees):
   0 = MFMA contraction (default), XG_METHOD_MOMENTS = moment form (FP64
   arithmetic, FP32 store) */
int xg_batch_method(xg_batch *b, int p, int method);
/* marking + plan only; then xg_batch_info */
int xg_batch_plan(xg_ctx *ctx, xg_batch *b);
/* synchronises; host outputs (nullable): cut elements per mesh (nmesh),
   output offsets in floats (nmesh + 1) */
int xg_batch_info(xg_ctx *ctx, xg_batch *b, int32_t *cut_per_mesh, long long *out_off);
/* device array of the cut element ids of the last plan / run */
const int32_t *xg_batch_cut_list(xg_batch *b);
/* full batch into out (cap floats; overflow -> XG_ERR_CAPACITY at xg_sync).
   use_graph != 0: capture once per (out, cap, tuning) and replay.
   only_p != 0: run the kernel of that degree only (eager; for tuning). */
int xg_batch_run(xg_ctx *ctx, xg_batch *b, float *out, long long cap, int use_graph, int only_p);

/* ---- learned rule: MLP I -> H -> H -> O (tanh), FP32 parameters --------
   No reference counterpart: the rebuild's learned quadrature (predicts the
   fixed-node weights of the NEG part of a P2-cut tet from its normalised
   nodal values; trained on xg_subdiv_weights).
   params layout: W1[H][I], b1[H], W2[H][H], b2[H], W3[O][H], b3[O]. */
#define XG_PREC_BF16 3   /* plain BF16 MFMA, FP32 accumulation */
int xg_mlp_nparams(int I, int H, int O);
/* precision XG_PREC_FP32 (VALU; work: 2*B*H floats) or XG_PREC_BF16 (fused
   MFMA kernel, I = 10, H = 128, O <= 96; work unused). */
int xg_mlp_forward(xg_ctx *ctx, int I, int H, int O, const float *params, int B, const float *x,
                   float *y, float *work, int precision);
/* One Adam step (FP32 forward + backward); loss: device float[1]; grad
   (nullable) receives the gradient; work: 4*B*H + 2*B*O floats.  lr <= 0
   skips the update (gradient only). */
int xg_mlp_train_step(xg_ctx *ctx, int I, int H, int O, float *params, float *adam_m, float *adam_v,
                      int step, float lr, int B, const float *x, const float *target, float *loss,
                      float *grad, float *work);
/* Moment-residual check: rel[e] = |scale*pred - ref|_2 / |ref|_2 (FP64);
   accept[e] = rel <= thr; out = accepted ? scale*pred : ref (fallback to the
   plain rule).  pred: FP32 nel x nq, ref: FP64 nel x nq; outputs nullable. */
int xg_moment_residual(xg_ctx *ctx, int nel, int nq, const float *pred, double scale, const double *ref,
                       double thr, double *rel, uint8_t *accept, double *out);

/* Learned rule as a correction of a cheap rule, checked without the plain
   rule of full degree.  No reference counterpart (the rebuild's learned
   quadrature).  out = base + corr: base = the degree-2p moments of a cheap
   plain rule (the subdivision rule at a low level), corr = the MLP's FP32
   correction (nel x NM); the check lowers out to degree-2 moments through
   sum lam = 1 and compares them with ref2 (nel x 10 FP64, the exact degree-2
   moments of the full plain rule, xg_subdiv_moments with p = 1, a fraction of
   its cost): rel[e] = |lowered - ref2| / |ref2| (nullable), accept[e] =
   rel <= thr.  Rejected elements fall back to the plain rule (the caller
   computes them).  p = 1..3. */
int xg_learned_moments(xg_ctx *ctx, int nel, int p, const double *base, const float *corr,
                       const double *ref2, double thr, double *out, double *rel, uint8_t *accept);

/* Per-kernel timing of this context (a measurement aid; no counterpart in
   the reference).  While enabled, the subdivision kernels record a HIP event
   pair on the context's stream around each launch, named "sdx_tree",
   "sdx_leaves_vol", "sdx_leaves_if", "sdx_rows" and "subdiv_exact" (the
   one-kernel path).  xg_timing_enable(ctx, 1) clears the records and starts
   recording, 0 stops; xg_timing_read synchronises the stream and returns the
   summed milliseconds and the launch count recorded under `name` (0 and 0
   for a name never recorded). */
int xg_timing_enable(xg_ctx *ctx, int on);
int xg_timing_read(xg_ctx *ctx, const char *name, double *total_ms, int *launches);

#ifdef __cplusplus
}
#endif
#endif
