/*
 * xg_oracle.h -- CPU FP64 restatement of ngsxfem's cut-element hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * product path.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it; the product (ngsxfem_amd/) never does.
 *
 * Every function states the reference file:line it restates.  NGSolve is not
 * available, so the pieces NGSolve supplies to the reference (base simplex
 * quadrature rules, the H1 basis, element maps) are the rebuild's own choices:
 *   - base rules: Stroud conical-product Gauss-Jacobi rules, n = order/2+1
 *     points per collapsed direction (exact to degree 2n-1 >= order);
 *   - H1 basis: equispaced Lagrange basis of degree p on the simplex;
 *   - element map: x = V_D + sum_m xi_m (V_{m} - V_D) with the NGSolve
 *     reference vertices (1,0,0),(0,1,0),(0,0,1),(0,0,0).
 * These choices are pinned by exactness / closed-form tests, not by an
 * import of the reference (which needs NGSolve and cannot be built here).
 *
 * Floating point: compiled with -ffp-contract=off so that every operation
 * is a single IEEE FP64 rounding in source order; the HIP kernels follow the
 * same order, which makes marking and sub-simplex counts bit-exact.
 */
#ifndef XG_ORACLE_H
#define XG_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define OR_NEG 0
#define OR_POS 1
#define OR_IF 2

#define OR_ERR_NONE 0
#define OR_ERR_BAD_SCUT (-2)   /* "Bad length of s_cut!" straightcutrule.cpp:148 */
#define OR_ERR_NOT_IMPL (-3)   /* "Cutting this part of a tetraeder is not implemented yet!" :193 */
#define OR_ERR_CAP (-4)        /* output capacity exceeded */
#define OR_ERR_SUBDIV (-5)     /* "did not expect this.. -3-" xintegration.cpp:1386 */
#define OR_ERR_BOUND (-6)      /* "if_scale_factor larger than bound" (straightcutrule.cpp:286),
                                  "Sufficient Criterion calculated nan Bound!" (:438) */
#define OR_ERR_PLAIN (-7)      /* "You tried to cut a simplex with a plain geometry lset function" (:45-49) */

/* ---- base rules ------------------------------------------------------ */
int or_rule_npts(int D, int order);
/* xi: npts x D (reference simplex coordinates), w: npts. returns npts */
int or_rule(int D, int order, double *xi, double *w);

/* ---- geometry ---------------------------------------------------------- */
/* V: (D+1) x D physical vertex coordinates; F (DxD row-major) = dx/dxi;
   Finv = F^{-1}; absdet = |det F| */
void or_jacobian(int D, const double *V, double *F, double *Finv, double *absdet);

/* ---- straight cut rule (straightcutrule.cpp:11-204, 503-677) ---------- */
int or_check_if_straight_cut(int n, const double *lset);
/* Rule on the reference simplex.  Points are written as xyz triples (z=0 /
   y=0 padding in 2D / 1D).  Finv (DxD) is only used for dt==IF.
   returns the number of points, -1 for "no rule" (nullptr in the
   reference), or a negative OR_ERR_* code. *nsub = number of sub-simplices
   of the decomposition (0 when the element is not cut). */
int or_straight_cut_rule(int D, const double *lset, int dt, int order,
                         const double *Finv, double *pts, double *w, int cap,
                         int *nsub);

/* ---- multiple level sets (mlsetintegration.cpp:8-158,
   straightcutrule.cpp:679-800, cutinfo.cpp:286-415) --------------------- */
/* One element: lsets M x (D+1) vertex values, dtts K x M domain tuples.
   Points (xyz triples, reference coordinates) and weights (physical weight
   = w * absdet) appended tuple by tuple.  Returns the number of points or a
   negative OR_ERR_* code. */
int or_mlset_rule(int D, int M, const double *lsets, int K, const int *dtts, int order, const double *F,
                  const double *Finv, double absdet, double *pts, double *w, int cap);
/* of_type[e] = the element's tuple of CheckIfStraightCut values is in dtts;
   contrib[e] (nullable) = the order-0 rule has a positive weight sum.
   lset_v: M x nv. */
int or_mlset_mark(int D, int ne, const int32_t *el2v, const double *coords, int M, const double *lset_v, int nv,
                  int K, const int *dtts, uint8_t *of_type, uint8_t *contrib);

/* ---- marking (cutinfo.cpp:61-133) -------------------------------------- */
/* bits: 8 arrays (indexed by COMBINED_DOMAIN_TYPE) of nwords=ceil(ne/32)
   uint32 words each, LSB-first.  returns 0 or an OR_ERR code. */
int or_mark(int D, int ne, const int32_t *el2v, const double *lset_v,
            uint8_t *dt_out, double *cut_ratio, uint32_t *bits);

/* ---- quadrilaterals / hexahedra (straightcutrule.cpp:74-129, 206-501) -- */
/* vals: 2^D vertex values in the reference element's vertex order (ET_QUAD:
   (0,0),(1,0),(1,1),(0,1); ET_HEX: the same square at z=0, then at z=1);
   pol: 0 FIRST_ALLOWED, 1 FIND_OPTIMAL, 2 ALWAYS_NONE.  V (nullable): 2^D x D
   physical vertex coordinates, used for the interface weights
   |F^{-T} n_ref| (identity map when NULL).  Points as xyz triples on the
   unit quad / hex.  returns the number of points, -1 for no rule, or an
   OR_ERR code. */
int or_quad_cut_rule(int D, const double *vals, int dt, int order, int pol, const double *V, double *pts,
                     double *w, int cap);
/* CutInformation::Update on a quad / hex mesh; outputs as or_mark */
int or_quad_mark(int D, int ne, const int32_t *el2v, const double *lset_v, int pol, uint8_t *dt_out,
                 double *cut_ratio, uint32_t *bits);
double or_quad_integrate_poly(int D, const double *V, const double *pts, const double *w, int n, int nterms,
                              const double *coef, const int32_t *expo);
/* Q1 element matrix (mass: mask 1, stiffness: 2, sum: 3) on a rule of the
   unit element; out: 2^D x 2^D */
void or_quad_elmat(int D, const double *V, const double *pts, const double *w, int n, int mask, double *out);

/* ---- H1 Lagrange basis ------------------------------------------------- */
int or_h1_ndof(int D, int p);
/* phi: ndof, dphi: ndof x D (reference gradients), either may be NULL */
void or_h1_shape(int D, int p, const double *xi, double *phi, double *dphi);
/* lattice node (reference coordinates, D values) of dof k */
void or_h1_node(int D, int p, int k, double *xi);

/* ---- element matrix (symboliccutbfi.cpp:76-304) ------------------------ */
/* mask bit0: mass u*v, bit1: stiffness grad(u)*grad(v).  out: ndof x ndof
   (row = test dof, col = trial dof), accumulated (+=). */
void or_elmat_add(int D, int p, const double *V, const double *pts,
                  const double *w, int n, int mask, double *out);

/* ---- integrate (python_cutint.cpp:66-98) ------------------------------- */
/* f = sum_t coef[t] * x^e[3t] y^e[3t+1] z^e[3t+2] in physical coords */
double or_integrate_poly(int D, const double *V, const double *pts,
                         const double *w, int n, int nterms,
                         const double *coef, const int *expo);

/* IntegrateX over a mesh: sum over elements of the straight cut rule of
   domain dt applied to the polynomial f (python_cutint.cpp:21-110).
   per_elem (nullable) receives the element contributions. */
int or_integrate_mesh(int D, int ne, const int32_t *el2v, const double *coords,
                      const double *lset_v, int dt, int order, int nterms,
                      const double *coef, const int *expo, double *per_elem,
                      double *total, int nthreads);

/* ---- fused element pipeline, threaded (CPU baseline) ------------------- */
/* For each listed element: straight cut rule for each dt in dts (NEG=0,
   POS=1, IF=2) and the element matrix on it.  out: nel x ndt x ndof^2.
   coords: nv x D.  returns 0 or error. */
int or_cut_pipeline(int D, int p, int order, int nel, const int32_t *elems,
                    const int32_t *el2v, const double *coords,
                    const double *lset_v, int ndt, const int *dts, int mask,
                    double *out, int nthreads);

/* ---- subdivision rule (xintegration.cpp:617-1063, 1084-1388, 1626-1754) */
/* Tetrahedron with a level set given by its Lagrange nodal values of
   degree lset_order (1 or 2).  L = subdivision levels, order = rule order.
   Finv: element Jacobian inverse for the interface weights (may be NULL,
   then the identity is used).  For each r in {NEG,POS,IF}: pts[r] (cap x 3),
   w[r] (cap), n[r] out.  returns the element domain (OR_NEG/OR_POS/OR_IF)
   or an OR_ERR code. */
int or_subdiv_rule3(const double *lset_nodal, int lset_order, int L, int order,
                    const double *Finv, double *pts[3], double *w[3], int cap,
                    int n[3]);
/* Same traversal, but only accumulate the fixed-node weights
   W_k = sum_points w * psi_k(x) of the degree-q Lagrange basis psi on the
   reference tet, for the NEG part (out_neg) and the interface (out_if,
   interface weights as above).  Avoids materialising large rules.
   returns the element domain or error. */
/* Order-0 subdivision rule on the reference triangle (boundary elements of
   a tetrahedral mesh with a P2 level set, xfem/cutinfo.cpp:80-124 with
   cutint/xintegration.cpp:620-838,1019-1061,1391-1545): part[0] / part[1]
   = NEG / POS areas; returns the lattice domain (0/1/2) or < 0 on error. */
int or_subdiv_trig_parts(const double *lset_nodal, int lset_order, int L, double part[2]);
int or_subdiv_moments3(const double *lset_nodal, int lset_order, int L,
                       int order, int q, const double *Finv, double *out_neg,
                       double *out_if);
/* Batched, threaded version of or_subdiv_moments3 over nel elements
   (lset_nodal: nel x 10 for P2).  returns 0 or error. */
int or_subdiv_moments3_batch(int nel, const double *lset_nodal, int lset_order,
                             int L, int order, int q, double *out_neg,
                             double *out_if, int nthreads);

/* Rule + element matrix on tetrahedral meshes with a P2 level set
   (lset_nodal: ne x 10 nodal values, indexed by element id): for each listed
   element (elems nullable) the level-L subdivision rule of every dts[t] and
   the degree-p mass/stiffness matrix on it (physical map from coords/el2v).
   out: nel x ndt x ndof^2.  returns 0 or an OR_ERR code. */
int or_subdiv_pipeline(int p, int L, int order, int nel, const int32_t *elems, const int32_t *el2v,
                       const double *coords, const double *lset_nodal, int ndt, const int *dts, int mask,
                       double *out, int nthreads);

/* ---- learned rule: MLP 10 -> H -> H -> O with tanh (FP64) -------------- */
/* params layout: W1[H][I], b1[H], W2[H][H], b2[H], W3[O][H], b3[O] */
int or_mlp_nparams(int I, int H, int O);
void or_mlp_forward(int I, int H, int O, const double *params, int batch,
                    const double *x, double *y);
/* loss = 0.5/batch * sum (y - target)^2 ; grad: same layout as params.
   returns the loss. */
double or_mlp_loss_grad(int I, int H, int O, const double *params, int batch,
                        const double *x, const double *target, double *grad);

#ifdef __cplusplus
}
#endif
#endif
