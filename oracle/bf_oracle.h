/*
 * bf_oracle.h -- CPU oracle for the active-strain element internal-force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, FP64, operation-by-operation
 * restatement of BioFEM-IA's FInternalAct path (src/active_strain.c, src/math.c)
 * and of the topology builders that feed it (src/bending.c IrrVer/Patch/
 * GlobalGhostInit/GlobalGhost, src/io.c AreaNormal).  It is the parity checker
 * for the HIP library and the timed CPU baseline.  Only tests/, the smoke entry
 * point and bench.py's cpu_baseline leg may load it; the product path never does.
 *
 * Parity status of each function is listed in the header comment of
 * bf_oracle.c (what pins it: golden vectors, closed forms, invariants).
 */
#ifndef BF_ORACLE_H
#define BF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define BFO_SENTINEL 1000000 /* empty patch slot, src/io.c:306, src/bending.c:1292 */

/* Status codes (first failing element is reported alongside). */
enum {
  BFO_OK = 0,
  BFO_ERR_ARG = 1,         /* bad argument (PETSC_ERR_ARG_*) */
  BFO_ERR_STENCIL = 2,     /* patch slot == sentinel (active_strain.c:498, :577) */
  BFO_ERR_IRE = 3,         /* unknown ire value (active_strain.c:721) */
  BFO_ERR_NEN = 4,         /* nen above the patch stride / v+12 > 64 (active_strain.c:599) */
  BFO_ERR_NQP = 5,         /* n_qp outside 1..5 (active_strain.c:224) */
  BFO_WARN_CC3333 = 6      /* |CC3333| <= 1e-14 or NaN (active_strain.c:1674-1681); non-fatal,
                              the reference ignores ModElemC33's return code (:2086) */
};

typedef struct {
  double x, y, z;
} bfo_vec3;

/* Mirror of SubdivGeomQP (include/act_variables.h:36-52) with the IN* arrays inline. */
typedef struct {
  int nen, v, is_irregular;
  bfo_vec3 ndx21, ndx31, nn, gc1, gc2, Aaa, Abb, Aab;
  double INa0[64], INa1[64], INab0[64], INab1[64], INab2[64];
} bfo_geom;

/* The subset of IBMNodes (include/variables.h) that the force path reads. */
typedef struct {
  int n_v;          /* real nodes */
  int n_nodes;      /* real + ghost nodes (coordinate array length) */
  int n_elmt;       /* real elements */
  int n_elmt_base;  /* > 0 enables the LV apex-cap CST branch for ec >= n_elmt_base */
  int patch_stride; /* 16 in the reference layout (src/io.c:306) */
  const int *patch; /* [patch_stride * n_elmt] */
  const int *ire;   /* [n_elmt] 0 regular, 1 irregular */
  const int *val;   /* [n_elmt] valence (6 for regular) */
  const int *nv1, *nv2, *nv3; /* [n_elmt] corner nodes (cap branch only; may be NULL otherwise) */
  const double *x0, *y0, *z0; /* reference coordinates [n_nodes] */
  const double *dA0;          /* [n_elmt] reference flat-triangle areas */
  const double *n_fib;        /* [3 * n_elmt] fibre directions (Cmpnts layout) */
  const double *gamma_scale;  /* [n_elmt] */
  const double *activation_delay; /* [n_elmt] or NULL (= 0) */
} bfo_mesh;

/* Globals and ActData/MuscleActParams fields the path reads (src/active_strain.c:47-58,
 * include/act_variables.h:78-105). */
typedef struct {
  int n_qp;            /* -num_gaussian_quad_points */
  int c33_iters;       /* -C33_subitr_nums */
  int cart_fib_act;    /* -cart_fib_act */
  int manufactured;    /* global `manufactured` */
  double h0, mu, K;    /* thickness, shear and bulk modulus */
  double gamma;        /* -muscle_act_gamma */
  int ramp_type;       /* 0..3 as in the reference; 4 = Gaussian pulse (new, defined here) */
  double gamma_T, tau_up, tau_down, apd;
  double pulse_t0, pulse_sigma; /* ramp type 4: gamma * exp(-0.5 ((t - t0)/sigma)^2) */
  double manufactured_gamma0, manufactured_T;
  int ti, tisteps;     /* time-step index and count */
  double dt;
} bfo_params;

typedef struct {
  int status;     /* first error code (BFO_*) */
  int elem;       /* element that raised it, -1 if none */
  int n_warn;     /* count of non-fatal CC3333 warnings */
} bfo_status;

/* ---- elementary pieces, exposed for unit tests ---- */
int bfo_gauss(int n_qp, double *theta, double *w);                                /* :174-230 */
double bfo_gamma_of_time(const bfo_params *p, double delay);                      /* :1156-1179 */
void bfo_irregular_tables(int v, double *INa0, double *INa1, double *INab0,
                          double *INab1, double *INab2);                          /* :594-666 */
int bfo_geom_subdiv(const bfo_mesh *m, int ec, const double *xb, const double *yb,
                    const double *zb, bfo_geom *G);                                /* :428-724 */

/* Per-element pipeline through the C33 solve (ElemUpdateG, ElemActDefGrad, ElemCGDefTens,
 * ElemC33Solve) followed by ElemUpdFint.  Fb receives 3*nen entries (zeroed here, as
 * PetscCalloc1 does at :2249).  Optional outputs: S (n_qp*9, total stress contravariant
 * after the last ElemTotStress) and C33 (n_qp, C.Cov[2][2] at the end). */
/* Fb must hold 3 * 64 doubles. */
int bfo_element_force(const bfo_mesh *m, const bfo_params *p, const bfo_geom *G0,
                      const bfo_geom *Gc, int ec, double *Fb, double *S_out, double *C33_out,
                      int *n_warn);

/* Whole-path entry: the equivalent of FInternalAct (active_strain.c:2226-2267) on one rank.
 * geom0 is built from x0 (FEM_DMPlexGeomSetup, dmplex_geom.c:859-865), the current geometry
 * from (x, y, z).  fint is interleaved [3 * n_nodes]; if accumulate == 0 it is zeroed first
 * (the caller's VecSet at main.c:829), else contributions are added.  Ghost rows (node >= n_v)
 * receive nothing (:2254).  c33 (nullable) receives n_elmt * n_qp final C33 values. */
int bfo_finternal_act(const bfo_mesh *m, const bfo_params *p, const double *x, const double *y,
                      const double *z, double *fint, int accumulate, double *c33,
                      bfo_status *st);

/* geom0 cache (dmplex_geom.c:859-865) and evaluation against it, so that timing
 * excludes the one-time setup as in the reference's per-residual cost. */
void *bfo_geom0_create(const bfo_mesh *m, bfo_status *st);
void bfo_geom0_free(void *g0);
int bfo_finternal_act_cached(const bfo_mesh *m, const bfo_params *p, const void *g0,
                             const double *x, const double *y, const double *z, double *fint,
                             int accumulate, double *c33, bfo_status *st, int nthreads);

/* Same, with an OpenMP-parallel element loop (per-element buffers, serial element-order
 * assembly) for the multi-core CPU baseline.  Bitwise identical to bfo_finternal_act. */
int bfo_finternal_act_omp(const bfo_mesh *m, const bfo_params *p, const double *x,
                          const double *y, const double *z, double *fint, int accumulate,
                          double *c33, bfo_status *st, int nthreads);

/* ---- topology builders (inputs to the path) ---- */

/* AreaNormal (src/io.c:1646-1669): dA[ec] = 0.5*|(x2-x1) x (x3-x1)|, for ec < n_tri. */
void bfo_area_normal(int n_tri, const int *nv1, const int *nv2, const int *nv3,
                     const double *x, const double *y, const double *z, double *dA);

/* IrrVer (src/bending.c:1188-1267). bnodes: boundary node list (sum_n_bnodes). Writes ire,
 * irv, val for n_elmt elements; n_elmt_base as in the reference (0 = off). */
void bfo_irrver(int n_v, int n_elmt, int n_elmt_base, const int *nv1, const int *nv2,
                const int *nv3, int sum_n_bnodes, const int *bnodes, int *ire, int *irv,
                int *val);

/* Patch (src/bending.c:1270-1667), O(E^2) search over n_tri triangles (real then ghost).
 * patch is [stride * n_elmt]; stride 16 reproduces the reference truncation. */
void bfo_patch(int n_v, int n_elmt, int n_elmt_base, int n_tri, const int *nv1, const int *nv2,
               const int *nv3, const int *ire, const int *irv, const int *val, int stride,
               int *patch);

/* GlobalGhostInit (src/bending.c:1670-1760): closed boundary rings, one ghost per boundary
 * edge.  nv1/nv2/nv3 must have room for n_elmt + 2*n_ghosts triangles; the ghost triangles are
 * appended.  Coordinates arrays must have room for n_v + n_ghosts nodes; ghost positions are
 * written to both x/y/z and x0/y0/z0. */
void bfo_global_ghost_init(int n_v, int n_elmt, int n_elmt_base, int n_edge, const int *n_bnodes,
                           const int *bnodes, int *nv1, int *nv2, int *nv3, double *x, double *y,
                           double *z, double *x0, double *y0, double *z0);

/* GlobalGhost (src/bending.c:1763-1814): recompute ghost positions from current coords. */
void bfo_global_ghost(int n_v, int n_elmt, int n_elmt_base, int n_edge, const int *n_bnodes,
                      const int *bnodes, const int *nv1, const int *nv2, const int *nv3,
                      double *x, double *y, double *z);

#ifdef __cplusplus
}
#endif
#endif
