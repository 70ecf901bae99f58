/*
 * biofem_fe.h -- C host layer that mirrors the reference's FInternalAct(FE *) call.
 *
 * The reference keeps mesh state in IBMNodes and options in FE/ActData
 * (include/variables.h, include/act_variables.h:55-105) and evaluates
 *     PetscErrorCode FInternalAct(FE *fem)     (src/active_strain.c:2226-2267)
 * which ADDs every owned element's contribution into the PETSc Vec fem->Fint.
 * bf_fe / bf_ibm hold the same fields as plain C arrays (PETSc is not a dependency of
 * this library), and bf_FInternalAct() has the same contract: it adds into fem->Fint
 * (length 3 * (n_v + n_ghosts), interleaved xyz as the PETSc Vec), never touches ghost
 * rows, and returns 0 or an error code.  The device context is created and the mesh
 * uploaded on the first call (the role of FEM_DMPlexGeomSetup's geom0 pass).
 */
#ifndef BIOFEM_FE_H
#define BIOFEM_FE_H

#include "biofem_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int n_v, n_ghosts, n_elmt, n_elmt_base;
  double *x_bp, *y_bp, *z_bp;       /* current coordinates [n_v + n_ghosts] */
  double *x_bp0, *y_bp0, *z_bp0;    /* reference coordinates */
  int *nv1, *nv2, *nv3;             /* element corners [n_elmt] */
  int *patch;                       /* [patch_stride * n_elmt] */
  int patch_stride;                 /* 16 as in src/io.c:306 (0 = 16) */
  int *ire, *val;                   /* IrrVer output */
  double *dA0;                      /* reference areas */
  double *n_fib;                    /* fibre directions (Cmpnts) */
  double *gamma_scale;              /* NULL = 1 */
  double *activation_delay;         /* NULL = 0 */
  int n_ghost_rules;                /* ghost mirror rules for bf_GlobalGhost (may be 0) */
  int *ghost_rules;
} bf_ibm;

typedef struct {
  bf_ibm *ibm;
  bf_params act;   /* options + globals (h0, mu, K, n_qp, C33 iterations, ti, dt, ...) */
  double *Fint;    /* stand-in for the PETSc Vec: [3 * (n_v + n_ghosts)] */
  int precision;   /* BF_PREC_FP64 or BF_PREC_MIXED */
  int device;
  bf_ctx *ctx;     /* owned; created by the first bf_FInternalAct */
} bf_fe;

/* FInternalAct: Fint += sum_e P_e^T f_e over the n_elmt elements (ghost rows untouched). */
int bf_FInternalAct(bf_fe *fem);

/* GlobalGhost (src/bending.c:1763-1814): refresh ghost coordinates in ibm->x_bp/y_bp/z_bp. */
int bf_GlobalGhost(bf_fe *fem);

/* UpdateElementThickness (src/active_strain.c:2198-2224) from the last evaluation. */
int bf_UpdateElementThickness(bf_fe *fem, double *thickness);

/* Release the device context (ActDataDestroy's role for the device state). */
int bf_FEDestroy(bf_fe *fem);

#ifdef __cplusplus
}
#endif
#endif
