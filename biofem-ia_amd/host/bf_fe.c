/*
 * bf_fe.c -- FInternalAct-shaped host entry over the C ABI (see include/biofem_fe.h).
 * Plain C: this is the layer a BioFEM-IA build links in place of the CPU element loop.
 */
#include "../../include/biofem_fe.h"

#include <stddef.h>

static int ensure_ctx(bf_fe *fem) {
  if (fem->ctx) return BF_OK;
  const bf_ibm *ibm = fem->ibm;
  bf_ctx *ctx = NULL;
  int rc = bf_ctx_create(fem->device, &ctx);
  if (rc != BF_OK) return rc;
  bf_mesh_desc m;
  m.n_v = ibm->n_v;
  m.n_nodes = ibm->n_v + ibm->n_ghosts;
  m.n_elmt = ibm->n_elmt;
  m.n_elmt_base = ibm->n_elmt_base;
  m.patch_stride = ibm->patch_stride > 0 ? ibm->patch_stride : 16;
  m.patch = ibm->patch;
  m.ire = ibm->ire;
  m.val = ibm->val;
  m.nv1 = ibm->nv1;
  m.nv2 = ibm->nv2;
  m.nv3 = ibm->nv3;
  m.x0 = ibm->x_bp0;
  m.y0 = ibm->y_bp0;
  m.z0 = ibm->z_bp0;
  m.dA0 = ibm->dA0;
  m.n_fib = ibm->n_fib;
  m.gamma_scale = ibm->gamma_scale;
  m.activation_delay = ibm->activation_delay;
  m.n_ghost_rules = ibm->n_ghost_rules;
  m.ghost_rules = ibm->ghost_rules;
  rc = bf_mesh_upload(ctx, &m);
  if (rc != BF_OK) {
    bf_ctx_destroy(ctx);
    return rc;
  }
  fem->ctx = ctx;
  return BF_OK;
}

int bf_FInternalAct(bf_fe *fem) {
  if (!fem || !fem->ibm || !fem->Fint) return BF_ERR_ARG;
  int rc = ensure_ctx(fem);
  if (rc != BF_OK) return rc;
  /* options and globals (ti changes every step) are re-read on every call, as the
   * reference reads its globals inside the element loop */
  rc = bf_params_set(fem->ctx, &fem->act);
  if (rc != BF_OK) return rc;
  const bf_ibm *ibm = fem->ibm;
  return bf_force_eval(fem->ctx, ibm->x_bp, ibm->y_bp, ibm->z_bp, fem->Fint, 1, fem->precision);
}

int bf_GlobalGhost(bf_fe *fem) {
  if (!fem || !fem->ibm) return BF_ERR_ARG;
  int rc = ensure_ctx(fem);
  if (rc != BF_OK) return rc;
  return bf_ghost_update(fem->ctx, fem->ibm->x_bp, fem->ibm->y_bp, fem->ibm->z_bp);
}

int bf_UpdateElementThickness(bf_fe *fem, double *thickness) {
  if (!fem || !fem->ctx || !thickness) return BF_ERR_STATE;
  return bf_update_thickness(fem->ctx, thickness);
}

int bf_FEDestroy(bf_fe *fem) {
  if (!fem) return BF_ERR_ARG;
  int rc = bf_ctx_destroy(fem->ctx);
  fem->ctx = NULL;
  return rc;
}
