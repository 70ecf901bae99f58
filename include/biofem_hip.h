/*
 * biofem_hip.h -- C ABI of the MI355X (gfx950) element internal-force library.
 *
 * Drop-in boundary for BioFEM-IA's active-strain force evaluation.  The reference entry
 * point is
 *     PetscErrorCode FInternalAct(FE *fem);          include/active_strain.h:68,
 *                                                    src/active_strain.c:2226-2267
 * called once per residual by FormFunctionFEM (src/main.c:834-836).  Its reads (current
 * coordinates x_bp/y_bp/z_bp incl. ghosts, the patch table, ire/val, dA0, n_fib,
 * gamma_scale, activation_delay, the geom0 cache and the globals h0/ti/dt/mu/K) become the
 * arguments below; its output Vec Fint becomes an interleaved [3 * n_nodes] double array.
 * The existing CUDA bridge (include/cuda_bridge.h:10-15) is the pattern: extern "C",
 * plain pointers, an error code per call.
 *
 * All functions return BF_OK (0) or a negative BF_ERR_* code.  Nothing throws across the
 * ABI.  Host pointers are borrowed for the duration of the call only.  A context owns its
 * device memory and may be used from one host thread at a time.
 */
#ifndef BIOFEM_HIP_H
#define BIOFEM_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes (mapping to the reference's PetscErrorCode classes) ---- */
#define BF_OK 0
#define BF_ERR_ARG (-1)         /* PETSC_ERR_ARG_* : bad size, NULL, option out of range */
#define BF_ERR_DEVICE (-2)      /* PETSC_ERR_LIB   : HIP runtime failure */
#define BF_ERR_NOMEM (-3)       /* PETSC_ERR_MEM   : device allocation failed */
#define BF_ERR_ELEMENT (-4)     /* PETSC_ERR_PLIB  : per-element failure, see bf_last_error */
#define BF_ERR_STATE (-5)       /* call order (e.g. eval before mesh upload) */
#define BF_ERR_UNSUPPORTED (-6) /* PETSC_ERR_SUP   : e.g. patch wider than 18 nodes */
#define BF_ERR_IO (-7)          /* PETSC_ERR_FILE_* : missing, unwritable or malformed file */

/* per-element codes reported by bf_last_error() */
#define BF_ELEM_STENCIL 2   /* patch slot empty (active_strain.c:498-500, :577-579) */
#define BF_ELEM_IRE 3       /* ire not 0/1 (active_strain.c:721-723) */
#define BF_ELEM_NEN 4       /* nen above the patch width (active_strain.c:599-600) */
#define BF_ELEM_CC3333 6    /* |CC3333| <= 1e-14 (active_strain.c:1674-1681); non-fatal,
                               the C33 update is skipped as in the reference */
#define BF_ELEM_NONFINITE 7 /* NaN/Inf in an element force */

/* precision modes of the force evaluation */
#define BF_PREC_FP64 0      /* plain FP64 in the reference's operation order (no contraction):
                               reproduces the CPU FInternalAct bit for bit */
#define BF_PREC_MIXED 1     /* FP32 per-quadrature-point geometry/constitutive arithmetic on the
                               staged contractions; FP64 coordinates, midsurface sums and nodal
                               accumulation */
#define BF_PREC_FP64_FAST 2 /* FP64 on the staged contractions (closed-form CC^{2222}, 3x3
                               products instead of the 6- and 8-deep index loops) */
/* Both staged modes hand every element whose local C33 iteration leaves its
 * well-conditioned range (ModElemC33's bisection clamp, the CC3333 guard, a non-finite
 * value, C33 outside [0.1, 10] -- [0.4, 2.5] for the FP32 chain -- at a qp, or an update
 * whose rounding bound exceeds the path's budget) to the plain kernel, which recomputes it
 * in the reference's operation order (bf_set_fallback); the FP32-mixed mode first retries
 * such elements with the FP64 staged chain at the mixed budget. */

#define BF_PATCH_SENTINEL 1000000 /* empty patch slot (src/io.c:306) */
#define BF_MAX_NEN 18             /* widest patch (valence 12) */

typedef struct bf_ctx bf_ctx;

/* Mesh arrays the force path reads (the IBMNodes fields of src/active_strain.c). */
typedef struct {
  int n_v;           /* real nodes; forces on nodes >= n_v are dropped (active_strain.c:2254) */
  int n_nodes;       /* real + ghost nodes (length of every coordinate array) */
  int n_elmt;        /* elements evaluated */
  int n_elmt_base;   /* > 0: elements >= n_elmt_base are LV apex-cap CST elements */
  int patch_stride;  /* row width of `patch`: 16 in the reference layout, at most 18 */
  const int *patch;  /* [patch_stride * n_elmt], BF_PATCH_SENTINEL for empty slots */
  const int *ire;    /* [n_elmt] 0 regular, 1 irregular */
  const int *val;    /* [n_elmt] valence (6 for regular elements) */
  const int *nv1, *nv2, *nv3;       /* [n_elmt] corners; needed only when n_elmt_base > 0 */
  const double *x0, *y0, *z0;       /* reference coordinates [n_nodes] (x_bp0 ...) */
  const double *dA0;                /* [n_elmt] reference flat areas (io.c:1646-1669) */
  const double *n_fib;              /* [3 * n_elmt] fibre directions, Cmpnts layout */
  const double *gamma_scale;        /* [n_elmt] or NULL (= 1) */
  const double *activation_delay;   /* [n_elmt] or NULL (= 0) */
  int n_ghost_rules;                /* ghosts recomputed by bf_ghost_update */
  const int *ghost_rules;           /* [3 * n_ghost_rules]: ghost n_v+g = x[a] + x[b] - x[c] */
} bf_mesh_desc;

/* Options and globals of the path (src/active_strain.c:90-110, 174-230, 1156-1179). */
typedef struct {
  int n_qp;          /* -num_gaussian_quad_points, 1..5 */
  int c33_iters;     /* -C33_subitr_nums, >= 1 */
  int cart_fib_act;  /* -cart_fib_act: Fa = I - gamma f (x) f */
  int manufactured;  /* manufactured-solution activation (active_strain.c:1240-1250) */
  double h0, mu, K;  /* thickness, shear modulus (-mu), bulk modulus (-bulk_modulus) */
  double gamma;      /* -muscle_act_gamma */
  int ramp_type;     /* -gamma_ramp_type 0..3; 4 = Gaussian pulse (no reference counterpart) */
  double gamma_T, tau_up, tau_down, apd; /* -gamma_T -gamma_tau_up -gamma_tau_down -gamma_apd */
  double pulse_t0, pulse_sigma;          /* ramp type 4: gamma exp(-((t-t0)/sigma)^2 / 2) */
  double manufactured_gamma0, manufactured_T;
  int ti, tisteps;   /* time-step index and count (globals) */
  double dt;
} bf_params;

/* ---- context ---- */
int bf_ctx_create(int device, bf_ctx **out);
int bf_ctx_destroy(bf_ctx *ctx);

/* Upload topology, reference coordinates and per-element data; builds the geom0 cache
 * (replaces ElemUpdateGeom0Subdiv over all cells in FEM_DMPlexGeomSetup,
 * src/dmplex_geom.c:859-865).  Validates every patch slot (the reference's
 * per-residual stencil check, active_strain.c:498/:577) once, here. */
int bf_mesh_upload(bf_ctx *ctx, const bf_mesh_desc *mesh);

/* Options/globals (GetUserActParams, SetGaussianQuadrature; globals h0, ti, dt). */
int bf_params_set(bf_ctx *ctx, const bf_params *p);

/* Replaces FInternalAct (active_strain.c:2226-2267) on one rank: current coordinates in,
 * interleaved forces out.  fint has 3 * n_nodes entries; accumulate = 0 writes
 * Fint = sum of element contributions (the caller's VecSet(Fint, 0) + ADD_VALUES), with
 * ghost rows zero; accumulate = 1 adds to fint and leaves ghost rows untouched.  Node
 * sums run in ascending element order, as VecSetValues(ADD_VALUES) does on one rank.
 * Host pointers; the call is synchronous. */
int bf_force_eval(bf_ctx *ctx, const double *x, const double *y, const double *z,
                  double *fint, int accumulate, int precision);

/* Same on device-resident arrays, asynchronous on `stream` (a hipStream_t, NULL = the
 * context's stream, which is non-blocking: a caller whose work runs on the legacy NULL
 * stream orders its own stream against the launches with events).  Status is collected
 * by bf_last_error(). */
int bf_force_eval_device(bf_ctx *ctx, const double *d_x, const double *d_y, const double *d_z,
                         double *d_fint, int accumulate, int precision, void *stream);

/* bf_force_eval_device without any device allocation, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * reference counterpart.  The per-evaluation scratch (element-force buffer, hand-over lists,
 * per-qp stresses) lives in one static device array of 64 MiB per GPU, held until the
 * process ends, instead of the context's buffers; meshes needing more (about 150 000
 * elements with 12-slot patches) get BF_ERR_UNSUPPORTED.  The context still owns the mesh,
 * the C33 state, the flags and the status.  All calls on one GPU share the array: calls that
 * overlap on one GPU (different streams, threads or contexts) race.  Same results as
 * bf_force_eval_device. */
int bf_force_eval_device_noalloc(bf_ctx *ctx, const double *d_x, const double *d_y,
                                 const double *d_z, double *d_fint, int accumulate,
                                 int precision, void *stream);

/* GlobalGhost (src/bending.c:1763-1814) as an O(n_ghosts) gather:
 * x[n_v + g] = x[a] + x[b] - x[c] for every ghost rule.  Host and device forms. */
int bf_ghost_update(bf_ctx *ctx, double *x, double *y, double *z);
int bf_ghost_update_device(bf_ctx *ctx, double *d_x, double *d_y, double *d_z, void *stream);

/* C33 state of the last evaluation, [n_elmt * n_qp] (ead->C[qp].Cov[2][2]). */
int bf_c33_download(bf_ctx *ctx, double *c33);

/* Per-element status bits of the last evaluation, [n_elmt]: 1 = CC3333 guard tripped
 * (BF_ELEM_CC3333), 2 = non-finite force, 4 = the det(C) > 0 bisection of ModElemC33
 * (active_strain.c:1696-1706) ran, i.e. the local C33 iteration left its well-conditioned
 * regime.  No reference counterpart: the reference clamps silently. */
int bf_elem_flags_download(bf_ctx *ctx, unsigned char *flags);

/* UpdateElementThickness (active_strain.c:2198-2224): thickness[ec] = h0 sqrt(mean C33). */
int bf_update_thickness(bf_ctx *ctx, double *thickness);

/* First failure of the last evaluation: element index (-1 if none), BF_ELEM_* code,
 * number of elements that raised the non-fatal CC3333 condition.  Synchronises. */
int bf_last_error(bf_ctx *ctx, int *elem, int *code, int *n_warn);

/* Number of elements the last staged evaluation handed to the reference's operation order
 * (their forces are formed in that order: BF_PREC_FP64_FAST solves every through-thickness
 * point of such an element in that order, BF_PREC_MIXED only the points that fail its
 * conditioning tests and keeps the staged stresses of the others); bf_set_fallback(ctx, 0)
 * disables that hand-over (default on).  The count read here (and by bf_force_eval,
 * bf_last_error) also steers the next evaluation of that precision: after a short list the
 * hand-over solve adds a kernel that spreads each (element, qp) over more lanes.  The
 * results do not depend on it.  No reference counterpart: the reference has a single
 * arithmetic path. */
int bf_last_fallback_count(bf_ctx *ctx, int *n);
int bf_set_fallback(bf_ctx *ctx, int on);

/* The (element, qp) pairs the last staged evaluation solved in the reference's operation
 * order: mask[e] bit q is set when through-thickness point q of element e was (elements
 * whose CC3333 guard tripped are redone with all their points in that order as well; their
 * flags carry BF_ELEM_CC3333).  *n_pairs: the number of pairs.  Synchronises.  No reference
 * counterpart: instrumentation of the hand-over. */
int bf_last_handed_qps(bf_ctx *ctx, unsigned char *mask, int *n_pairs);

/* Wall-clock duration of the last bf_force_eval* kernels on the device, in ms
 * (HIP events on the evaluation stream): element kernel and node-gather kernel.  The events
 * are recorded while timing is on (bf_timing_enable); otherwise BF_ERR_STATE. */
int bf_last_timing(bf_ctx *ctx, float *ms_element, float *ms_gather);

/* Kernel timing over many evaluations: bf_timing_enable(ctx, 1) resets the totals and
 * records HIP events around the element and gather kernels of every later evaluation (each
 * record costs a few microseconds of device time between the kernels, so an untimed
 * evaluation records none but the one after its last kernel);
 * bf_timing_read() synchronises and returns the count and total milliseconds
 * (ms_element includes the hand-over kernel of the staged precisions). */
int bf_timing_enable(bf_ctx *ctx, int on);
int bf_timing_read(bf_ctx *ctx, long long *count, double *ms_element, double *ms_gather);
/* Same totals with the element phase split: the element kernel alone and the hand-over
 * kernel (reference-order recomputation of flagged elements; 0 for BF_PREC_FP64).
 * No reference counterpart (instrumentation). */
int bf_timing_read_split(bf_ctx *ctx, long long *count, double *ms_element,
                         double *ms_handover, double *ms_gather);
/* Total milliseconds of the hand-over solve kernels alone (the reference-order C33 solve of
 * the handed-over (element, qp) items, inside ms_handover) over the same evaluations.  No
 * reference counterpart (instrumentation). */
int bf_timing_read_solve(bf_ctx *ctx, long long *count, double *ms_solve);

/* Device pointers owned by the context, for callers that keep state resident:
 * coordinates and force buffers of n_nodes (3 * n_nodes for forces). */
int bf_device_buffers(bf_ctx *ctx, double **d_x, double **d_y, double **d_z, double **d_fint);
int bf_sync(bf_ctx *ctx);

/* ---- host topology (inputs of the path; no device needed) ---- */
/* AreaNormal (src/io.c:1646-1669): dA[e] = 0.5 |(x2-x1) x (x3-x1)|. */
int bf_topo_area0(int n_tri, const int *nv1, const int *nv2, const int *nv3, const double *x,
                  const double *y, const double *z, double *dA);
/* IrrVer (src/bending.c:1188-1267). */
int bf_topo_irrver(int n_v, int n_elmt, int n_elmt_base, const int *nv1, const int *nv2,
                   const int *nv3, int n_bnodes, const int *bnodes, int *ire, int *irv,
                   int *val);
/* Patch (src/bending.c:1270-1667) over n_tri triangles (real first, then ghost). */
int bf_topo_patch(int n_v, int n_nodes, int n_elmt, int n_elmt_base, int n_tri, const int *nv1,
                  const int *nv2, const int *nv3, const int *ire, const int *irv,
                  const int *val, int stride, int *patch);
/* GlobalGhostInit (src/bending.c:1670-1760): appends 2 ghost triangles per boundary edge
 * and writes the ghost rules (b, a, c). */
int bf_topo_ghost_ring(int n_v, int n_elmt, int n_elmt_base, int n_edge, const int *n_bnodes,
                       const int *bnodes, int *nv1, int *nv2, int *nv3, int *rules);
/* IN* rows of a valence-v irregular patch (active_strain.c:594-666), out[5][v+6]. */
int bf_loop_tables(int v, double *out);

const char *bf_version(void);

/* ---- time stepping on resident state ----
 * The reference advances FormFunctionFEM's residual with an implicit Newmark step solved by
 * PETSc SNES/KSP (src/main.c:461-502, 757-1032), unavailable here.  Its substitute is an
 * explicit damped step on the device, in FP64 without contraction:
 *     f = Fint(x)  (the force of this library, GlobalGhost first),
 *     v <- damping * v - (dt / m_n) f,   x <- x + dt v    (nodes not fixed),
 * with lumped masses m_n supplied by the caller.  bf_state_init sets the state (v0 NULL =
 * at rest; fixed[n] != 0 clamps node n, as EdgeDirectionalFix would); bf_step advances
 * n_steps steps asynchronously on the context stream; bf_state_download copies x, y, z
 * (n_nodes each) and v (3 * n_nodes, interleaved) back, any of them NULL to skip. */
int bf_state_init(bf_ctx *ctx, const double *mass, const unsigned char *fixed,
                  const double *v0, double dt, double damping);
int bf_step(bf_ctx *ctx, int precision, int n_steps);
int bf_state_download(bf_ctx *ctx, double *x, double *y, double *z, double *v);

/* Host copies to and from a context's resident buffers (the arrays bf_device_buffers
 * returns): coordinates [n_nodes] each, forces [3 * n_nodes].  Synchronous.  No reference
 * counterpart (the reference's arrays live in host memory). */
int bf_coords_upload(bf_ctx *ctx, const double *x, const double *y, const double *z);
int bf_fint_download(bf_ctx *ctx, double *fint);

/* ---- batches: many meshes in the same launches ----
 * A batch evaluates the force of every case (a context with a mesh and params) in one
 * step of five launches -- status reset + GlobalGhost, the element phase (a persistent
 * kernel that takes element tiles of all cases from a device work list), the compaction
 * of hand-over lists, the reference-order hand-over kernel, the node gather -- optionally
 * replayed from a captured hipGraph.  Each case reads its resident coordinates and writes
 * its resident forces (accumulate = 0), as bf_force_eval_device on its own buffers would;
 * per-case status, C33 and flags land in the case's context.  The batching of
 * `biofem sweep` cases has no reference counterpart (the reference runs one case per
 * process). */
typedef struct bf_batch bf_batch;
int bf_batch_create(int device, bf_batch **out);
int bf_batch_destroy(bf_batch *batch);
/* The cases of the batch (contexts on the batch's device, each with a mesh and params;
 * they must stay alive while the batch uses them).  Rebuilds the work lists. */
int bf_batch_set_cases(bf_batch *batch, bf_ctx *const *ctxs, int n);
/* Re-reads the cases' params (after bf_params_set on a case). */
int bf_batch_refresh(bf_batch *batch);
/* One step for all cases; ghosts != 0 runs GlobalGhost on the resident coordinates first;
 * use_graph != 0 captures the step once per (precision, ghosts) and replays it.
 * Asynchronous on the batch's stream. */
int bf_batch_eval(bf_batch *batch, int precision, int ghosts, int use_graph);
/* n_steps time steps of every case that holds a state (bf_state_init before
 * bf_batch_set_cases / bf_batch_refresh); cases without one are only evaluated. */
int bf_batch_step(bf_batch *batch, int precision, int n_steps, int use_graph);
int bf_batch_sync(bf_batch *batch);
/* Device time of the last bf_batch_eval / bf_batch_step call, in ms (synchronises): HIP
 * events around the call, recorded while bf_batch_timing_enable(batch, 1) is on (each
 * record costs a few microseconds of device time); otherwise BF_ERR_STATE.  No reference
 * counterpart (instrumentation). */
int bf_batch_timing_enable(bf_batch *batch, int on);
int bf_batch_last_ms(bf_batch *batch, float *ms);
/* Elements of all cases the last staged step recomputed in the reference's order. */
int bf_batch_last_fallback_count(bf_batch *batch, int *n);
/* Per case of the batch (n_cases entries each, either array may be NULL): the first element
 * whose force was not finite and the first element whose CC3333 guard tripped in the last
 * step, -1 for none -- what bf_force_eval reports per context as BF_ERR_ELEMENT and the
 * warning count.  Lets a request loop retire a case whose state diverged.  Synchronises. */
int bf_batch_case_status(bf_batch *batch, int *first_nonfinite, int *first_cc3333);

/* A served round without a host round trip in between: bf_batch_step_async enqueues
 * bf_batch_step (ghosts on) followed by the per-case status of its last step into round
 * slot (round & 1) and returns; bf_batch_step_wait waits for that round and returns its GPU
 * time and per-case status as bf_batch_case_status does, in the case order the round ran
 * with.  At most two rounds in flight (round r's slot must be waited for before round
 * r + 2 is enqueued); bf_batch_set_cases refuses (BF_ERR_STATE) while a round is in flight.
 * No reference counterpart (the serving loop replaces `biofem sweep`). */
int bf_batch_step_async(bf_batch *batch, int precision, int n_steps, int use_graph,
                        long long round);
int bf_batch_step_wait(bf_batch *batch, long long round, float *ms, int *first_nonfinite,
                       int *first_cc3333);
int bf_batch_info(bf_batch *batch, int *n_cases, long long *n_elmt, int *n_tiles, int *tile_cap);
/* Launch shape of the element phase: `passes` element groups per work item (a tile holds
 * passes * floor(64 / n_qp) elements), threads per block (64, 128 or 256), blocks
 * (0 = one wave per SIMD slot of the kernel's occupancy), persistent (1: waves stride over
 * the work list, 0: one work item per wave).  Until it is set or tuned, passes follows the
 * case set: 1 to 4, as many as keep every wave slot of the element phase busy. */
int bf_batch_set_launch(bf_batch *batch, int passes, int block, int blocks, int persistent);
int bf_batch_get_launch(bf_batch *batch, int *passes, int *block, int *blocks, int *persistent);
/* Autotuner: times every launch shape of a fixed candidate set on the batch's own cases
 * (reps steps each, precision as given) and keeps the fastest. */
int bf_batch_tune(bf_batch *batch, int precision, int reps, int *passes, int *block,
                  int *blocks, int *persistent, float *best_ms);

#ifdef __cplusplus
}
#endif
#endif
