/*
 * bf_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY) for BioFEM-IA's active-strain
 * element internal-force path.  Plain C, FP64, built with -ffp-contract=off so every
 * operation rounds as the reference's C source does.  Each function cites the
 * reference lines it restates; arithmetic order follows those lines.
 *
 * Parity pins, per function (tests/test_oracle.py = T, tests/test_oracle_pins.py = P,
 * tests/test_topology.py = Y); none retypes the formula it checks:
 *  - Na_c/Nab_c tables (:69-88): derivatives of the 12 Loop box-spline basis functions at
 *    the centre, symbolic, rounded to 4 decimals (T).
 *  - bfo_gauss (:174-230): Gauss-Legendre moments and numpy's leggauss (T).
 *  - bfo_gamma_of_time (:1156-1179): closed-form ramp values of types 0-4 (T).
 *  - bfo_geom_subdiv regular (:477-558): affine and quadratic reproduction (T); irregular
 *    (:563-719) and bfo_irregular_tables: one independent Loop subdivision step, affine row
 *    sums (T); LV cap CST (:448-472): uniform stretch against the textbook CST force (P).
 *  - a3_alpha (:745-762) and elem_update_g (:896-982) off the mid-surface: prescribed
 *    polynomial surface, symbolic derivative of the unnormalised normal, closed-form C33
 *    root and stress per through-thickness point (P).
 *  - elem_act_def_grad (:1196-1329): stress-free active state x = Fa X for the fibre law (T),
 *    C33 closed form for the isotropic law (T), stress-free manufactured state
 *    x = lambda(t) X (P).
 *  - elem_cg_def_tens .. elem_tot_tang, elem_c33_solve (:1332-1648, :2048-2138): uniaxial
 *    and isotropic C33 roots and the rate 1/2 of the half-Newton step (T), in-plane
 *    stress closed form at every qp (P), objectivity, force-free reference state (T).
 *  - mod_elem_c33 (:1654-1721): bisection clamp against its grid closed form, CC3333 guard
 *    skip and trip count (P).
 *  - elem_upd_fint (:1762-1995): membrane and bending rows against finite differences of
 *    the metric and curvature, regular and irregular patches (P); cap CST (P).
 *  - assemble / bfo_finternal_act (:2226-2267): brute-force element-order sum, accumulate,
 *    OpenMP build bitwise (T).
 *  - bfo_area_normal, bfo_irrver, bfo_patch, bfo_global_ghost_init/bfo_global_ghost:
 *    lattice areas (T), the hashed native builders and geometric one-rings (Y).
 * The reference binary cannot be built here (PETSc/MPI absent), so no function has
 * golden vectors from the reference itself; bitwise agreement with the reference
 * executable (VecSetValues summation included) is therefore parity unpinned.
 */
#include "bf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI_ 3.14159265358979323846264338327950288 /* PETSC_PI */

/* Loop box-spline derivatives at the element centre, 4 decimals, verbatim values of
 * src/active_strain.c:69-88 (kept truncated: the truncation is part of the behaviour). */
static const double Na_c[2][12] = {
    {-0.0247, -0.0309, 0.0000, -0.4815, -0.1852, 0.0247, 0.4815, 0.0000, -0.0062, 0.0309,
     0.1852, 0.0062},
    {-0.0309, -0.0247, -0.1852, -0.4815, 0.0000, -0.0062, 0.0000, 0.4815, 0.0247, 0.0062,
     0.1852, 0.0309}};
static const double Nab_c[3][12] = {
    {0.1111, 0.2222, -0.2222, -0.2222, 0.4444, 0.1111, -0.2222, -0.8889, 0.0000, 0.2222,
     0.4444, 0.0000},
    {0.2222, 0.1111, 0.4444, -0.2222, -0.2222, 0.0000, -0.8889, -0.2222, 0.1111, 0.0000,
     0.4444, 0.2222},
    {0.1667, 0.1667, -0.1111, 0.2222, -0.1111, -0.0556, -0.4444, -0.4444, -0.0556, 0.0556,
     0.5556, 0.0556}};

/* ------------------------------------------------------------------ math.c:5-245 ---- */
static bfo_vec3 PLUS(bfo_vec3 a, bfo_vec3 b) {
  bfo_vec3 r = {a.x + b.x, a.y + b.y, a.z + b.z};
  return r;
}
static bfo_vec3 CROSS(bfo_vec3 v1, bfo_vec3 v2) { /* math.c:29-37 */
  bfo_vec3 r;
  r.x = v1.y * v2.z - v2.y * v1.z;
  r.y = -v1.x * v2.z + v2.x * v1.z;
  r.z = v1.x * v2.y - v2.x * v1.y;
  return r;
}
static double DOT(bfo_vec3 a, bfo_vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static bfo_vec3 UNIT(bfo_vec3 v) { /* math.c:47-55 */
  double n = sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
  bfo_vec3 r = {v.x / n, v.y / n, v.z / n};
  return r;
}
static double SIZE(bfo_vec3 v) { return sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
static bfo_vec3 AMULT(double a, bfo_vec3 v) {
  bfo_vec3 r = {a * v.x, a * v.y, a * v.z};
  return r;
}
static void INV(const double T[3][3], double Ti[3][3]) { /* math.c:77-91, no det guard */
  double T11 = T[0][0], T12 = T[0][1], T13 = T[0][2];
  double T21 = T[1][0], T22 = T[1][1], T23 = T[1][2];
  double T31 = T[2][0], T32 = T[2][1], T33 = T[2][2];
  double det = T11 * (T33 * T22 - T32 * T23) - T21 * (T33 * T12 - T32 * T13) +
               T31 * (T23 * T12 - T22 * T13);
  double c = 1 / det;
  Ti[0][0] = c * (T33 * T22 - T32 * T23);
  Ti[0][1] = -c * (T33 * T12 - T32 * T13);
  Ti[0][2] = c * (T23 * T12 - T22 * T13);
  Ti[1][0] = -c * (T33 * T21 - T31 * T23);
  Ti[1][1] = c * (T33 * T11 - T31 * T13);
  Ti[1][2] = -c * (T23 * T11 - T21 * T13);
  Ti[2][0] = c * (T32 * T21 - T31 * T22);
  Ti[2][1] = -c * (T32 * T11 - T31 * T12);
  Ti[2][2] = c * (T22 * T11 - T21 * T12);
}
static void MATMULT(const double A[3][3], const double B[3][3], double C[3][3]) { /* :95-120 */
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      C[i][j] = 0.0;
      for (int k = 0; k < 3; k++) C[i][j] += A[i][k] * B[k][j];
    }
}
static void TRANS(const double A[3][3], double AT[3][3]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) AT[i][j] = A[j][i];
}
static double DET3x3(const double A[3][3]) { /* math.c:150-156 */
  return A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
         A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
         A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
}
static void Raise2(const double gInv[3][3], const double A[3][3], double out[3][3]) {
  double tmp[3][3]; /* math.c:159-172 */
  MATMULT(gInv, A, tmp);
  MATMULT(tmp, gInv, out);
}
static void Lower2(const double gCov[3][3], const double A[3][3], double out[3][3]) {
  double tmp[3][3]; /* math.c:174-186 */
  MATMULT(gCov, A, tmp);
  MATMULT(tmp, gCov, out);
}
static void copy33(double d[3][3], const double s[3][3]) { memcpy(d, s, 9 * sizeof(double)); }

/* --------------------------------------------------- SetGaussianQuadrature :174-230 -- */
int bfo_gauss(int n_qp, double *theta, double *w) {
  for (int i = 0; i < n_qp && i < 5; i++) theta[i] = w[i] = 0.0;
  switch (n_qp) {
  case 1:
    theta[0] = 0.0;
    w[0] = 2.0;
    break;
  case 2:
    theta[0] = -0.5773502691896257;
    theta[1] = 0.5773502691896257;
    w[0] = w[1] = 1.0;
    break;
  case 3:
    theta[0] = -0.7745966692414834;
    theta[1] = 0.0;
    theta[2] = 0.7745966692414834;
    w[0] = w[2] = 0.5555555555555556;
    w[1] = 0.8888888888888888;
    break;
  case 4:
    theta[0] = -0.8611363115940526;
    theta[1] = -0.3399810435848563;
    theta[2] = 0.3399810435848563;
    theta[3] = 0.8611363115940526;
    w[0] = w[3] = 0.3478548451374538;
    w[1] = w[2] = 0.6521451548625461;
    break;
  case 5:
    theta[0] = -0.9061798459386640;
    theta[1] = -0.5384693101056831;
    theta[2] = 0.0;
    theta[3] = 0.5384693101056831;
    theta[4] = 0.9061798459386640;
    w[0] = w[4] = 0.2369268850561891;
    w[1] = w[3] = 0.4786286704993665;
    w[2] = 0.5688888888888889;
    break;
  default:
    return BFO_ERR_NQP;
  }
  return BFO_OK;
}

/* ------------------------------------------------ GammaOfTimeDelayed :1156-1179 ----- */
double bfo_gamma_of_time(const bfo_params *p, double delay) {
  if (p->ramp_type == 0) return p->gamma;
  double t = (double)p->ti * p->dt - delay;
  if (p->ramp_type == 3) {
    double tau_up = (p->tau_up > 0.0) ? p->tau_up : 0.02;
    double tau_down = (p->tau_down > 0.0) ? p->tau_down : 0.08;
    double apd = (p->apd > 0.0) ? p->apd : 0.30;
    double t_up = 4.0 * tau_up;
    double t_down = t_up + apd;
    double phi_up = 1.0 / (1.0 + exp(-(t - t_up) / tau_up));
    double phi_down = 1.0 / (1.0 + exp((t - t_down) / tau_down));
    return p->gamma * phi_up * phi_down;
  }
  if (p->ramp_type == 4) { /* Gaussian pulse in time: not a reference ramp type */
    double s = (t - p->pulse_t0) / p->pulse_sigma;
    return p->gamma * exp(-0.5 * s * s);
  }
  double T = (p->gamma_T > 0.0) ? p->gamma_T : (double)p->tisteps * p->dt;
  if (t <= 0.0) return 0.0;
  if (p->ramp_type == 1) {
    double s = sin(PI_ * t / T);
    return p->gamma * s * s;
  }
  double s = sin(PI_ * 0.5 * t / T);
  return p->gamma * s * s;
}

/* ------------------------------- irregular IN* tables, active_strain.c:594-666 ------- */
void bfo_irregular_tables(int v, double *INa0, double *INa1, double *INab0, double *INab1,
                          double *INab2) {
  const int nen = v + 6;
  double w = (1.0 / (double)v) * (0.625 - pow(0.375 + 0.25 * cos(2 * PI_ / (double)v), 2.0));
  double B2[12][64], B1[64][64], B3[12][64];
  for (int i = 0; i < 12; i++)
    for (int j = 0; j < v + 12; j++) B2[i][j] = 0.0;
  for (int i = 0; i < v + 12; i++)
    for (int j = 0; j < nen; j++) B1[i][j] = 0.0;
  B2[0][v + 9] = 1.;
  B2[1][v + 6] = 1.;
  B2[2][v + 4] = 1.;
  B2[3][v + 1] = 1.;
  B2[4][v + 2] = 1.;
  B2[5][v + 5] = 1.;
  B2[6][v] = 1.;
  B2[7][1] = 1.;
  B2[8][v + 3] = 1.;
  B2[9][v - 1] = 1.;
  B2[10][0] = 1.;
  B2[11][2] = 1.;
  B1[0][0] = 1 - v * w;
  for (int j = 0; j < v; j++) B1[0][1 + j] = w;
  B1[1][0] = 0.375; B1[1][1] = 0.375; B1[1][2] = 0.125; B1[1][v] = 0.125;
  B1[2][0] = 0.375; B1[2][1] = 0.125; B1[2][2] = 0.375; B1[2][3] = 0.125;
  if (v > 5) { B1[v - 4][0] = 0.375; B1[v - 4][v - 5] = 0.125; B1[v - 4][v - 4] = 0.375; B1[v - 4][v - 3] = 0.125; }
  if (v > 4) { B1[v - 3][0] = 0.375; B1[v - 3][v - 4] = 0.125; B1[v - 3][v - 3] = 0.375; B1[v - 3][v - 2] = 0.125; }
  B1[v - 2][0] = 0.375; B1[v - 2][v - 3] = 0.125; B1[v - 2][v - 2] = 0.375; B1[v - 2][v - 1] = 0.125;
  B1[v - 1][0] = 0.375; B1[v - 1][v - 2] = 0.125; B1[v - 1][v - 1] = 0.375; B1[v - 1][v] = 0.125;
  B1[v][0] = 0.375; B1[v][1] = 0.125; B1[v][v - 1] = 0.125; B1[v][v] = 0.375;
  B1[v + 1][0] = 0.125; B1[v + 1][1] = 0.375; B1[v + 1][v] = 0.375; B1[v + 1][v + 1] = 0.125;
  B1[v + 2][0] = 0.0625; B1[v + 2][1] = 0.625; B1[v + 2][2] = 0.0625; B1[v + 2][v] = 0.0625;
  B1[v + 2][v + 1] = 0.0625; B1[v + 2][v + 2] = 0.0625; B1[v + 2][v + 3] = 0.0625;
  B1[v + 3][0] = 0.125; B1[v + 3][1] = 0.375; B1[v + 3][2] = 0.375; B1[v + 3][v + 3] = 0.125;
  B1[v + 4][0] = 0.0625; B1[v + 4][1] = 0.0625; B1[v + 4][v - 1] = 0.0625; B1[v + 4][v] = 0.625;
  B1[v + 4][v + 1] = 0.0625; B1[v + 4][v + 4] = 0.0625; B1[v + 4][v + 5] = 0.0625;
  B1[v + 5][0] = 0.125; B1[v + 5][v - 1] = 0.375; B1[v + 5][v] = 0.375; B1[v + 5][v + 5] = 0.125;
  B1[v + 6][1] = 0.375; B1[v + 6][v] = 0.125; B1[v + 6][v + 1] = 0.375; B1[v + 6][v + 2] = 0.125;
  B1[v + 7][1] = 0.375; B1[v + 7][v + 1] = 0.125; B1[v + 7][v + 2] = 0.375; B1[v + 7][v + 3] = 0.125;
  B1[v + 8][1] = 0.375; B1[v + 8][2] = 0.125; B1[v + 8][v + 2] = 0.125; B1[v + 8][v + 3] = 0.375;
  B1[v + 9][1] = 0.125; B1[v + 9][v] = 0.375; B1[v + 9][v + 1] = 0.375; B1[v + 9][v + 4] = 0.125;
  B1[v + 10][v] = 0.375; B1[v + 10][v + 1] = 0.125; B1[v + 10][v + 4] = 0.375; B1[v + 10][v + 5] = 0.125;
  B1[v + 11][v - 1] = 0.125; B1[v + 11][v] = 0.375; B1[v + 11][v + 4] = 0.125; B1[v + 11][v + 5] = 0.375;
  for (int i = 0; i < 12; i++)
    for (int j = 0; j < nen; j++) {
      double s = 0.0;
      for (int m = 0; m < v + 12; m++) s += B2[i][m] * B1[m][j];
      B3[i][j] = s;
    }
  for (int j = 0; j < nen; j++) {
    double s1 = 0., s2 = 0., s3 = 0., s4 = 0., s5 = 0.;
    for (int i = 0; i < 12; i++) {
      s1 += B3[i][j] * Na_c[0][i];
      s2 += B3[i][j] * Na_c[1][i];
      s3 += B3[i][j] * Nab_c[0][i];
      s4 += B3[i][j] * Nab_c[1][i];
      s5 += B3[i][j] * Nab_c[2][i];
    }
    INa0[j] = -2.0 * s1;
    INa1[j] = -2.0 * s2;
    INab0[j] = 4.0 * s3;
    INab1[j] = 4.0 * s4;
    INab2[j] = 4.0 * s5;
  }
}

/* Tail shared by every branch: a3 = unit(a1 x a2), contravariant in-plane basis
 * (active_strain.c:535-541 / :693-699 / :452-456). */
static void geom_tail(bfo_geom *G) {
  G->nn = UNIT(CROSS(G->ndx21, G->ndx31));
  G->gc1 = CROSS(G->ndx31, G->nn);
  G->gc1 = AMULT(1.0 / DOT(G->ndx21, G->gc1), G->gc1);
  G->gc2 = CROSS(G->nn, G->ndx21);
  G->gc2 = AMULT(1.0 / DOT(G->ndx31, G->gc2), G->gc2);
}

/* ------------------------- ElemUpdateGeomSubdivFromCoords_, active_strain.c:428-724 ---- */
int bfo_geom_subdiv(const bfo_mesh *m, int ec, const double *xb, const double *yb,
                    const double *zb, bfo_geom *G) {
  const int *pt = m->patch + (size_t)m->patch_stride * ec;
  if (m->n_elmt_base > 0 && ec >= m->n_elmt_base) { /* LV apex cap: CST geometry :448-472 */
    const int n1 = m->nv1[ec], n2 = m->nv2[ec], n3 = m->nv3[ec];
    G->ndx21.x = xb[n2] - xb[n1]; G->ndx21.y = yb[n2] - yb[n1]; G->ndx21.z = zb[n2] - zb[n1];
    G->ndx31.x = xb[n3] - xb[n1]; G->ndx31.y = yb[n3] - yb[n1]; G->ndx31.z = zb[n3] - zb[n1];
    geom_tail(G);
    G->Aaa.x = G->Aaa.y = G->Aaa.z = 0.0;
    G->Abb.x = G->Abb.y = G->Abb.z = 0.0;
    G->Aab.x = G->Aab.y = G->Aab.z = 0.0;
    G->is_irregular = 0;
    G->v = 0;
    G->nen = 12;
    return BFO_OK;
  }
  if (m->ire[ec] == 0) { /* regular patch :477-558 */
    double x[12] = {0}, y[12] = {0}, z[12] = {0};
    int nob = 1;
    for (int i = 0; i < 12; i++) {
      if (pt[i] != BFO_SENTINEL) {
        int node = pt[i];
        x[i] = xb[node];
        y[i] = yb[node];
        z[i] = zb[node];
      } else {
        nob = 0;
      }
    }
    if (!nob) return BFO_ERR_STENCIL;
    G->Aaa.x = G->Aaa.y = G->Aaa.z = 0.0;
    G->Abb.x = G->Abb.y = G->Abb.z = 0.0;
    G->Aab.x = G->Aab.y = G->Aab.z = 0.0;
    for (int i = 0; i < 12; i++) {
      G->Aaa.x += Nab_c[0][i] * x[i]; G->Aaa.y += Nab_c[0][i] * y[i]; G->Aaa.z += Nab_c[0][i] * z[i];
      G->Abb.x += Nab_c[1][i] * x[i]; G->Abb.y += Nab_c[1][i] * y[i]; G->Abb.z += Nab_c[1][i] * z[i];
      G->Aab.x += Nab_c[2][i] * x[i]; G->Aab.y += Nab_c[2][i] * y[i]; G->Aab.z += Nab_c[2][i] * z[i];
    }
    G->ndx21.x = G->ndx21.y = G->ndx21.z = 0.0;
    G->ndx31.x = G->ndx31.y = G->ndx31.z = 0.0;
    for (int i = 0; i < 12; i++) {
      G->ndx21.x += Na_c[0][i] * x[i]; G->ndx21.y += Na_c[0][i] * y[i]; G->ndx21.z += Na_c[0][i] * z[i];
      G->ndx31.x += Na_c[1][i] * x[i]; G->ndx31.y += Na_c[1][i] * y[i]; G->ndx31.z += Na_c[1][i] * z[i];
    }
    geom_tail(G);
    G->is_irregular = 0;
    G->v = 0;
    G->nen = 12;
    return BFO_OK;
  }
  if (m->ire[ec] == 1) { /* irregular patch :563-719 */
    const int v = m->val[ec];
    const int nen = v + 6;
    if (v < 3 || v + 12 > 64 || nen > m->patch_stride) return BFO_ERR_NEN;
    for (int i = 0; i < nen; i++)
      if (pt[i] == BFO_SENTINEL) return BFO_ERR_STENCIL;
    double INa0[64], INa1[64], INab0[64], INab1[64], INab2[64];
    bfo_irregular_tables(v, INa0, INa1, INab0, INab1, INab2);
    G->Aaa.x = G->Aaa.y = G->Aaa.z = 0.0;
    G->Abb.x = G->Abb.y = G->Abb.z = 0.0;
    G->Aab.x = G->Aab.y = G->Aab.z = 0.0;
    for (int i = 0; i < nen; i++) {
      const int n = pt[i];
      const double X0 = xb[n], X1 = yb[n], X2 = zb[n];
      G->Aaa.x += INab0[i] * X0; G->Aaa.y += INab0[i] * X1; G->Aaa.z += INab0[i] * X2;
      G->Abb.x += INab1[i] * X0; G->Abb.y += INab1[i] * X1; G->Abb.z += INab1[i] * X2;
      G->Aab.x += INab2[i] * X0; G->Aab.y += INab2[i] * X1; G->Aab.z += INab2[i] * X2;
    }
    G->ndx21.x = G->ndx21.y = G->ndx21.z = 0.0;
    G->ndx31.x = G->ndx31.y = G->ndx31.z = 0.0;
    for (int i = 0; i < nen; i++) {
      const int n = pt[i];
      G->ndx21.x += INa0[i] * xb[n]; G->ndx21.y += INa0[i] * yb[n]; G->ndx21.z += INa0[i] * zb[n];
      G->ndx31.x += INa1[i] * xb[n]; G->ndx31.y += INa1[i] * yb[n]; G->ndx31.z += INa1[i] * zb[n];
    }
    geom_tail(G);
    for (int i = 0; i < nen; i++) {
      G->INa0[i] = INa0[i]; G->INa1[i] = INa1[i];
      G->INab0[i] = INab0[i]; G->INab1[i] = INab1[i]; G->INab2[i] = INab2[i];
    }
    G->is_irregular = 1;
    G->v = v;
    G->nen = nen;
    return BFO_OK;
  }
  return BFO_ERR_IRE;
}

/* Compute_a3_alpha_ :745-762 */
static void a3_alpha(bfo_vec3 a1, bfo_vec3 a2, bfo_vec3 Aaa, bfo_vec3 Aab, bfo_vec3 Abb,
                     bfo_vec3 *a3_1, bfo_vec3 *a3_2) {
  double size_a3 = SIZE(CROSS(a1, a2));
  *a3_1 = PLUS(CROSS(Aaa, a2), CROSS(a1, Aab));
  *a3_2 = PLUS(CROSS(Aab, a2), CROSS(a1, Abb));
  a3_1->x /= size_a3; a3_1->y /= size_a3; a3_1->z /= size_a3;
  a3_2->x /= size_a3; a3_2->y /= size_a3; a3_2->z /= size_a3;
}

/* Per-qp state of ElemActData (include/act_variables.h:55-75). */
typedef struct {
  double gCov[3][3];  /* rows: g_i */
  double gCont[3][3]; /* rows: g^i */
  double g0Cov[3][3], g0Cont[3][3];
  double gm[3][3], gmI[3][3];   /* gm.Cov, gm.Cont */
  double gm0[3][3], gm0I[3][3]; /* gm0.Cov, gm0.Cont */
  double FaCov[3][3], FaCont[3][3], FaiCov[3][3], FaiCont[3][3];
  double CCov[3][3], CCont[3][3], CiCov[3][3], CiCont[3][3];
  double CeCov[3][3], CeCont[3][3], CeiCov[3][3], CeiCont[3][3];
  double Se[3][3], S[3][3];
  double CCe[3][3][3][3];
  double CC2222;
} qp_state;

static void vec_to_row(double r[3], bfo_vec3 v) { r[0] = v.x; r[1] = v.y; r[2] = v.z; }

/* ComputeMetricTensor :769-779, INV, ComputeContravariantBasis :784-800 */
static void metric_and_cont(const double gCov[3][3], double gm[3][3], double gmI[3][3],
                            double gCont[3][3]) {
  bfo_vec3 g[3];
  for (int i = 0; i < 3; i++) { g[i].x = gCov[i][0]; g[i].y = gCov[i][1]; g[i].z = gCov[i][2]; }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) gm[i][j] = DOT(g[i], g[j]);
  INV(gm, gmI);
  for (int i = 0; i < 3; i++) {
    double cx = 0.0, cy = 0.0, cz = 0.0;
    for (int j = 0; j < 3; j++) {
      cx += gmI[i][j] * g[j].x;
      cy += gmI[i][j] * g[j].y;
      cz += gmI[i][j] * g[j].z;
    }
    gCont[i][0] = cx; gCont[i][1] = cy; gCont[i][2] = cz;
  }
}

/* ElemUpdateG :896-982 */
static void elem_update_g(const bfo_geom *Gc, const bfo_geom *G0, double h0, const double *theta,
                          int n_qp, qp_state *Q) {
  const double hhalf = 0.5 * h0;
  bfo_vec3 a3c_1, a3c_2, a30_1, a30_2;
  a3_alpha(Gc->ndx21, Gc->ndx31, Gc->Aaa, Gc->Aab, Gc->Abb, &a3c_1, &a3c_2);
  a3_alpha(G0->ndx21, G0->ndx31, G0->Aaa, G0->Aab, G0->Abb, &a30_1, &a30_2);
  for (int qp = 0; qp < n_qp; qp++) {
    const double theta3 = hhalf * theta[qp];
    qp_state *q = &Q[qp];
    vec_to_row(q->gCov[0], PLUS(Gc->ndx21, AMULT(theta3, a3c_1)));
    vec_to_row(q->gCov[1], PLUS(Gc->ndx31, AMULT(theta3, a3c_2)));
    vec_to_row(q->gCov[2], Gc->nn);
    metric_and_cont(q->gCov, q->gm, q->gmI, q->gCont);
    vec_to_row(q->g0Cov[0], PLUS(G0->ndx21, AMULT(theta3, a30_1)));
    vec_to_row(q->g0Cov[1], PLUS(G0->ndx31, AMULT(theta3, a30_2)));
    vec_to_row(q->g0Cov[2], G0->nn);
    metric_and_cont(q->g0Cov, q->gm0, q->gm0I, q->g0Cont);
  }
}

/* ElemActDefGrad :1196-1329 */
static void elem_act_def_grad(const bfo_mesh *m, const bfo_params *p, int ec, int n_qp,
                              qp_state *Q) {
  double Fa_cart[3][3] = {{0.0}}, Finv_cart[3][3], S[3][3], ST[3][3], tmp[3][3];
  double delay = m->activation_delay ? m->activation_delay[ec] : 0.0;
  double gamma = bfo_gamma_of_time(p, delay);
  gamma *= m->gamma_scale[ec];
  if (p->cart_fib_act) {
    double nf[3] = {m->n_fib[3 * ec + 0], m->n_fib[3 * ec + 1], m->n_fib[3 * ec + 2]};
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) Fa_cart[i][j] = (i == j ? 1.0 : 0.0) - gamma * nf[i] * nf[j];
  } else if (p->manufactured) {
    double t = p->ti * p->dt;
    gamma = p->manufactured_gamma0 * pow(sin(PI_ * t / p->manufactured_T), 2.0);
    double lambda = 1.0 - gamma;
    Fa_cart[0][0] = lambda;
    Fa_cart[1][1] = 1.0;
    Fa_cart[2][2] = 1.0;
  } else {
    Fa_cart[0][0] = 1.0 - gamma;
    Fa_cart[1][1] = 1.0 - gamma;
    Fa_cart[2][2] = 1.0 / (Fa_cart[0][0] * Fa_cart[1][1]);
  }
  for (int qp = 0; qp < n_qp; qp++) {
    qp_state *q = &Q[qp];
    for (int j = 0; j < 3; j++) {
      S[0][j] = q->g0Cov[j][0];
      S[1][j] = q->g0Cov[j][1];
      S[2][j] = q->g0Cov[j][2];
    }
    TRANS(S, ST);
    MATMULT(ST, Fa_cart, tmp);
    MATMULT(tmp, S, q->FaCov);
    INV(Fa_cart, Finv_cart);
    MATMULT(ST, Finv_cart, tmp);
    MATMULT(tmp, S, q->FaiCov);
    Raise2(q->gm0I, q->FaCov, q->FaCont);
    Raise2(q->gm0I, q->FaiCov, q->FaiCont);
  }
}

/* ElemCGDefTens :1332-1378 */
static void elem_cg_def_tens(int n_qp, qp_state *Q) {
  for (int qp = 0; qp < n_qp; qp++) {
    qp_state *q = &Q[qp];
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) {
        q->CCov[i][j] = q->gm[i][j];
        q->CiCont[i][j] = q->gmI[i][j];
      }
    for (int a = 0; a < 2; a++) {
      q->CCov[a][2] = 0.0;
      q->CCov[2][a] = 0.0;
      q->CiCont[a][2] = 0.0;
      q->CiCont[2][a] = 0.0;
    }
    q->CCov[2][2] = 1.0;
    q->CiCont[2][2] = 1.0;
    Raise2(q->gmI, q->CCov, q->CCont);
    Lower2(q->gm, q->CiCont, q->CiCov);
  }
}

/* ElemElasCGDefTens :1381-1434 */
static void elem_elas_cg(qp_state *q) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) q->CeiCont[i][j] = 0.0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int mm = 0; mm < 3; mm++)
        for (int pp = 0; pp < 3; pp++)
          for (int qq = 0; qq < 3; qq++)
            for (int l = 0; l < 3; l++)
              q->CeiCont[i][j] += q->FaCont[i][mm] * q->gm0[mm][pp] * q->CiCont[pp][qq] *
                                  q->FaCont[l][j] * q->gm0[l][qq];
  Lower2(q->gm0, q->CeiCont, q->CeiCov);
  INV(q->CeiCov, q->CeCont);
  Lower2(q->gm0, q->CeCont, q->CeCov);
}

/* ElemElasStress :1437-1491 */
static void elem_elas_stress(const bfo_params *p, qp_state *q) {
  double detCe = DET3x3(q->CeCov);
  double detG0 = DET3x3(q->gm0);
  double Je = sqrt(detCe / detG0);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double term1 = p->mu * (q->gm0I[i][j] - q->CeiCont[i][j]);
      double term2 = p->K * q->CeiCont[i][j] * (Je * Je - Je);
      q->Se[i][j] = term1 + term2;
    }
}

/* ElemTotStress :1494-1532 */
static void elem_tot_stress(qp_state *q) {
  const double (*Fi)[3] = (const double (*)[3])q->FaiCov;
  const double (*G)[3] = (const double (*)[3])q->gm0I;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) q->S[i][j] = 0.0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int w = 0; w < 3; w++)
        for (int pp = 0; pp < 3; pp++)
          for (int z = 0; z < 3; z++)
            for (int s = 0; s < 3; s++)
              q->S[i][j] += 0.5 * Fi[w][pp] * Fi[z][s] * G[s][j] *
                            (q->Se[pp][z] * G[w][i] + q->Se[pp][i] * G[w][z]);
}

/* ElemElsTangMatTens :1535-1589 */
static void elem_els_tang(const bfo_params *p, qp_state *q) {
  double detCe = DET3x3(q->CeCov);
  double detG0 = DET3x3(q->gm0);
  double Je = sqrt(detCe / detG0);
  const double Je2 = Je * Je;
  const double alpha = p->K * (Je2 - Je) - p->mu;
  const double beta = 0.5 * p->K * Je * (2.0 * Je - 1.0);
  const double (*Cb)[3] = (const double (*)[3])q->CeiCont;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++)
        for (int l = 0; l < 3; l++)
          q->CCe[i][j][k][l] = (2.0 * beta) * Cb[i][j] * Cb[k][l] -
                               alpha * (Cb[i][k] * Cb[l][j] + Cb[i][l] * Cb[k][j]);
}

/* ElemTotTangMatTens :1592-1648, only [2][2][2][2] */
static void elem_tot_tang(qp_state *q) {
  const double (*Fi)[3] = (const double (*)[3])q->FaiCov;
  const double (*G)[3] = (const double (*)[3])q->gm0I;
  const int i = 2, j = 2, k = 2, l = 2;
  double acc = 0.0;
  for (int g = 0; g < 3; g++)
    for (int h = 0; h < 3; h++)
      for (int mm = 0; mm < 3; mm++)
        for (int n = 0; n < 3; n++)
          for (int pp = 0; pp < 3; pp++)
            for (int s = 0; s < 3; s++)
              for (int w = 0; w < 3; w++)
                for (int z = 0; z < 3; z++)
                  acc += 0.25 * Fi[w][pp] * Fi[z][s] * Fi[g][h] * Fi[n][mm] * G[s][j] *
                         G[mm][l] *
                         (G[g][k] * (q->CCe[h][n][pp][z] * G[w][i] +
                                     q->CCe[h][n][pp][i] * G[w][z]) +
                          G[g][n] * (q->CCe[h][k][pp][z] * G[w][i] +
                                     q->CCe[h][k][pp][i] * G[w][z]));
  q->CC2222 = acc;
}

/* ModElemC33 :1654-1721.  Returns BFO_WARN_CC3333 at the first qp whose tangent fails the
 * guard; as in the reference (PetscCheck returns early, :1679), that qp and the rest of
 * this iteration's qps keep their C33. */
static int mod_elem_c33(int n_qp, qp_state *Q) {
  for (int qp = 0; qp < n_qp; qp++) {
    qp_state *q = &Q[qp];
    double S33 = q->S[2][2];
    double CC3333 = q->CC2222;
    double denom_tol = 1e-14;
    if (!(fabs(CC3333) > denom_tol)) return BFO_WARN_CC3333;
    double DeltaC33 = -1.0 * S33 / CC3333;
    double C33_old = q->CCov[2][2];
    q->CCov[2][2] = C33_old + DeltaC33;
    if (DET3x3(q->CCov) <= 0.0) {
      double lo = 0.0, hi = 1.0;
      for (int bk = 0; bk < 16; bk++) {
        double mid = 0.5 * (lo + hi);
        q->CCov[2][2] = C33_old + mid * DeltaC33;
        if (DET3x3(q->CCov) > 0.0) lo = mid; else hi = mid;
      }
      DeltaC33 *= lo;
      q->CCov[2][2] = C33_old + DeltaC33;
    }
  }
  return BFO_OK;
}

/* ElemC33Solve :2048-2138 (sub-step return codes ignored, as at :2060-2086) */
static int elem_c33_solve(const bfo_params *p, int n_qp, qp_state *Q) {
  int warn = 0;
  for (int it = 0; it < p->c33_iters; it++) {
    for (int qp = 0; qp < n_qp; qp++) elem_elas_cg(&Q[qp]);
    for (int qp = 0; qp < n_qp; qp++) elem_elas_stress(p, &Q[qp]);
    for (int qp = 0; qp < n_qp; qp++) elem_tot_stress(&Q[qp]);
    for (int qp = 0; qp < n_qp; qp++) elem_els_tang(p, &Q[qp]);
    for (int qp = 0; qp < n_qp; qp++) elem_tot_tang(&Q[qp]);
    if (mod_elem_c33(n_qp, Q) != BFO_OK) warn++;
    for (int qp = 0; qp < n_qp; qp++) {
      qp_state *q = &Q[qp];
      Raise2(q->gm0I, q->CCov, q->CCont);
      INV(q->CCov, q->CiCont);
      Lower2(q->gm0, q->CiCont, q->CiCov);
    }
  }
  return warn;
}

/* ElemUpdFint :1762-1995 */
static void elem_upd_fint(const bfo_mesh *m, const bfo_params *p, int ec, const bfo_geom *G,
                          const double *theta, const double *wq, int n_qp, const qp_state *Q,
                          double *Fb) {
  const bfo_vec3 ndx21 = G->ndx21, ndx31 = G->ndx31, nn = G->nn, gc1 = G->gc1, gc2 = G->gc2;
  const bfo_vec3 Aaa = G->Aaa, Abb = G->Abb, Aab = G->Aab;
  double A0 = m->dA0[ec];
  const double half_h0 = 0.5 * p->h0;
  const double pref0 = 0.5 * p->h0 * A0;
  if (m->n_elmt_base > 0 && ec >= m->n_elmt_base) { /* cap CST :1807-1854 */
    const int *pt = m->patch + (size_t)m->patch_stride * ec;
    const int tri[3] = {m->nv1[ec], m->nv2[ec], m->nv3[ec]};
    const double dN[3][2] = {{-1.0, -1.0}, {1.0, 0.0}, {0.0, 1.0}};
    int slot[3] = {-1, -1, -1};
    const int nslot = m->val[ec] + 6;
    for (int a = 0; a < 3; a++)
      for (int i = 0; i < nslot; i++)
        if (pt[i] == tri[a]) { slot[a] = i; break; }
    for (int qp = 0; qp < n_qp; qp++) {
      const double w_qp = wq[qp];
      double Sv[3] = {Q[qp].S[0][0], Q[qp].S[1][1], Q[qp].S[0][1]};
      for (int a = 0; a < 3; a++) {
        if (slot[a] < 0) continue;
        const int base = 3 * slot[a];
        const double Na0 = dN[a][0], Na1 = dN[a][1];
        const double Bm0x = Na0 * ndx21.x, Bm0y = Na0 * ndx21.y, Bm0z = Na0 * ndx21.z;
        const double Bm1x = Na1 * ndx31.x, Bm1y = Na1 * ndx31.y, Bm1z = Na1 * ndx31.z;
        const double Bm2x = Na0 * ndx31.x + Na1 * ndx21.x;
        const double Bm2y = Na0 * ndx31.y + Na1 * ndx21.y;
        const double Bm2z = Na0 * ndx31.z + Na1 * ndx21.z;
        Fb[base + 0] += pref0 * w_qp * (Bm0x * Sv[0] + Bm1x * Sv[1] + Bm2x * Sv[2]);
        Fb[base + 1] += pref0 * w_qp * (Bm0y * Sv[0] + Bm1y * Sv[1] + Bm2y * Sv[2]);
        Fb[base + 2] += pref0 * w_qp * (Bm0z * Sv[0] + Bm1z * Sv[1] + Bm2z * Sv[2]);
      }
    }
    return;
  }
  for (int qp = 0; qp < n_qp; qp++) {
    const double w_qp = wq[qp];
    const double xi_qp = theta[qp];
    double Sv[3] = {Q[qp].S[0][0], Q[qp].S[1][1], Q[qp].S[0][1]};
    if (m->ire[ec] == 0) {
      double Bm[3][36], Bs[3][36];
      for (int a = 0; a < 12; a++) {
        Bm[0][3 * a + 0] = Na_c[0][a] * ndx21.x; Bm[0][3 * a + 1] = Na_c[0][a] * ndx21.y; Bm[0][3 * a + 2] = Na_c[0][a] * ndx21.z;
        Bm[1][3 * a + 0] = Na_c[1][a] * ndx31.x; Bm[1][3 * a + 1] = Na_c[1][a] * ndx31.y; Bm[1][3 * a + 2] = Na_c[1][a] * ndx31.z;
        Bm[2][3 * a + 0] = Na_c[0][a] * ndx31.x + Na_c[1][a] * ndx21.x;
        Bm[2][3 * a + 1] = Na_c[0][a] * ndx31.y + Na_c[1][a] * ndx21.y;
        Bm[2][3 * a + 2] = Na_c[0][a] * ndx31.z + Na_c[1][a] * ndx21.z;
      }
      for (int a = 0; a < 12; a++) {
        const bfo_vec3 ng = PLUS(AMULT(Na_c[0][a], gc1), AMULT(Na_c[1][a], gc2));
        const double b1 = DOT(ng, Aaa), b2 = DOT(ng, Abb), b3 = DOT(ng, Aab);
        Bs[0][3 * a + 0] = -(Nab_c[0][a] + b1) * nn.x; Bs[0][3 * a + 1] = -(Nab_c[0][a] + b1) * nn.y; Bs[0][3 * a + 2] = -(Nab_c[0][a] + b1) * nn.z;
        Bs[1][3 * a + 0] = -(Nab_c[1][a] + b2) * nn.x; Bs[1][3 * a + 1] = -(Nab_c[1][a] + b2) * nn.y; Bs[1][3 * a + 2] = -(Nab_c[1][a] + b2) * nn.z;
        Bs[2][3 * a + 0] = -2.0 * (Nab_c[2][a] + b3) * nn.x; Bs[2][3 * a + 1] = -2.0 * (Nab_c[2][a] + b3) * nn.y; Bs[2][3 * a + 2] = -2.0 * (Nab_c[2][a] + b3) * nn.z;
      }
      for (int ii = 0; ii < 36; ii++) {
        double mem = 0.0, bend = 0.0;
        for (int mm = 0; mm < 3; mm++) {
          mem += Bm[mm][ii] * Sv[mm];
          bend += Bs[mm][ii] * Sv[mm];
        }
        Fb[ii] += pref0 * w_qp * (mem + bend * (half_h0 * xi_qp));
      }
      continue;
    }
    /* irregular :1922-1987 */
    const int nen = G->nen;
    for (int a = 0; a < nen; a++) {
      double Bm0x = G->INa0[a] * ndx21.x, Bm0y = G->INa0[a] * ndx21.y, Bm0z = G->INa0[a] * ndx21.z;
      double Bm1x = G->INa1[a] * ndx31.x, Bm1y = G->INa1[a] * ndx31.y, Bm1z = G->INa1[a] * ndx31.z;
      double Bm2x = G->INa0[a] * ndx31.x + G->INa1[a] * ndx21.x;
      double Bm2y = G->INa0[a] * ndx31.y + G->INa1[a] * ndx21.y;
      double Bm2z = G->INa0[a] * ndx31.z + G->INa1[a] * ndx21.z;
      const bfo_vec3 ng = PLUS(AMULT(G->INa0[a], gc1), AMULT(G->INa1[a], gc2));
      const double b1 = DOT(ng, Aaa), b2 = DOT(ng, Abb), b3 = DOT(ng, Aab);
      double Bs0 = -(G->INab0[a] + b1);
      double Bs1 = -(G->INab1[a] + b2);
      double Bs2 = -2.0 * (G->INab2[a] + b3);
      const int base = 3 * a;
      {
        double mem = Bm0x * Sv[0] + Bm1x * Sv[1] + Bm2x * Sv[2];
        double bend = (Bs0 * nn.x) * Sv[0] + (Bs1 * nn.x) * Sv[1] + (Bs2 * nn.x) * Sv[2];
        Fb[base + 0] += pref0 * w_qp * (mem + bend * (half_h0 * xi_qp));
      }
      {
        double mem = Bm0y * Sv[0] + Bm1y * Sv[1] + Bm2y * Sv[2];
        double bend = (Bs0 * nn.y) * Sv[0] + (Bs1 * nn.y) * Sv[1] + (Bs2 * nn.y) * Sv[2];
        Fb[base + 1] += pref0 * w_qp * (mem + bend * (half_h0 * xi_qp));
      }
      {
        double mem = Bm0z * Sv[0] + Bm1z * Sv[1] + Bm2z * Sv[2];
        double bend = (Bs0 * nn.z) * Sv[0] + (Bs1 * nn.z) * Sv[1] + (Bs2 * nn.z) * Sv[2];
        Fb[base + 2] += pref0 * w_qp * (mem + bend * (half_h0 * xi_qp));
      }
    }
  }
}

static int elem_nen(const bfo_mesh *m, int ec) { /* active_strain.c:2245 */
  return (m->ire[ec] == 0) ? 12 : (m->val[ec] + 6);
}

int bfo_element_force(const bfo_mesh *m, const bfo_params *p, const bfo_geom *G0,
                      const bfo_geom *Gc, int ec, double *Fb, double *S_out, double *C33_out,
                      int *n_warn) {
  double theta[5], wq[5];
  if (bfo_gauss(p->n_qp, theta, wq) != BFO_OK) return BFO_ERR_NQP;
  const int n_qp = p->n_qp;
  qp_state Q[5];
  memset(Q, 0, sizeof(Q));
  elem_update_g(Gc, G0, p->h0, theta, n_qp, Q);
  elem_act_def_grad(m, p, ec, n_qp, Q);
  elem_cg_def_tens(n_qp, Q);
  int w = elem_c33_solve(p, n_qp, Q);
  if (n_warn) *n_warn += w;
  for (int i = 0; i < 3 * 64; i++) Fb[i] = 0.0;
  elem_upd_fint(m, p, ec, Gc, theta, wq, n_qp, Q, Fb);
  if (S_out)
    for (int qp = 0; qp < n_qp; qp++) memcpy(S_out + 9 * qp, Q[qp].S, 9 * sizeof(double));
  if (C33_out)
    for (int qp = 0; qp < n_qp; qp++) C33_out[qp] = Q[qp].CCov[2][2];
  return BFO_OK;
}

static void assemble(const bfo_mesh *m, int ec, const double *Fb, double *fint) {
  const int nen = elem_nen(m, ec);
  const int *pt = m->patch + (size_t)m->patch_stride * ec;
  for (int i = 0; i < nen; i++) { /* :2252-2260 (gdof0 = 3*node on one rank) */
    int node = pt[i];
    if (node == BFO_SENTINEL || node < 0 || node >= m->n_v) continue;
    fint[3 * node + 0] += Fb[3 * i + 0];
    fint[3 * node + 1] += Fb[3 * i + 1];
    fint[3 * node + 2] += Fb[3 * i + 2];
  }
}

/* geom0 cache (FEM_DMPlexGeomSetup, dmplex_geom.c:859-865): one bfo_geom per element. */
void *bfo_geom0_create(const bfo_mesh *m, bfo_status *st) {
  bfo_status loc = {BFO_OK, -1, 0};
  const int E = m->n_elmt;
  bfo_geom *G0 = (bfo_geom *)malloc(sizeof(bfo_geom) * (size_t)(E > 0 ? E : 1));
  if (!G0) return NULL;
  for (int ec = 0; ec < E; ec++) {
    int rc = bfo_geom_subdiv(m, ec, m->x0, m->y0, m->z0, &G0[ec]);
    if (rc != BFO_OK && loc.status == BFO_OK) { loc.status = rc; loc.elem = ec; }
  }
  if (st) *st = loc;
  return G0;
}

void bfo_geom0_free(void *g0) { free(g0); }

int bfo_finternal_act_cached(const bfo_mesh *m, const bfo_params *p, const void *g0,
                             const double *x, const double *y, const double *z, double *fint,
                             int accumulate, double *c33, bfo_status *st, int nthreads) {
  bfo_status loc = {BFO_OK, -1, 0};
  if (!m || !p || !g0 || !x || !y || !z || !fint) return BFO_ERR_ARG;
  if (p->n_qp < 1 || p->n_qp > 5) { loc.status = BFO_ERR_NQP; if (st) *st = loc; return loc.status; }
  const bfo_geom *G0 = (const bfo_geom *)g0;
  const int E = m->n_elmt;
  if (!accumulate)
    for (int i = 0; i < 3 * m->n_nodes; i++) fint[i] = 0.0;
  const int maxnen = 64;
  double *Fb = (double *)malloc(sizeof(double) * 3 * maxnen * (size_t)(E > 0 ? E : 1));
  int *rcs = (int *)calloc((size_t)(E > 0 ? E : 1), sizeof(int));
  int *warns = (int *)calloc((size_t)(E > 0 ? E : 1), sizeof(int));
  const int q = p->n_qp;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
#endif
  for (int ec = 0; ec < E; ec++) {
    bfo_geom Gc;
    int rc = bfo_geom_subdiv(m, ec, x, y, z, &Gc);
    if (rc != BFO_OK) { rcs[ec] = rc; continue; }
    double C33[5];
    bfo_element_force(m, p, &G0[ec], &Gc, ec, Fb + (size_t)3 * maxnen * ec, NULL, C33,
                      &warns[ec]);
    if (c33)
      for (int k = 0; k < q; k++) c33[(size_t)ec * q + k] = C33[k];
  }
  (void)nthreads;
  /* element-order assembly (active_strain.c:2238-2260) */
  for (int ec = 0; ec < E; ec++) {
    if (rcs[ec] != BFO_OK) {
      if (loc.status == BFO_OK) { loc.status = rcs[ec]; loc.elem = ec; }
      continue;
    }
    if (warns[ec]) {
      loc.n_warn += warns[ec];
      if (loc.elem < 0) loc.elem = ec;
    }
    assemble(m, ec, Fb + (size_t)3 * maxnen * ec, fint);
  }
  free(Fb);
  free(rcs);
  free(warns);
  if (st) *st = loc;
  return loc.status;
}

static int run_path(const bfo_mesh *m, const bfo_params *p, const double *x, const double *y,
                    const double *z, double *fint, int accumulate, double *c33, bfo_status *st,
                    int nthreads) {
  if (!m || !p || !x || !y || !z || !fint) return BFO_ERR_ARG;
  bfo_status g = {BFO_OK, -1, 0};
  void *G0 = bfo_geom0_create(m, &g);
  if (!G0) return BFO_ERR_ARG;
  if (g.status != BFO_OK) {
    if (!accumulate)
      for (int i = 0; i < 3 * m->n_nodes; i++) fint[i] = 0.0;
    bfo_geom0_free(G0);
    if (st) *st = g;
    return g.status;
  }
  int rc = bfo_finternal_act_cached(m, p, G0, x, y, z, fint, accumulate, c33, st, nthreads);
  bfo_geom0_free(G0);
  return rc;
}

int bfo_finternal_act(const bfo_mesh *m, const bfo_params *p, const double *x, const double *y,
                      const double *z, double *fint, int accumulate, double *c33,
                      bfo_status *st) {
  return run_path(m, p, x, y, z, fint, accumulate, c33, st, 1);
}

int bfo_finternal_act_omp(const bfo_mesh *m, const bfo_params *p, const double *x,
                          const double *y, const double *z, double *fint, int accumulate,
                          double *c33, bfo_status *st, int nthreads) {
  return run_path(m, p, x, y, z, fint, accumulate, c33, st, nthreads);
}

/* ------------------------------------------------------------ topology --------------- */
void bfo_area_normal(int n_tri, const int *nv1, const int *nv2, const int *nv3,
                     const double *x, const double *y, const double *z, double *dA) {
  for (int ec = 0; ec < n_tri; ec++) { /* io.c:1651-1666 */
    int n1 = nv1[ec], n2 = nv2[ec], n3 = nv3[ec];
    bfo_vec3 x1 = {x[n1], y[n1], z[n1]}, x2 = {x[n2], y[n2], z[n2]}, x3 = {x[n3], y[n3], z[n3]};
    bfo_vec3 d21 = {x2.x - x1.x, x2.y - x1.y, x2.z - x1.z};
    bfo_vec3 d31 = {x3.x - x1.x, x3.y - x1.y, x3.z - x1.z};
    bfo_vec3 c = CROSS(d21, d31);
    dA[ec] = 0.5 * SIZE(c);
  }
}

void bfo_irrver(int n_v, int n_elmt, int n_elmt_base, const int *nv1, const int *nv2,
                const int *nv3, int sum_n_bnodes, const int *bnodes, int *ire, int *irv,
                int *val) {
  int *irr = (int *)malloc(sizeof(int) * (size_t)(n_v > 0 ? n_v : 1));
  int *elmt = (int *)malloc(sizeof(int) * (size_t)(n_elmt > 0 ? n_elmt : 1));
  for (int ec = 0; ec < n_elmt; ec++) { ire[ec] = 0; irv[ec] = 0; val[ec] = 6; }
  int ec_limit = (n_elmt_base > 0) ? n_elmt_base : n_elmt;
  for (int nc = 0; nc < n_v; nc++) { /* bending.c:1212-1252 */
    irr[nc] = 0;
    int count = 0, bcount = 0, is_bnode = 0;
    for (int i = 0; i < sum_n_bnodes; i++)
      if (nc == bnodes[i]) { is_bnode = 1; break; }
    for (int ec = 0; ec < ec_limit; ec++)
      if (nc == nv1[ec] || nc == nv2[ec] || nc == nv3[ec]) elmt[count++] = ec;
    if (is_bnode) bcount = count + 3;
    if (bcount == 0) {
      if (count != 6) {
        irr[nc] = 1;
        for (int i = 0; i < count; i++) { ire[elmt[i]] = 1; val[elmt[i]] = count; }
      }
    } else if (bcount != 6) {
      irr[nc] = 1;
      for (int i = 0; i < count; i++) { ire[elmt[i]] = 1; val[elmt[i]] = bcount; }
    }
  }
  for (int ec = 0; ec < n_elmt; ec++) /* :1254-1262 */
    for (int nc = 0; nc < n_v; nc++)
      if (irr[nc] == 1) {
        if (nc == nv1[ec]) irv[ec] = 1;
        if (nc == nv2[ec]) irv[ec] = 2;
        if (nc == nv3[ec]) irv[ec] = 3;
      }
  free(irr);
  free(elmt);
}

/* One triangle-scan step of Patch: if triangle (n1,n2,n3) holds both a and b, return its
 * first vertex (in n1,n2,n3 order) that is none of ex0, ex1, ex2 and passes the `ok` flag;
 * this is the if/else-if chain repeated throughout bending.c:1297-1640. */
static int pick(int n1, int n2, int n3, int a, int b, int ex0, int ex1, int ex2, int ok,
                int *out) {
  int mcount = 0;
  if (a == n1 || a == n2 || a == n3) mcount++;
  if (b == n1 || b == n2 || b == n3) mcount++;
  if (mcount != 2) return 0;
  if (n1 != ex0 && n1 != ex1 && n1 != ex2 && ok) { *out = n1; return 1; }
  if (n2 != ex0 && n2 != ex1 && n2 != ex2 && ok) { *out = n2; return 1; }
  if (n3 != ex0 && n3 != ex1 && n3 != ex2 && ok) { *out = n3; return 1; }
  return 0;
}

void bfo_patch(int n_v, int n_elmt, int n_elmt_base, int n_tri, const int *nv1, const int *nv2,
               const int *nv3, const int *ire, const int *irv, const int *val, int stride,
               int *patch) {
  for (int ec = 0; ec < n_elmt; ec++) {
    const int v = val[ec];
    int *P = patch + (size_t)stride * ec;
    if (n_elmt_base > 0 && ec >= n_elmt_base) { /* :1281-1294 */
      for (int i = 0; i < stride; i++) P[i] = BFO_SENTINEL;
      P[3] = nv1[ec];
      P[6] = nv2[ec];
      P[7] = nv3[ec];
      continue;
    }
    int p[64];
    for (int i = 0; i < v + 6 && i < 64; i++) p[i] = BFO_SENTINEL;
    if (ire[ec] == 0) { /* :1300-1430 */
      p[3] = nv1[ec]; p[6] = nv2[ec]; p[7] = nv3[ec];
      for (int i = 0; i < n_tri; i++) {
        int a = nv1[i], b = nv2[i], c = nv3[i], ok = (i != ec);
        pick(a, b, c, p[3], p[6], p[3], p[6], p[6], ok, &p[2]);
        pick(a, b, c, p[3], p[7], p[3], p[7], p[7], ok, &p[4]);
        pick(a, b, c, p[6], p[7], p[6], p[7], p[7], ok, &p[10]);
      }
      for (int i = 0; i < n_tri; i++) {
        int a = nv1[i], b = nv2[i], c = nv3[i];
        pick(a, b, c, p[3], p[2], p[3], p[2], p[6], 1, &p[0]);
        pick(a, b, c, p[3], p[4], p[3], p[4], p[7], 1, &p[1]);
        pick(a, b, c, p[6], p[2], p[6], p[2], p[3], 1, &p[5]);
        pick(a, b, c, p[6], p[10], p[6], p[10], p[7], 1, &p[9]);
        pick(a, b, c, p[7], p[10], p[7], p[10], p[6], 1, &p[11]);
        pick(a, b, c, p[7], p[4], p[7], p[4], p[3], 1, &p[8]);
      }
    } else if (ire[ec] == 1) { /* :1432-1620 */
      if (irv[ec] == 1) { p[0] = nv1[ec]; p[1] = nv2[ec]; p[v] = nv3[ec]; }
      else if (irv[ec] == 2) { p[0] = nv2[ec]; p[1] = nv3[ec]; p[v] = nv1[ec]; }
      else if (irv[ec] == 3) { p[0] = nv3[ec]; p[1] = nv1[ec]; p[v] = nv2[ec]; }
      for (int i = 0; i < n_tri; i++) {
        int a = nv1[i], b = nv2[i], c = nv3[i], ok = (i != ec);
        pick(a, b, c, p[0], p[1], p[0], p[1], p[1], ok, &p[2]);
        pick(a, b, c, p[0], p[v], p[0], p[v], p[v], ok, &p[v - 1]);
        pick(a, b, c, p[1], p[v], p[1], p[v], p[v], ok, &p[v + 1]);
      }
      for (int i = 0; i < n_tri; i++) {
        int a = nv1[i], b = nv2[i], c = nv3[i];
        pick(a, b, c, p[0], p[2], p[0], p[2], p[1], 1, &p[3]);
        pick(a, b, c, p[0], p[v - 1], p[0], p[v - 1], p[v], 1, &p[v - 2]);
        pick(a, b, c, p[1], p[2], p[1], p[2], p[0], 1, &p[v + 3]);
        pick(a, b, c, p[1], p[v + 1], p[1], p[v + 1], p[v], 1, &p[v + 2]);
        pick(a, b, c, p[v], p[v + 1], p[v], p[v + 1], p[1], 1, &p[v + 4]);
        pick(a, b, c, p[v], p[v - 1], p[v], p[v - 1], p[0], 1, &p[v + 5]);
      }
      if (val[ec] > 6) { /* :1622-1660 */
        for (int k = 0; k < v - 6; k++)
          for (int i = 0; i < n_tri; i++) {
            int a = nv1[i], b = nv2[i], c = nv3[i];
            pick(a, b, c, p[0], p[3 + k], p[0], p[3 + k], p[2 + k], 1, &p[k + 4]);
            pick(a, b, c, p[0], p[v - 2 - k], p[0], p[v - 2 - k], p[v - 1 - k], 1, &p[v - 3 - k]);
          }
      }
    }
    /* sentinel fix-up :1662-1680 */
    int fixup = -1, ghost_fixup = -1;
    for (int i = 0; i < v + 6 && i < stride; i++) {
      if (p[i] == BFO_SENTINEL) continue;
      if (fixup < 0) fixup = p[i];
      if (p[i] >= n_v) { ghost_fixup = p[i]; break; }
    }
    if (ghost_fixup >= 0) fixup = ghost_fixup;
    if (fixup >= 0)
      for (int i = 0; i < v + 6 && i < stride; i++)
        if (p[i] == BFO_SENTINEL) p[i] = fixup;
    for (int i = 0; i < stride; i++) P[i] = BFO_SENTINEL;
    for (int i = 0; i < v + 6 && i < stride; i++) P[i] = p[i];
  }
}

static void ghost_ring_walk(int n_v, int n_elmt, int n_elmt_base, int n_edge, const int *n_bnodes,
                            const int *bnodes, int *nv1, int *nv2, int *nv3, double *x,
                            double *y, double *z, double *x0, double *y0, double *z0,
                            int init) {
  int count = 0;
  for (int edge_n = 0; edge_n < n_edge; edge_n++) { /* bending.c:1674-1756 */
    int end = 0;
    for (int e = 0; e < edge_n + 1; e++) end += n_bnodes[e];
    int start = end - n_bnodes[edge_n];
    int n_local = end - start;
    int ec_limit = (edge_n == 0 && n_elmt_base > 0) ? n_elmt_base : n_elmt;
    int ring_first_ghost = count;
    for (int local_i = 0; local_i < n_local; local_i++) {
      int i = start + local_i;
      int i_next = start + (local_i + 1) % n_local;
      int a = bnodes[i_next], b = bnodes[i], c = -1;
      for (int ec = 0; ec < ec_limit; ec++) {
        int catch_ = 0;
        if (a == nv1[ec] || a == nv2[ec] || a == nv3[ec]) catch_++;
        if (b == nv1[ec] || b == nv2[ec] || b == nv3[ec]) catch_++;
        if (catch_ == 2) {
          if (nv1[ec] != a && nv1[ec] != b) c = nv1[ec];
          else if (nv2[ec] != a && nv2[ec] != b) c = nv2[ec];
          else c = nv3[ec];
        }
      }
      if (c < 0) c = a; /* reference leaves nv3 uninitialised here; never hit on closed rings */
      double gx = x[b] + x[a] - x[c], gy = y[b] + y[a] - y[c], gz = z[b] + z[a] - z[c];
      x[n_v + count] = gx; y[n_v + count] = gy; z[n_v + count] = gz;
      if (init) {
        x0[n_v + count] = gx; y0[n_v + count] = gy; z0[n_v + count] = gz;
        nv2[n_elmt + 2 * count] = b;
        nv1[n_elmt + 2 * count] = a;
        nv3[n_elmt + 2 * count] = n_v + count;
        int ghost_prev = (local_i == 0) ? (n_v + ring_first_ghost + n_local - 1) : (n_v + count - 1);
        nv1[n_elmt + 2 * count + 1] = n_v + count;
        nv2[n_elmt + 2 * count + 1] = b;
        nv3[n_elmt + 2 * count + 1] = ghost_prev;
      }
      count++;
    }
  }
}

void bfo_global_ghost_init(int n_v, int n_elmt, int n_elmt_base, int n_edge, const int *n_bnodes,
                           const int *bnodes, int *nv1, int *nv2, int *nv3, double *x, double *y,
                           double *z, double *x0, double *y0, double *z0) {
  ghost_ring_walk(n_v, n_elmt, n_elmt_base, n_edge, n_bnodes, bnodes, nv1, nv2, nv3, x, y, z,
                  x0, y0, z0, 1);
}

void bfo_global_ghost(int n_v, int n_elmt, int n_elmt_base, int n_edge, const int *n_bnodes,
                      const int *bnodes, const int *nv1, const int *nv2, const int *nv3,
                      double *x, double *y, double *z) {
  ghost_ring_walk(n_v, n_elmt, n_elmt_base, n_edge, n_bnodes, bnodes, (int *)nv1, (int *)nv2,
                  (int *)nv3, x, y, z, NULL, NULL, NULL, 0);
}
