// bf_exact.h -- per-element active-strain force in the reference's operation order.
//
// Every sum, product and inverse below is evaluated in the order written in
// src/active_strain.c (ElemUpdateG :896-982, ElemActDefGrad :1196-1329, ElemCGDefTens
// :1332-1378, ElemC33Solve :2048-2138 with its five sub-steps :1381-1721, ElemUpdFint
// :1762-1995) and src/math.c.  Compiled without floating-point contraction, it rounds
// exactly as the reference's C does, so this path reproduces the CPU result bit for bit
// (up to the libm calls of the activation ramp).  It is the plain FP64 path; the staged
// arithmetic of bf_elem.h is the fast path and states its error against this one.
//
// Differences from the reference that do not change any rounded value:
//  * the qp loop is outermost (the reference loops qp inside each C33 iteration); qps
//    do not interact except through the CC3333 guard, handled by `skip_mask` as below;
//  * CCe entries are recomputed where the 8-deep CC^{2222} sum reads them instead of
//    being stored (same expression, same rounding);
//  * quantities no force depends on (C.Cont, C_inv.Cov, Fa_inv.Cont, g^i) are skipped.
#pragma once

#include "bf_elem.h"

namespace bf {
namespace exact {

struct Mat {
  double m[3][3];
};

BF_HD void matmult(const Mat& A, const Mat& B, Mat& C) {  // src/math.c:95-120
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double s = 0.0;
      for (int k = 0; k < 3; k++) s += A.m[i][k] * B.m[k][j];
      C.m[i][j] = s;
    }
}
BF_HD void inv(const Mat& T, Mat& Ti) {  // src/math.c:77-91
  const double T11 = T.m[0][0], T12 = T.m[0][1], T13 = T.m[0][2];
  const double T21 = T.m[1][0], T22 = T.m[1][1], T23 = T.m[1][2];
  const double T31 = T.m[2][0], T32 = T.m[2][1], T33 = T.m[2][2];
  const double det = T11 * (T33 * T22 - T32 * T23) - T21 * (T33 * T12 - T32 * T13) +
                     T31 * (T23 * T12 - T22 * T13);
  const double c = 1 / det;
  Ti.m[0][0] = c * (T33 * T22 - T32 * T23);
  Ti.m[0][1] = -c * (T33 * T12 - T32 * T13);
  Ti.m[0][2] = c * (T23 * T12 - T22 * T13);
  Ti.m[1][0] = -c * (T33 * T21 - T31 * T23);
  Ti.m[1][1] = c * (T33 * T11 - T31 * T13);
  Ti.m[1][2] = -c * (T23 * T11 - T21 * T13);
  Ti.m[2][0] = c * (T32 * T21 - T31 * T22);
  Ti.m[2][1] = -c * (T32 * T11 - T31 * T12);
  Ti.m[2][2] = c * (T22 * T11 - T21 * T12);
}
BF_HD double det3(const Mat& A) {  // src/math.c:150-156
  return A.m[0][0] * (A.m[1][1] * A.m[2][2] - A.m[1][2] * A.m[2][1]) -
         A.m[0][1] * (A.m[1][0] * A.m[2][2] - A.m[1][2] * A.m[2][0]) +
         A.m[0][2] * (A.m[1][0] * A.m[2][1] - A.m[1][1] * A.m[2][0]);
}
BF_HD void raise2(const Mat& gInv, const Mat& A, Mat& out) {  // src/math.c:159-172
  Mat t;
  matmult(gInv, A, t);
  matmult(t, gInv, out);
}
BF_HD void lower2(const Mat& gCov, const Mat& A, Mat& out) {  // src/math.c:174-186
  Mat t;
  matmult(gCov, A, t);
  matmult(t, gCov, out);
}
// ComputeMetricTensor (active_strain.c:769-779) from basis rows
BF_HD void metric(const V3<double> g[3], Mat& gm) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) gm.m[i][j] = dot(g[i], g[j]);
}

// Element-level inputs of the exact pipeline.
struct ElemIn {
  MidGeom cur;                 // current midsurface (geom[0])
  V3<double> a31c, a32c;       // Compute_a3_alpha_ of the current midsurface
  RefGeom ref;                 // a1, a2, a3, a3_1, a3_2 of geom0
  double Fa_cart[3][3];        // active deformation in the Cartesian frame (:1227-1256)
  double h0, mu, K;
  int n_qp, iters;
  const double* theta;
  const double* wq;
};

// Stress S^{00}, S^{11}, S^{01} per qp after the C33 loop, final C33 per qp, status bits.
struct ElemOut {
  double Sv[kMaxQp][3];
  double c33[kMaxQp];
  int status;
};

// Pick element i of (a0, a1, a2) for a lane-uniform index: keeps 3x3 operands in
// registers where a runtime-indexed array would spill to scratch.
BF_HD double sel3(int i, double a0, double a1, double a2) { return i == 0 ? a0 : (i == 1 ? a1 : a2); }
BF_HD void row3(const Mat& M, int i, double r[3]) {
  r[0] = sel3(i, M.m[0][0], M.m[1][0], M.m[2][0]);
  r[1] = sel3(i, M.m[0][1], M.m[1][1], M.m[2][1]);
  r[2] = sel3(i, M.m[0][2], M.m[1][2], M.m[2][2]);
}
BF_HD void col3(const Mat& M, int j, double c[3]) {
  c[0] = sel3(j, M.m[0][0], M.m[0][1], M.m[0][2]);
  c[1] = sel3(j, M.m[1][0], M.m[1][1], M.m[1][2]);
  c[2] = sel3(j, M.m[2][0], M.m[2][1], M.m[2][2]);
}

// Zero-skip guard of the CC^{2222} sum: factors below this magnitude cannot overflow the
// 8-factor products, so a zero factor makes an exact zero term.
constexpr double kZeroSkipMax = 1e25;
BF_HD bool bounded(const Mat& A) {
  bool ok = true;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) ok = ok && fabs(A.m[i][j]) < kZeroSkipMax;
  return ok;
}

// Returns v unchanged, but the device compiler can no longer see that it is a constant:
// products that start from it stay inside the loop that calls this instead of being
// hoisted out of it and kept live (81 hoisted FP64 products per contraction would
// exceed the register file and spill).
BF_HD double opaque(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" : "+v"(v));
#endif
  return v;
}

// ---- ElemTotTangMatTens :1603-1643: CC^{2222}, one sum of 3^8 terms in the loop order
// g, h, m, n, p, s, w, z.  The 81 blocks (g, h, m, n) of 81 terms (p, s, w, z) each are
// formed by cc_block_terms; the sum itself is cc_sum_serial here (one lane adds every term
// in order) or a split form on the device that forms several blocks at once and still adds
// the terms in this order.
//
// Blocks of terms with an exactly zero outer factor (Fi[g][h], G[m][2], Fi[n][m]) are
// skipped when every factor is finite and below 1e25 in magnitude: each such term is then an
// exact (signed) zero, and CC, which starts at +0 and can never become -0 under
// round-to-nearest addition, is unchanged by it -- the sum stays bit for bit the
// reference's.  (Flat reference shells have G[a][2] = 0 and Fa_inv.Cov[a][2] =
// Fa_inv.Cov[2][a] = 0 at the mid-surface point, theta = 0: 5 of the 81 blocks remain.  The
// inner factors are left alone: a branch there would split the unrolled block and cost more
// than it saves.)
struct CcIn {
  Mat Fi, G, Ch;  // Fa_inv.Cov, G^{ij} (contravariant reference metric), C_h = Ce_inv.Cont
  double alpha, b2;
  bool zskip;
};
// Factors of block (g, h, m, n).
struct CcBlock {
  double Figh, Finm, Gm2, Gg2, Ggn, Chhn;
  double Chr[3], Chc[3];  // Ch[h][*], Ch[*][n]
};
BF_HD bool cc_block_active(const CcIn& c, const CcBlock& k) {
  return !c.zskip || (k.Figh != 0.0 && k.Gm2 != 0.0 && k.Finm != 0.0);
}
// Factors of block b = 27 g + 9 h + 3 m + n.
BF_HD CcBlock cc_block_at(const CcIn& c, int b) {
  const int gg = b / 27, h = (b / 9) % 3, mm = (b / 3) % 3, n = b % 3;
  double Gr[3], Fgr[3], Fmc[3];
  CcBlock k;
  row3(c.G, gg, Gr);
  row3(c.Fi, gg, Fgr);
  row3(c.Ch, h, k.Chr);
  col3(c.Fi, mm, Fmc);
  col3(c.Ch, n, k.Chc);
  k.Figh = sel3(h, Fgr[0], Fgr[1], Fgr[2]);
  k.Gm2 = sel3(mm, c.G.m[0][2], c.G.m[1][2], c.G.m[2][2]);
  k.Finm = sel3(n, Fmc[0], Fmc[1], Fmc[2]);
  k.Gg2 = Gr[2];
  k.Ggn = sel3(n, Gr[0], Gr[1], Gr[2]);
  k.Chhn = sel3(n, k.Chr[0], k.Chr[1], k.Chr[2]);
  return k;
}
// The 81 terms of one block in the sum's order: emit(i, term), i = 27 p + 9 s + 3 w + z.
template <class Emit>
BF_HD void cc_block_terms(const CcIn& c, const CcBlock& k, Emit&& emit) {
  const Mat& Fi = c.Fi;
  const Mat& G = c.G;
  const Mat& Ch = c.Ch;
  const double alpha = c.alpha, b2 = c.b2;
  const double quarter = opaque(0.25);
#pragma unroll
  for (int pp = 0; pp < 3; pp++) {
    // CCe(h,n,p,2), CCe(h,2,p,2) (:1578-1581)
    const double e_np2 = b2 * k.Chhn * Ch.m[pp][2] - alpha * (k.Chr[pp] * k.Chc[2] + k.Chr[2] * k.Chc[pp]);
    const double e_2p2 = b2 * k.Chr[2] * Ch.m[pp][2] - alpha * (k.Chr[pp] * Ch.m[2][2] + k.Chr[2] * Ch.m[pp][2]);
#pragma unroll
    for (int s = 0; s < 3; s++)
#pragma unroll
      for (int w = 0; w < 3; w++)
#pragma unroll
        for (int z = 0; z < 3; z++) {
          const double e_npz = b2 * k.Chhn * Ch.m[pp][z] - alpha * (k.Chr[pp] * k.Chc[z] + k.Chr[z] * k.Chc[pp]);
          const double e_2pz = b2 * k.Chr[2] * Ch.m[pp][z] - alpha * (k.Chr[pp] * Ch.m[z][2] + k.Chr[z] * Ch.m[pp][2]);
          emit(27 * pp + 9 * s + 3 * w + z,
               quarter * Fi.m[w][pp] * Fi.m[z][s] * k.Figh * k.Finm * G.m[s][2] * k.Gm2 *
                   (k.Gg2 * (e_npz * G.m[w][2] + e_np2 * G.m[w][z]) +
                    k.Ggn * (e_2pz * G.m[w][2] + e_2p2 * G.m[w][z])));
        }
  }
}
BF_HD double cc_sum_serial(const CcIn& c) {
  double CC = 0.0;
#pragma unroll 1
  for (int gg = 0; gg < 3; gg++) {
    double Gr[3], Fgr[3];
    row3(c.G, gg, Gr);    // G[g][*]
    row3(c.Fi, gg, Fgr);  // Fi[g][*]
#pragma unroll 1
    for (int h = 0; h < 3; h++) {
      CcBlock k;
      k.Gg2 = Gr[2];
      row3(c.Ch, h, k.Chr);
      k.Figh = sel3(h, Fgr[0], Fgr[1], Fgr[2]);
      if (c.zskip && k.Figh == 0.0) continue;
#pragma unroll 1
      for (int mm = 0; mm < 3; mm++) {
        k.Gm2 = sel3(mm, c.G.m[0][2], c.G.m[1][2], c.G.m[2][2]);
        if (c.zskip && k.Gm2 == 0.0) continue;
        double Fmc[3];
        col3(c.Fi, mm, Fmc);  // Fi[*][m]
#pragma unroll 1
        for (int n = 0; n < 3; n++) {
          col3(c.Ch, n, k.Chc);
          k.Finm = sel3(n, Fmc[0], Fmc[1], Fmc[2]);
          if (c.zskip && k.Finm == 0.0) continue;
          k.Ggn = sel3(n, Gr[0], Gr[1], Gr[2]);
          k.Chhn = sel3(n, k.Chr[0], k.Chr[1], k.Chr[2]);
          cc_block_terms(c, k, [&](int, double t) { CC += t; });
        }
      }
    }
  }
  return CC;
}
// ---- ElemElasCGDefTens :1389-1430: entry (i, j) of Ce_inv.Cont, one sum over m, p, q, l
// in that order (i, j may differ between the lanes that call it) ----
BF_HD double ce_entry(const Mat& FaCont, const Mat& gm0, const Mat& Ci, int i, int j) {
  double Fr[3], Fcj[3];
  row3(FaCont, i, Fr);
  col3(FaCont, j, Fcj);
  const double Fo[3] = {opaque(Fr[0]), opaque(Fr[1]), opaque(Fr[2])};
  double acc = 0.0;
#pragma unroll
  for (int mm = 0; mm < 3; mm++)
#pragma unroll
    for (int pp = 0; pp < 3; pp++)
#pragma unroll
      for (int qq = 0; qq < 3; qq++)
#pragma unroll
        for (int l = 0; l < 3; l++)
          acc += Fo[mm] * gm0.m[mm][pp] * Ci.m[pp][qq] * Fcj[l] * gm0.m[l][qq];
  return acc;
}

// ---- ElemTotStress :1506-1528: entry (i, j) of the total stress, one sum over w, p, z, s in
// that order (i, j may differ between the lanes that call it) ----
BF_HD double tot_entry(const Mat& Fi, const Mat& G, const Mat& Se, int i, int j) {
  double Gi[3], Gj[3], Sei[3];
  col3(G, i, Gi);
  col3(G, j, Gj);
  col3(Se, i, Sei);
  const double half = opaque(0.5);
  double acc = 0.0;
#pragma unroll
  for (int w = 0; w < 3; w++)
#pragma unroll
    for (int pp = 0; pp < 3; pp++)
#pragma unroll
      for (int z = 0; z < 3; z++)
#pragma unroll
        for (int s = 0; s < 3; s++)
          acc += half * Fi.m[w][pp] * Fi.m[z][s] * Gj[s] *
                 (Se.m[pp][z] * Gi[w] + Sei[pp] * G.m[w][z]);
  return acc;
}
// The total-stress entries a C33 iteration reads: S^22, and in the last iteration (whose
// values are the ones kept) the in-plane S^00, S^11, S^01 as well.
struct TotOut {
  double S00, S11, S01, S22;
};

struct CcSerial {
  BF_HD double operator()(const CcIn& c) const { return cc_sum_serial(c); }
  BF_HD void stresses(const Mat& Fi, const Mat& G, const Mat& Se, bool last, TotOut& o) const {
    if (last) {
      o.S00 = tot_entry(Fi, G, Se, 0, 0);
      o.S11 = tot_entry(Fi, G, Se, 1, 1);
      o.S01 = tot_entry(Fi, G, Se, 0, 1);
    }
    o.S22 = tot_entry(Fi, G, Se, 2, 2);
  }
  // the nine entries of Ce_inv.Cont, one after the other
  BF_HD void ce_inv(const Mat& FaCont, const Mat& gm0, const Mat& Ci, Mat& Ch) const {
#pragma unroll 1
    for (int i = 0; i < 3; i++)
#pragma unroll 1
      for (int j = 0; j < 3; j++) {
        const double v = ce_entry(FaCont, gm0, Ci, i, j);
        if (i == 0) {
          if (j == 0) Ch.m[0][0] = v; else if (j == 1) Ch.m[0][1] = v; else Ch.m[0][2] = v;
        } else if (i == 1) {
          if (j == 0) Ch.m[1][0] = v; else if (j == 1) Ch.m[1][1] = v; else Ch.m[1][2] = v;
        } else {
          if (j == 0) Ch.m[2][0] = v; else if (j == 1) Ch.m[2][1] = v; else Ch.m[2][2] = v;
        }
      }
  }
};

// One through-thickness point: ElemUpdateG, ElemActDefGrad, ElemCGDefTens and the C33
// loop for qp `qp`.  cc_sum forms CC^{2222}, the nine Ce_inv entries (cc_sum.ce_inv) and
// the total-stress entries (cc_sum.stresses): lanes that solve the same qp together divide
// these among themselves, each entry still one sum in the reference's order.  ModElemC33
// returns at the first qp whose |CC3333| fails the guard (:1674-1681), so in iteration `it` this qp keeps its C33 if it or any earlier qp failed:
// skip(fail, it) receives this qp's guard result and answers that question (sequentially
// from a bit mask, or across the lanes that hold the element's other qps).
struct QpOut {
  double S00, S11, S01, c33;
  int status;
};

template <class Skip, class CcSum>
BF_HD QpOut integrate_qp(const ElemIn& in, int qp, Skip&& skip, CcSum&& cc_sum,
                         bool stop_on_guard = false) {
  QpOut out;
  out.status = 0;
  // Fa_cart and its inverse (:1227-1256, :1291): rebuilt per qp (same values every time)
  // instead of being held live across the qp loop
  Mat Fc, Finv_c;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) Fc.m[i][j] = in.Fa_cart[i][j];
  inv(Fc, Finv_c);
  const double hhalf = 0.5 * in.h0;  // :913
  const double theta3 = hhalf * in.theta[qp];  // :936
  // ---- ElemUpdateG :940-978 ----
  V3<double> g[3], g0[3];
  g[0] = add(in.cur.a1, scale(theta3, in.a31c));
  g[1] = add(in.cur.a2, scale(theta3, in.a32c));
  g[2] = in.cur.a3;
  Mat gm, gmI;
  metric(g, gm);
  inv(gm, gmI);
  g0[0] = add(in.ref.a1, scale(theta3, in.ref.a31));
  g0[1] = add(in.ref.a2, scale(theta3, in.ref.a32));
  g0[2] = in.ref.a3;
  Mat gm0, G;
  metric(g0, gm0);
  inv(gm0, G);
  // ---- ElemActDefGrad :1272-1320 ----
  Mat S, ST, tmp, FaCov, Fi, FaCont;
  for (int j = 0; j < 3; j++) {
    S.m[0][j] = g0[j].x;
    S.m[1][j] = g0[j].y;
    S.m[2][j] = g0[j].z;
  }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) ST.m[i][j] = S.m[j][i];
  matmult(ST, Fc, tmp);
  matmult(tmp, S, FaCov);
  matmult(ST, Finv_c, tmp);
  matmult(tmp, S, Fi);
  raise2(G, FaCov, FaCont);
  // ---- ElemCGDefTens :1341-1363 ----
  Mat CCov, Ci;
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) {
      CCov.m[i][j] = gm.m[i][j];
      Ci.m[i][j] = gmI.m[i][j];
    }
  for (int a = 0; a < 2; a++) {
    CCov.m[a][2] = 0.0;
    CCov.m[2][a] = 0.0;
    Ci.m[a][2] = 0.0;
    Ci.m[2][a] = 0.0;
  }
  CCov.m[2][2] = 1.0;
  Ci.m[2][2] = 1.0;
  double S00 = 0.0, S11 = 0.0, S01 = 0.0;
#pragma unroll 1
  for (int it = 0; it < in.iters; it++) {
    // ---- ElemElasCGDefTens :1389-1430: sum over m, p, q, l per entry ----
    Mat Ch, CeiCov, CeCont, CeCov;
    cc_sum.ce_inv(FaCont, gm0, Ci, Ch);
    lower2(gm0, Ch, CeiCov);
    inv(CeiCov, CeCont);
    lower2(gm0, CeCont, CeCov);
    // ---- ElemElasStress :1458-1487 ----
    const double detCe = det3(CeCov);
    const double detG0 = det3(gm0);
    const double Je = sqrt(detCe / detG0);
    Mat Se;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        const double term1 = in.mu * (G.m[i][j] - Ch.m[i][j]);
        const double term2 = in.K * Ch.m[i][j] * (Je * Je - Je);
        Se.m[i][j] = term1 + term2;
      }
    // ---- ElemTotStress :1506-1528, the four entries later steps read: S^22 in every
    // iteration, the in-plane ones only in the last, whose values are the ones kept ----
    TotOut st;
    const bool last = it == in.iters - 1;
    cc_sum.stresses(Fi, G, Se, last, st);
    if (last) {
      S00 = st.S00;
      S11 = st.S11;
      S01 = st.S01;
    }
    const double S22 = st.S22;
    // ---- ElemElsTangMatTens :1548-1585 (alpha, beta; entries formed on demand) ----
    const double Je2 = Je * Je;
    const double alpha = in.K * (Je2 - Je) - in.mu;
    const double beta = 0.5 * in.K * Je * (2.0 * Je - 1.0);
    const double b2 = 2.0 * beta;
    // ---- ElemTotTangMatTens :1603-1643, [2][2][2][2] ----
    CcIn cc_in;
    cc_in.Fi = Fi;
    cc_in.G = G;
    cc_in.Ch = Ch;
    cc_in.alpha = alpha;
    cc_in.b2 = b2;
    cc_in.zskip = bounded(Fi) && bounded(G) && bounded(Ch) && fabs(alpha) < kZeroSkipMax &&
                  fabs(beta) < kZeroSkipMax;
    const double CC = cc_sum(cc_in);
    // ---- ModElemC33 :1669-1707 ----
    const bool fail = !(fabs(CC) > 1e-14);
    if (fail) out.status |= kStatCC3333;
    // the caller redoes the element with coupled qps and drops this qp's stresses
    if (fail && stop_on_guard) break;
    if (!skip(fail, it)) {
      double DeltaC33 = -1.0 * S22 / CC;
      const double C33_old = CCov.m[2][2];
      CCov.m[2][2] = C33_old + DeltaC33;
      if (det3(CCov) <= 0.0) {
        double lo = 0.0, hi = 1.0;
        for (int bk = 0; bk < 16; bk++) {
          const double mid = 0.5 * (lo + hi);
          CCov.m[2][2] = C33_old + mid * DeltaC33;
          if (det3(CCov) > 0.0) lo = mid; else hi = mid;
        }
        DeltaC33 *= lo;
        CCov.m[2][2] = C33_old + DeltaC33;
        out.status |= kStatClamp;
      }
    }
    // ---- C_inv.Cont = INV(C.Cov) :2101 ----
    inv(CCov, Ci);
  }
  out.S00 = S00;
  out.S11 = S11;
  out.S01 = S01;
  out.c33 = CCov.m[2][2];
  return out;
}

template <class Skip>
BF_HD QpOut integrate_qp(const ElemIn& in, int qp, Skip&& skip) {
  return integrate_qp(in, qp, skip, CcSerial{});
}

BF_HD void integrate(const ElemIn& in, ElemOut& out) {
  out.status = 0;
  unsigned skip_mask = 0u;
  auto skip = [&](bool fail, int it) {
    if (fail) skip_mask |= 1u << it;
    return ((skip_mask >> it) & 1u) != 0u;
  };
#pragma unroll 1
  for (int qp = 0; qp < in.n_qp; qp++) {
    const QpOut q = integrate_qp(in, qp, skip);
    out.Sv[qp][0] = q.S00;
    out.Sv[qp][1] = q.S11;
    out.Sv[qp][2] = q.S01;
    out.c33[qp] = q.c33;
    out.status |= q.status;
  }
}

// ElemUpdFint for one patch node a, regular/irregular branch (:1879-1913 / :1934-1983):
// accumulates over qp in the reference order.  Na0..Nb2 are the node's table entries.
BF_HD V3<double> node_force_exact(double Na0, double Na1, double Nb0, double Nb1, double Nb2,
                                  const MidGeom& G, double pref0, double half_h0,
                                  const double (*Sq)[3], const double* theta, const double* wq,
                                  int n_qp, bool irregular) {
  double F[3] = {0.0, 0.0, 0.0};
  const double nn[3] = {G.a3.x, G.a3.y, G.a3.z};
  const double a1[3] = {G.a1.x, G.a1.y, G.a1.z};
  const double a2[3] = {G.a2.x, G.a2.y, G.a2.z};
  const V3<double> ng = add(scale(Na0, G.gc1), scale(Na1, G.gc2));
  const double b1 = dot(ng, G.Aaa), b2 = dot(ng, G.Abb), b3 = dot(ng, G.Aab);
  for (int qp = 0; qp < n_qp; qp++) {
    const double w_qp = wq[qp], xi_qp = theta[qp];
    const double* Sv = Sq[qp];
    for (int c = 0; c < 3; c++) {
      const double Bm0 = Na0 * a1[c];
      const double Bm1 = Na1 * a2[c];
      const double Bm2 = Na0 * a2[c] + Na1 * a1[c];
      double mem, bend;
      if (!irregular) {  // :1879-1913, loop form starting from 0.0
        const double Bs0 = -(Nb0 + b1) * nn[c];
        const double Bs1 = -(Nb1 + b2) * nn[c];
        const double Bs2 = -2.0 * (Nb2 + b3) * nn[c];
        mem = 0.0;
        bend = 0.0;
        mem += Bm0 * Sv[0];
        bend += Bs0 * Sv[0];
        mem += Bm1 * Sv[1];
        bend += Bs1 * Sv[1];
        mem += Bm2 * Sv[2];
        bend += Bs2 * Sv[2];
      } else {  // :1959-1983
        const double Bs0 = -(Nb0 + b1);
        const double Bs1 = -(Nb1 + b2);
        const double Bs2 = -2.0 * (Nb2 + b3);
        mem = Bm0 * Sv[0] + Bm1 * Sv[1] + Bm2 * Sv[2];
        bend = (Bs0 * nn[c]) * Sv[0] + (Bs1 * nn[c]) * Sv[1] + (Bs2 * nn[c]) * Sv[2];
      }
      F[c] += pref0 * w_qp * (mem + bend * (half_h0 * xi_qp));
    }
  }
  return mk(F[0], F[1], F[2]);
}

// LV apex-cap CST membrane force (:1825-1850) for corner a with shape gradient (Na0, Na1).
BF_HD V3<double> cap_force_exact(double Na0, double Na1, const MidGeom& G, double pref0,
                                 const double (*Sq)[3], const double* wq, int n_qp) {
  double F[3] = {0.0, 0.0, 0.0};
  const double a1[3] = {G.a1.x, G.a1.y, G.a1.z};
  const double a2[3] = {G.a2.x, G.a2.y, G.a2.z};
  for (int qp = 0; qp < n_qp; qp++) {
    const double w_qp = wq[qp];
    const double* Sv = Sq[qp];
    for (int c = 0; c < 3; c++) {
      const double Bm0 = Na0 * a1[c];
      const double Bm1 = Na1 * a2[c];
      const double Bm2 = Na0 * a2[c] + Na1 * a1[c];
      F[c] += pref0 * w_qp * (Bm0 * Sv[0] + Bm1 * Sv[1] + Bm2 * Sv[2]);
    }
  }
  return mk(F[0], F[1], F[2]);
}

}  // namespace exact
}  // namespace bf
