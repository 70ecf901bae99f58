// bf_elem.h -- per-element active-strain Kirchhoff-Love force arithmetic, shared by the
// gfx950 kernels (bf_force.hip) and by the CPU consistency harness in tests/.
//
// Everything here is templated on the per-quadrature-point real type R (double for the
// plain FP64 path, float for the FP32-mixed path).  Node coordinates, the midsurface
// geometry sums and the nodal force accumulation stay in double in both paths.
//
// The reference evaluates the constitutive chain with literal index loops
// (src/active_strain.c:1381-1648: a 6-deep loop for Ce^-1, a 6-deep loop for S and an
// 8-deep, 6561-term loop for CC^{2222}).  Here the same tensors are produced by staged
// 3x3 products and closed forms of the contractions:
//   Ce^-1     = T Ci T^T,          T  = Fa^{..} g0_{..}        (:1394-1417)
//   Je        = 1/sqrt(det(Ce^-1) det(g0_{..}))              (:1458-1461; det Ce = 1/det Ce^-1)
//   S         = 1/2 (H^T Se H + Se^T H H),  H = Fa^-1_{..} g0^{..}   (:1509-1528)
//   CC^{2222} = 1/4 [2(beta-alpha) b^2 + 4 beta b (P.r) - 4 alpha B(h,P)(h.r)
//                    + 2 beta (P.r)^2 - alpha B(P,P) C^22 - alpha (P.r)^2]   (:1614-1643)
//               with h = H e3, P = H h, r = Ce^-1 e3, b = h.Ce^-1 h, B(u,v) = u.Ce^-1 v.
// The staged forms are exact algebraic identities of the reference's sums (given the
// symmetry of Ce^-1, g0 and Fa^-1 that the reference's own construction produces); they
// differ from the loops only by rounding order.
#pragma once

#if defined(__HIPCC__) || defined(__HIP__)
#define BF_HD __host__ __device__ __forceinline__
#else
#define BF_HD inline
#endif

#include <math.h>

namespace bf {

constexpr int kSentinel = 1000000;  // empty patch slot (src/io.c:306)
constexpr int kMaxNen = 18;         // widest patch handled (valence 12)
constexpr int kMaxQp = 5;

// element kinds packed in meta bits 0..1
enum : int { kRegular = 0, kIrregular = 1, kCap = 2 };

// per-element status bits (OR-ed into a flag word, first element kept by atomicMin)
enum : int { kStatCC3333 = 1, kStatNonFinite = 2, kStatClamp = 4, kStatFallback = 8,
             kStatIllCond = 16, kStatRange = 32 };
// kStatClamp: the det(C) > 0 bisection of ModElemC33 (active_strain.c:1696-1706) ran: the
// element is outside the regime where the local C33 iteration is well conditioned.
// kStatRange: C33 left [kC33Lo, kC33Hi].

template <class R>
struct V3 {
  R x, y, z;
};

template <class R>
BF_HD V3<R> mk(R x, R y, R z) {
  V3<R> v;
  v.x = x;
  v.y = y;
  v.z = z;
  return v;
}
template <class R>
BF_HD V3<R> add(V3<R> a, V3<R> b) {
  return mk(a.x + b.x, a.y + b.y, a.z + b.z);
}
template <class R>
BF_HD V3<R> scale(R s, V3<R> a) {
  return mk(s * a.x, s * a.y, s * a.z);
}
template <class R>
BF_HD V3<R> cross(V3<R> v1, V3<R> v2) {  // src/math.c:29-37
  return mk(v1.y * v2.z - v2.y * v1.z, -v1.x * v2.z + v2.x * v1.z, v1.x * v2.y - v2.x * v1.y);
}
template <class R>
BF_HD R dot(V3<R> a, V3<R> b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class R>
BF_HD R norm(V3<R> a) {
  return sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
}
template <class R>
BF_HD V3<R> unit(V3<R> a) {
  R n = norm(a);
  return mk(a.x / n, a.y / n, a.z / n);
}
template <class T, class R>
BF_HD V3<T> cvt(V3<R> a) {
  return mk((T)a.x, (T)a.y, (T)a.z);
}

// symmetric 3x3, storage (00, 11, 22, 01, 02, 12)
template <class R>
struct Sym {
  R a00, a11, a22, a01, a02, a12;
};
template <class R>
struct M33 {
  R m[3][3];
};

template <class R>
BF_HD R sdet(const Sym<R>& A) {  // DET3x3 row expansion, src/math.c:150-156
  return A.a00 * (A.a11 * A.a22 - A.a12 * A.a12) - A.a01 * (A.a01 * A.a22 - A.a12 * A.a02) +
         A.a02 * (A.a01 * A.a12 - A.a11 * A.a02);
}
template <class R>
BF_HD Sym<R> sinv(const Sym<R>& T, R* det_out) {  // cofactor inverse, src/math.c:77-91
  const R T11 = T.a00, T12 = T.a01, T13 = T.a02, T22 = T.a11, T23 = T.a12, T33 = T.a22;
  const R det = T11 * (T33 * T22 - T23 * T23) - T12 * (T33 * T12 - T23 * T13) +
                T13 * (T23 * T12 - T22 * T13);
  const R c = R(1) / det;
  Sym<R> I;
  I.a00 = c * (T33 * T22 - T23 * T23);
  I.a01 = -c * (T33 * T12 - T23 * T13);
  I.a02 = c * (T23 * T12 - T22 * T13);
  I.a11 = c * (T33 * T11 - T13 * T13);
  I.a12 = -c * (T23 * T11 - T12 * T13);
  I.a22 = c * (T22 * T11 - T12 * T12);
  if (det_out) *det_out = det;
  return I;
}
template <class R>
BF_HD M33<R> full(const Sym<R>& S) {
  M33<R> M;
  M.m[0][0] = S.a00; M.m[0][1] = S.a01; M.m[0][2] = S.a02;
  M.m[1][0] = S.a01; M.m[1][1] = S.a11; M.m[1][2] = S.a12;
  M.m[2][0] = S.a02; M.m[2][1] = S.a12; M.m[2][2] = S.a22;
  return M;
}
template <class R>
BF_HD M33<R> mul(const M33<R>& A, const M33<R>& B) {
  M33<R> C;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) C.m[i][j] = A.m[i][0] * B.m[0][j] + A.m[i][1] * B.m[1][j] + A.m[i][2] * B.m[2][j];
  return C;
}
// |A| |B| with element-wise absolute values
template <class R>
BF_HD M33<R> abs_mul(const M33<R>& A, const M33<R>& B) {
  M33<R> C;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      C.m[i][j] = fabs(A.m[i][0]) * fabs(B.m[0][j]) + fabs(A.m[i][1]) * fabs(B.m[1][j]) +
                  fabs(A.m[i][2]) * fabs(B.m[2][j]);
  return C;
}

// Rounding noise of one ModElemC33 update as the reference computes it.  The reference sums
// CC^{2222} over 6561 terms (ElemTotTangMatTens :1603-1643) and S^{22} over 81
// (ElemTotStress :1506-1528); recursive summation of n products of k factors errs by at most
// (n + k) eps times the sum of the terms' magnitudes.  Those magnitude sums are the staged
// closed forms evaluated on element-wise absolute values (Ha = |Fa_inv.Cov| |gm0.Cont|,
// H2a = Ha Ha, |Ce_inv.Cont|, |Se|, |alpha|, |beta|), an upper bound because the staged
// derivation only regroups terms.  Returns the bound on the resulting change of C33,
// (|dC33| rel.err(CC) + err(S22) / |CC|), from inv_aCC = 1 / |CC|.  Elements where it
// exceeds the fast paths' budget are ill-conditioned: the reference's own C33 is
// noise-dominated there.
template <class R>
BF_HD R c33_update_noise(const M33<R>& Ha, const M33<R>& H2a, const M33<R>& Ch,
                         const M33<R>& Se, R alpha, R beta, R S22, R inv_aCC) {
  const R eps = R(1.1102230246251565e-16);  // 2^-53, the reference's FP64
  const R hx = Ha.m[0][2], hy = Ha.m[1][2], hz = Ha.m[2][2];
  const R Px = H2a.m[0][2], Py = H2a.m[1][2], Pz = H2a.m[2][2];
  M33<R> Ca;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) Ca.m[i][j] = fabs(Ch.m[i][j]);
  const R rx = Ca.m[0][2], ry = Ca.m[1][2], rz = Ca.m[2][2];
  const R Chx = Ca.m[0][0] * hx + Ca.m[0][1] * hy + Ca.m[0][2] * hz;
  const R Chy = Ca.m[1][0] * hx + Ca.m[1][1] * hy + Ca.m[1][2] * hz;
  const R Chz = Ca.m[2][0] * hx + Ca.m[2][1] * hy + Ca.m[2][2] * hz;
  const R b = hx * Chx + hy * Chy + hz * Chz;
  const R BhP = Px * Chx + Py * Chy + Pz * Chz;
  const R BPP = Px * (Ca.m[0][0] * Px + Ca.m[0][1] * Py + Ca.m[0][2] * Pz) +
                Py * (Ca.m[1][0] * Px + Ca.m[1][1] * Py + Ca.m[1][2] * Pz) +
                Pz * (Ca.m[2][0] * Px + Ca.m[2][1] * Py + Ca.m[2][2] * Pz);
  const R Pr = Px * rx + Py * ry + Pz * rz;
  const R hr = hx * rx + hy * ry + hz * rz;
  const R aa = fabs(alpha), ab = fabs(beta);
  const R CCabs = R(0.25) * (R(2) * (ab + aa) * b * b + R(4) * ab * b * Pr + R(4) * aa * BhP * hr +
                             R(2) * ab * Pr * Pr + aa * BPP * Ca.m[2][2] + aa * Pr * Pr);
  // S^22 = 1/2 (H^T Se H + Se H2)_22 on magnitudes
  R SHz[3];
#pragma unroll
  for (int p = 0; p < 3; p++)
    SHz[p] = fabs(Se.m[p][0]) * hx + fabs(Se.m[p][1]) * hy + fabs(Se.m[p][2]) * hz;
  const R S22abs = R(0.5) * (hx * SHz[0] + hy * SHz[1] + hz * SHz[2] + fabs(Se.m[0][2]) * Px +
                             fabs(Se.m[1][2]) * Py + fabs(Se.m[2][2]) * Pz);
  return (fabs(S22) * (R(6569) * eps * CCabs) * inv_aCC + R(89) * eps * S22abs) * inv_aCC;
}

// Gram matrix of three vectors g_i . g_j (ComputeMetricTensor, active_strain.c:769-779)
template <class R>
BF_HD Sym<R> gram(V3<R> g0, V3<R> g1, V3<R> g2) {
  Sym<R> G;
  G.a00 = dot(g0, g0);
  G.a11 = dot(g1, g1);
  G.a22 = dot(g2, g2);
  G.a01 = dot(g0, g1);
  G.a02 = dot(g0, g2);
  G.a12 = dot(g1, g2);
  return G;
}

// Compute_a3_alpha_ (active_strain.c:745-762): un-normalised derivative of a1 x a2 over |a1 x a2|
template <class R>
BF_HD void a3_alpha(V3<R> a1, V3<R> a2, V3<R> Aaa, V3<R> Aab, V3<R> Abb, V3<R>& a31, V3<R>& a32) {
  const R s = norm(cross(a1, a2));
  a31 = add(cross(Aaa, a2), cross(a1, Aab));
  a32 = add(cross(Aab, a2), cross(a1, Abb));
  a31 = mk(a31.x / s, a31.y / s, a31.z / s);
  a32 = mk(a32.x / s, a32.y / s, a32.z / s);
}

// Midsurface data of one configuration (SubdivGeomQP without the IN* arrays).
struct MidGeom {
  V3<double> a1, a2, a3, gc1, gc2, Aaa, Abb, Aab;
};

// a3 = unit(a1 x a2) and the contravariant in-plane basis (active_strain.c:535-541)
BF_HD void geom_tail(MidGeom& G) {
  G.a3 = unit(cross(G.a1, G.a2));
  G.gc1 = cross(G.a2, G.a3);
  G.gc1 = scale(1.0 / dot(G.a1, G.gc1), G.gc1);
  G.gc2 = cross(G.a3, G.a1);
  G.gc2 = scale(1.0 / dot(G.a2, G.gc2), G.gc2);
}

// Midsurface geometry of a patch (ElemUpdateGeomSubdivFromCoords_, active_strain.c:477-541
// regular, :668-699 irregular): second-derivative sums with Nab rows, tangents with Na rows,
// each accumulated from 0 in slot order.  `xyz(i)` returns the coordinates of slot i.
template <class CoordFn>
BF_HD MidGeom mid_geom(int nen, const double* na0, const double* na1, const double* nb0,
                       const double* nb1, const double* nb2, CoordFn xyz) {
  MidGeom G;
  G.Aaa = mk(0.0, 0.0, 0.0);
  G.Abb = G.Aaa;
  G.Aab = G.Aaa;
  G.a1 = G.Aaa;
  G.a2 = G.Aaa;
  for (int i = 0; i < nen; i++) {
    const V3<double> X = xyz(i);
    G.Aaa.x += nb0[i] * X.x; G.Aaa.y += nb0[i] * X.y; G.Aaa.z += nb0[i] * X.z;
    G.Abb.x += nb1[i] * X.x; G.Abb.y += nb1[i] * X.y; G.Abb.z += nb1[i] * X.z;
    G.Aab.x += nb2[i] * X.x; G.Aab.y += nb2[i] * X.y; G.Aab.z += nb2[i] * X.z;
    G.a1.x += na0[i] * X.x; G.a1.y += na0[i] * X.y; G.a1.z += na0[i] * X.z;
    G.a2.x += na1[i] * X.x; G.a2.y += na1[i] * X.y; G.a2.z += na1[i] * X.z;
  }
  geom_tail(G);
  return G;
}

// mid_geom for coordinates read from HBM: every slot's coordinates are loaded before the
// first sum (slots past nen read slot nen - 1, so the loads are branch-free and overlap
// instead of forming one round trip per slot), then the same sums run in the same slot
// order as mid_geom.
template <class CoordFn>
BF_HD MidGeom mid_geom_loaded(int nen, const double* na0, const double* na1, const double* nb0,
                              const double* nb1, const double* nb2, CoordFn xyz) {
  V3<double> X[kMaxNen];
#pragma unroll
  for (int i = 0; i < kMaxNen; i++) X[i] = xyz(i < nen ? i : nen - 1);
  MidGeom G;
  G.Aaa = mk(0.0, 0.0, 0.0);
  G.Abb = G.Aaa;
  G.Aab = G.Aaa;
  G.a1 = G.Aaa;
  G.a2 = G.Aaa;
#pragma unroll
  for (int i = 0; i < kMaxNen; i++) {
    if (i < nen) {
      G.Aaa.x += nb0[i] * X[i].x; G.Aaa.y += nb0[i] * X[i].y; G.Aaa.z += nb0[i] * X[i].z;
      G.Abb.x += nb1[i] * X[i].x; G.Abb.y += nb1[i] * X[i].y; G.Abb.z += nb1[i] * X[i].z;
      G.Aab.x += nb2[i] * X[i].x; G.Aab.y += nb2[i] * X[i].y; G.Aab.z += nb2[i] * X[i].z;
      G.a1.x += na0[i] * X[i].x; G.a1.y += na0[i] * X[i].y; G.a1.z += na0[i] * X[i].z;
      G.a2.x += na1[i] * X[i].x; G.a2.y += na1[i] * X[i].y; G.a2.z += na1[i] * X[i].z;
    }
  }
  geom_tail(G);
  return G;
}

// Loop box-spline derivatives at the element centre of a regular patch, the verbatim
// 4-decimal values of src/active_strain.c:69-88 (Na_center, Nab_center).
constexpr double kLoopNa[2][12] = {
    {-0.0247, -0.0309, 0.0000, -0.4815, -0.1852, 0.0247, 0.4815, 0.0000, -0.0062, 0.0309,
     0.1852, 0.0062},
    {-0.0309, -0.0247, -0.1852, -0.4815, 0.0000, -0.0062, 0.0000, 0.4815, 0.0247, 0.0062,
     0.1852, 0.0309}};
constexpr double kLoopNab[3][12] = {
    {0.1111, 0.2222, -0.2222, -0.2222, 0.4444, 0.1111, -0.2222, -0.8889, 0.0000, 0.2222,
     0.4444, 0.0000},
    {0.2222, 0.1111, 0.4444, -0.2222, -0.2222, 0.0000, -0.8889, -0.2222, 0.1111, 0.0000,
     0.4444, 0.2222},
    {0.1667, 0.1667, -0.1111, 0.2222, -0.1111, -0.0556, -0.4444, -0.4444, -0.0556, 0.0556,
     0.5556, 0.0556}};

// mid_geom of a regular patch with the tables folded in: all 12 slot coordinates are
// fetched before any arithmetic (their loads overlap instead of forming a chain of 12
// round trips), then the same sums run in the same slot order.
template <class CoordFn>
BF_HD MidGeom mid_geom_regular(CoordFn xyz) {
  V3<double> X[12];
#pragma unroll
  for (int i = 0; i < 12; i++) X[i] = xyz(i);
  MidGeom G;
  G.Aaa = mk(0.0, 0.0, 0.0);
  G.Abb = G.Aaa;
  G.Aab = G.Aaa;
  G.a1 = G.Aaa;
  G.a2 = G.Aaa;
#pragma unroll
  for (int i = 0; i < 12; i++) {
    const double nb0 = kLoopNab[0][i], nb1 = kLoopNab[1][i], nb2 = kLoopNab[2][i];
    const double na0 = kLoopNa[0][i], na1 = kLoopNa[1][i];
    G.Aaa.x += nb0 * X[i].x; G.Aaa.y += nb0 * X[i].y; G.Aaa.z += nb0 * X[i].z;
    G.Abb.x += nb1 * X[i].x; G.Abb.y += nb1 * X[i].y; G.Abb.z += nb1 * X[i].z;
    G.Aab.x += nb2 * X[i].x; G.Aab.y += nb2 * X[i].y; G.Aab.z += nb2 * X[i].z;
    G.a1.x += na0 * X[i].x; G.a1.y += na0 * X[i].y; G.a1.z += na0 * X[i].z;
    G.a2.x += na1 * X[i].x; G.a2.y += na1 * X[i].y; G.a2.z += na1 * X[i].z;
  }
  geom_tail(G);
  return G;
}

// Flat-triangle (CST) geometry of an LV apex-cap element (active_strain.c:448-472).
BF_HD MidGeom cap_geom(V3<double> x1, V3<double> x2, V3<double> x3) {
  MidGeom G;
  G.a1 = mk(x2.x - x1.x, x2.y - x1.y, x2.z - x1.z);
  G.a2 = mk(x3.x - x1.x, x3.y - x1.y, x3.z - x1.z);
  geom_tail(G);
  G.Aaa = mk(0.0, 0.0, 0.0);
  G.Abb = G.Aaa;
  G.Aab = G.Aaa;
  return G;
}

// Activation description of one element: Fa_cart = diag(d) - gam * f f^T.
struct Activation {
  double d0, d1, d2;  // diagonal part
  double gam;         // rank-1 coefficient
  V3<double> f;       // fibre
};

// Per-element reference data cached once (geom0): a_alpha, a3 and a3,alpha of the
// reference midsurface.
struct RefGeom {
  V3<double> a1, a2, a3, a31, a32;
};

// Result of the through-thickness integration of one element: the stress resultants
//   N = sum_q w_q Sv_q   and   M = sum_q w_q theta_q Sv_q,  Sv = (S^11, S^22, S^12),
// plus the final C33 per qp and status bits.
template <class R>
struct Resultants {
  double N0, N1, N2, M0, M1, M2;
  double c33[kMaxQp];
  int status;
  float noise;  // largest c33_update_noise of the element's updates
};

// Bound on the reference's rounding-driven change of C33 in one update above which a
// staged chain gives the element up (the FP32 chain too: with a looser 1e-8 the mixed path
// stays within its 1e-5 budget on the 256 x 192 membrane but not on noisier states of the
// same mesh).
constexpr double kC33NoiseMax = 1e-11;
// The FP64 chain standing in for the FP32 one (second tier of the FP32-mixed path) answers
// to the mixed path's 1e-5 budget: 1e-8 (staged FP64 errors there stay below 3e-12).
constexpr double kC33NoiseMaxTier2 = 1e-8;

// Result of one through-thickness point of a staged chain: the total stress entries
// ElemUpdFint reads, the final C33 and the status bits.
template <class R>
struct StagedQp {
  R S00, S11, S01;
  double c33;
  int status;
  float noise;
};

// Staged constitutive chain for one through-thickness point (ElemUpdateG, ElemActDefGrad,
// ElemCGDefTens, ElemC33Solve; active_strain.c:896-982, 1196-1378, 1381-1721, 2048-2138).
//   cur: current a1,a2,a3,a3_1,a3_2   ref: reference ones   R: per-qp arithmetic type.
// skip(fail, it) decides whether the CC3333 guard of iteration it keeps this qp's C33 (see
// integrate_element for the reference's cross-qp rule).
template <class R, class Skip>
BF_HD StagedQp<R> staged_qp(const V3<double>& a1c, const V3<double>& a2c,
                            const V3<double>& a3c, const V3<double>& a31c,
                            const V3<double>& a32c, const RefGeom& ref, const Activation& act,
                            double h0, double mu_d, double K_d, int iters, double theta_q,
                            Skip&& skip, bool stop_on_clamp = false,
                            double noise_max = kC33NoiseMax) {
  StagedQp<R> out;
  out.status = 0;
  out.noise = 0.0f;
  out.S00 = out.S11 = out.S01 = R(0);
  out.c33 = 1.0;
  const R mu = (R)mu_d, K = (R)K_d;
  const double hhalf = 0.5 * h0;
  // activation in the Cartesian frame, its inverse:  Fa^-1 = diag(1/d) + kap f f^T
  const bool rank1 = act.gam != 0.0;
  const R gam = (R)act.gam;
  const V3<R> f = cvt<R>(act.f);
  const R ff = dot(f, f);
  const R kap = rank1 ? gam / (R(1) - gam * ff) : R(0);
  const R d0 = (R)act.d0, d1 = (R)act.d1, d2 = (R)act.d2;
  const bool diag_id = (act.d0 == 1.0 && act.d1 == 1.0 && act.d2 == 1.0);
  {
    const double t3d = hhalf * theta_q;
    const R t3 = (R)t3d;
    // ---- current metric (ElemUpdateG :940-957) ----
    const V3<R> g0 = add(cvt<R>(a1c), scale(t3, cvt<R>(a31c)));
    const V3<R> g1 = add(cvt<R>(a2c), scale(t3, cvt<R>(a32c)));
    const V3<R> g2 = cvt<R>(a3c);
    const Sym<R> gm = gram(g0, g1, g2);
    const Sym<R> gmI = sinv(gm, (R*)nullptr);
    // ---- reference metric (:961-978) ----
    const V3<R> h0v = add(cvt<R>(ref.a1), scale(t3, cvt<R>(ref.a31)));
    const V3<R> h1v = add(cvt<R>(ref.a2), scale(t3, cvt<R>(ref.a32)));
    const V3<R> h2v = cvt<R>(ref.a3);
    const Sym<R> gm0 = gram(h0v, h1v, h2v);
    const Sym<R> G = sinv(gm0, (R*)nullptr);  // gm0.Cont
    const R det0 = sdet(gm0);
    // ---- activation (ElemActDefGrad :1272-1320) ----
    // Fa.Cov = S^T Fa_cart S, Fa_inv.Cov = S^T Fa_cart^-1 S with S = [g0_1 g0_2 g0_3]
    const R phi0 = dot(f, h0v), phi1 = dot(f, h1v), phi2 = dot(f, h2v);
    Sym<R> FaCov, FiCov;
    if (diag_id) {
      FaCov = gm0;
      FiCov = gm0;
    } else {
      const R e0 = R(1) / d0, e1 = R(1) / d1, e2 = R(1) / d2;
      FaCov.a00 = d0 * h0v.x * h0v.x + d1 * h0v.y * h0v.y + d2 * h0v.z * h0v.z;
      FaCov.a11 = d0 * h1v.x * h1v.x + d1 * h1v.y * h1v.y + d2 * h1v.z * h1v.z;
      FaCov.a22 = d0 * h2v.x * h2v.x + d1 * h2v.y * h2v.y + d2 * h2v.z * h2v.z;
      FaCov.a01 = d0 * h0v.x * h1v.x + d1 * h0v.y * h1v.y + d2 * h0v.z * h1v.z;
      FaCov.a02 = d0 * h0v.x * h2v.x + d1 * h0v.y * h2v.y + d2 * h0v.z * h2v.z;
      FaCov.a12 = d0 * h1v.x * h2v.x + d1 * h1v.y * h2v.y + d2 * h1v.z * h2v.z;
      FiCov.a00 = e0 * h0v.x * h0v.x + e1 * h0v.y * h0v.y + e2 * h0v.z * h0v.z;
      FiCov.a11 = e0 * h1v.x * h1v.x + e1 * h1v.y * h1v.y + e2 * h1v.z * h1v.z;
      FiCov.a22 = e0 * h2v.x * h2v.x + e1 * h2v.y * h2v.y + e2 * h2v.z * h2v.z;
      FiCov.a01 = e0 * h0v.x * h1v.x + e1 * h0v.y * h1v.y + e2 * h0v.z * h1v.z;
      FiCov.a02 = e0 * h0v.x * h2v.x + e1 * h0v.y * h2v.y + e2 * h0v.z * h2v.z;
      FiCov.a12 = e0 * h1v.x * h2v.x + e1 * h1v.y * h2v.y + e2 * h1v.z * h2v.z;
    }
    if (rank1) {
      FaCov.a00 -= gam * phi0 * phi0; FaCov.a11 -= gam * phi1 * phi1; FaCov.a22 -= gam * phi2 * phi2;
      FaCov.a01 -= gam * phi0 * phi1; FaCov.a02 -= gam * phi0 * phi2; FaCov.a12 -= gam * phi1 * phi2;
      FiCov.a00 += kap * phi0 * phi0; FiCov.a11 += kap * phi1 * phi1; FiCov.a22 += kap * phi2 * phi2;
      FiCov.a01 += kap * phi0 * phi1; FiCov.a02 += kap * phi0 * phi2; FiCov.a12 += kap * phi1 * phi2;
    }
    const M33<R> Gm = full(G), g0m = full(gm0);
    // Fa.Cont = G FaCov G ; T = Fa.Cont gm0 (mixed)
    const M33<R> FaCont = mul(mul(Gm, full(FaCov)), Gm);
    const M33<R> T = mul(FaCont, g0m);
    // H = Fa_inv.Cov G  (mixed), H2 = H H
    const M33<R> H = mul(full(FiCov), Gm);
    const M33<R> H2 = mul(H, H);
    const M33<R> Ha = abs_mul(full(FiCov), Gm);
    const M33<R> H2a = abs_mul(Ha, Ha);
    // ---- ElemCGDefTens (:1341-1363): C.Cov in-plane block from gm, C33 = 1;
    //      C_inv.Cont in-plane block from gm.Cont, 1 at [2][2]
    const R c00 = gm.a00, c01 = gm.a01, c11 = gm.a11;
    R c33 = R(1);
    R ci00 = gmI.a00, ci01 = gmI.a01, ci11 = gmI.a11, ci22 = R(1);
    // after the first update C.Cov is block-diagonal: its in-plane inverse no longer
    // depends on C33
    const R rD = R(1) / (c00 * c11 - c01 * c01);
    R S00 = 0, S11 = 0, S01 = 0;
    for (int it = 0; it < iters; it++) {
      const bool last = it == iters - 1;
      // ---- Ce^-1 = T Ci T^T (ElemElasCGDefTens :1394-1417) ----
      M33<R> TC;  // T * Ci, Ci block-diagonal
#pragma unroll
      for (int i = 0; i < 3; i++) {
        TC.m[i][0] = T.m[i][0] * ci00 + T.m[i][1] * ci01;
        TC.m[i][1] = T.m[i][0] * ci01 + T.m[i][1] * ci11;
        TC.m[i][2] = T.m[i][2] * ci22;
      }
      Sym<R> Ch;
      Ch.a00 = TC.m[0][0] * T.m[0][0] + TC.m[0][1] * T.m[0][1] + TC.m[0][2] * T.m[0][2];
      Ch.a11 = TC.m[1][0] * T.m[1][0] + TC.m[1][1] * T.m[1][1] + TC.m[1][2] * T.m[1][2];
      Ch.a22 = TC.m[2][0] * T.m[2][0] + TC.m[2][1] * T.m[2][1] + TC.m[2][2] * T.m[2][2];
      Ch.a01 = TC.m[0][0] * T.m[1][0] + TC.m[0][1] * T.m[1][1] + TC.m[0][2] * T.m[1][2];
      Ch.a02 = TC.m[0][0] * T.m[2][0] + TC.m[0][1] * T.m[2][1] + TC.m[0][2] * T.m[2][2];
      Ch.a12 = TC.m[1][0] * T.m[2][0] + TC.m[1][1] * T.m[2][1] + TC.m[1][2] * T.m[2][2];
      // ---- Je (ElemElasStress :1458-1461) ----
      const R detCh = sdet(Ch);
      const R Je = R(1) / sqrt(detCh * det0);
      const R JJ = Je * Je - Je;
      // ---- Se = mu (G - Ch) + K Ch (Je^2 - Je) (:1473-1487) ----
      const R kj = K * JJ;
      Sym<R> Se;
      Se.a00 = mu * (G.a00 - Ch.a00) + kj * Ch.a00;
      Se.a11 = mu * (G.a11 - Ch.a11) + kj * Ch.a11;
      Se.a22 = mu * (G.a22 - Ch.a22) + kj * Ch.a22;
      Se.a01 = mu * (G.a01 - Ch.a01) + kj * Ch.a01;
      Se.a02 = mu * (G.a02 - Ch.a02) + kj * Ch.a02;
      Se.a12 = mu * (G.a12 - Ch.a12) + kj * Ch.a12;
      // ---- S = 1/2 (H^T Se H + Se H2) (ElemTotStress :1509-1528): S^22 in every
      //      iteration, the in-plane entries in the last (the ones kept) ----
      const M33<R> SeF = full(Se);
      // S^ij = 1/2 (sum_p H[p][i] SeH[p][j] + sum_p Se[p][i] H2[p][j]), SeH = Se H
      auto seh = [&](int p, int j) {
        return SeF.m[p][0] * H.m[0][j] + SeF.m[p][1] * H.m[1][j] + SeF.m[p][2] * H.m[2][j];
      };
      auto tot = [&](int i, int j) {
        return R(0.5) * (H.m[0][i] * seh(0, j) + H.m[1][i] * seh(1, j) + H.m[2][i] * seh(2, j) +
                         SeF.m[0][i] * H2.m[0][j] + SeF.m[1][i] * H2.m[1][j] +
                         SeF.m[2][i] * H2.m[2][j]);
      };
      if (last) {
        S00 = tot(0, 0);
        S11 = tot(1, 1);
        S01 = tot(0, 1);
      }
      const R S22 = tot(2, 2);
      // ---- CC^{2222} closed form (ElemElsTangMatTens + ElemTotTangMatTens) ----
      const R alpha = K * JJ - mu;
      const R beta = R(0.5) * K * Je * (R(2) * Je - R(1));
      const R hx = H.m[0][2], hy = H.m[1][2], hz = H.m[2][2];     // h = H e3
      const R Px = H2.m[0][2], Py = H2.m[1][2], Pz = H2.m[2][2];  // P = H h
      const R rx = Ch.a02, ry = Ch.a12, rz = Ch.a22;              // r = Ch e3
      const R Chx = Ch.a00 * hx + Ch.a01 * hy + Ch.a02 * hz;       // Ch h
      const R Chy = Ch.a01 * hx + Ch.a11 * hy + Ch.a12 * hz;
      const R Chz = Ch.a02 * hx + Ch.a12 * hy + Ch.a22 * hz;
      const R b = hx * Chx + hy * Chy + hz * Chz;
      const R BhP = Px * Chx + Py * Chy + Pz * Chz;
      const R CPx = Ch.a00 * Px + Ch.a01 * Py + Ch.a02 * Pz;
      const R CPy = Ch.a01 * Px + Ch.a11 * Py + Ch.a12 * Pz;
      const R CPz = Ch.a02 * Px + Ch.a12 * Py + Ch.a22 * Pz;
      const R BPP = Px * CPx + Py * CPy + Pz * CPz;
      const R Pr = Px * rx + Py * ry + Pz * rz;
      const R hr = hx * rx + hy * ry + hz * rz;
      const R CC = R(0.25) * (R(2) * (beta - alpha) * b * b + R(4) * beta * b * Pr -
                              R(4) * alpha * BhP * hr + R(2) * beta * Pr * Pr -
                              alpha * BPP * Ch.a22 - alpha * Pr * Pr);
      const R rCC = R(1) / CC;
      {
        const R nz = c33_update_noise(Ha, H2a, full(Ch), SeF, alpha, beta, S22, fabs(rCC));
        if (!(nz <= R(noise_max))) out.status |= kStatIllCond;
        out.noise = fmaxf(out.noise, (float)nz);
      }
      // ---- ModElemC33 (:1669-1707) ----
      const bool fail = !(fabs(CC) > R(1e-14));  // error ignored by the caller (:2086)
      if (fail) out.status |= kStatCC3333;
      if (!skip(fail, it)) {
        R dC = -S22 * rCC;
        const R cold = c33;
        R trial = cold + dC;
        // det(C.Cov) of the block-diagonal C: C00 (C11 C33) - C01 (C01 C33)
        if (c00 * (c11 * trial) - c01 * (c01 * trial) <= R(0)) {
          if (stop_on_clamp) {  // the element is handed over; its staged result is unused
            out.status |= kStatClamp;
            return out;
          }
          R lo = 0, hi = 1;
          for (int bk = 0; bk < 16; bk++) {
            const R mid = R(0.5) * (lo + hi);
            const R cm = cold + mid * dC;
            if (c00 * (c11 * cm) - c01 * (c01 * cm) > R(0)) lo = mid; else hi = mid;
          }
          dC *= lo;
          trial = cold + dC;
          out.status |= kStatClamp;
        }
        c33 = trial;
      }
      // ---- C_inv.Cont = INV(C.Cov) (:2101), block-diagonal: the in-plane inverse of
      //      the in-plane block, 1 / C33 ----
      ci00 = c11 * rD;
      ci01 = -c01 * rD;
      ci11 = c00 * rD;
      ci22 = R(1) / c33;
    }
    out.c33 = (double)c33;
    out.S00 = S00;
    out.S11 = S11;
    out.S01 = S01;
  }
  return out;
}

// Guard rule of ModElemC33 across qps: the reference returns at the first qp that fails the
// CC3333 guard (:1679), so that qp and every later qp keep their C33 in that iteration
// (iteration-outer loop order, :2058/:1660).  Bit `it` of the mask is set once some qp
// failed in iteration `it`.
struct SeqGuard {
  unsigned mask = 0u;
  BF_HD bool operator()(bool fail, int it) {
    if (fail) mask |= 1u << it;
    return ((mask >> it) & 1u) != 0u;
  }
};

// Through-thickness resultants for ElemUpdFint (:1859-1914), qp contributions added in qp
// order: N += w_q S_q, M += w_q theta_q S_q.
template <class R, class Res>
BF_HD void add_resultants(Res& out, const StagedQp<R>& s, double w, double theta_q) {
  const double wt = w * theta_q;
  out.N0 += w * (double)s.S00;
  out.N1 += w * (double)s.S11;
  out.N2 += w * (double)s.S01;
  out.M0 += wt * (double)s.S00;
  out.M1 += wt * (double)s.S11;
  out.M2 += wt * (double)s.S01;
}

// Staged chain of a whole element: every qp in order, resultants summed in qp order.
template <class R>
BF_HD Resultants<R> integrate_element(const V3<double>& a1c, const V3<double>& a2c,
                                      const V3<double>& a3c, const V3<double>& a31c,
                                      const V3<double>& a32c, const RefGeom& ref,
                                      const Activation& act, double h0, double mu_d,
                                      double K_d, int n_qp, int iters, const double* theta,
                                      const double* wq, double noise_max = kC33NoiseMax) {
  Resultants<R> out;
  out.N0 = out.N1 = out.N2 = out.M0 = out.M1 = out.M2 = 0.0;
  out.status = 0;
  out.noise = 0.0f;
  SeqGuard guard;
  for (int qp = 0; qp < n_qp; qp++) {
    const StagedQp<R> s =
        staged_qp<R>(a1c, a2c, a3c, a31c, a32c, ref, act, h0, mu_d, K_d, iters, theta[qp], guard,
                     false, noise_max);
    out.c33[qp] = s.c33;
    out.status |= s.status;
    out.noise = fmaxf(out.noise, s.noise);
    add_resultants(out, s, wq[qp], theta[qp]);
  }
  return out;
}

// FP32-mixed constitutive chain, written in difference form so that no FP32 subtraction
// cancels.  The reference forms G - Ce^-1 and Je^2 - Je directly (ElemElasStress
// :1477-1478), which in FP32 loses log10(1/strain) digits.  Here, with G = g0^{..} and
// M = (Ce^-1 - G) g0_{..}:
//   C^-1 - G      = -C^-1 (C - g0) G                  (C - g0 formed in FP64)
//   Ce^-1 - G     = T (C^-1 - G) T^T + (T G T^T - G)
//   T G T^T - G   = sum_r (d_r^2 - 1) psi_r psi_r^T    (Fa_cart = diag(d))
//                 = -gam (2 - gam f.f) psi psi^T,  psi = G (f . g0_i)   (Fa_cart = I - gam f f)
//   det(Ce^-1) det(g0) = det(I + M) = 1 + d,  d = tr M + minors2(M) + det M
//   Je - 1        = -d / (sqrt(1 + d) (1 + sqrt(1 + d)))
//   Se            = -mu (Ce^-1 - G) + K Je (Je - 1) Ce^-1
// The metrics, their inverses, C33 and its Newton update stay in FP64; stress, tangent and
// the 3x3 products run in FP32.  Algebraically identical to integrate_element.
template <class Skip>
BF_HD StagedQp<float> staged_qp_mixed(const V3<double>& a1c, const V3<double>& a2c,
                                      const V3<double>& a3c, const V3<double>& a31c,
                                      const V3<double>& a32c, const RefGeom& ref,
                                      const Activation& act, double h0, double mu_d,
                                      double K_d, int iters, double theta_q, Skip&& skip,
                                      bool stop_on_clamp = false) {
  typedef float R;
  StagedQp<float> out;
  out.status = 0;
  out.noise = 0.0f;
  out.S00 = out.S11 = out.S01 = 0.0f;
  out.c33 = 1.0;
  const R mu = (R)mu_d, K = (R)K_d;
  const double hhalf = 0.5 * h0;
  const bool rank1 = act.gam != 0.0;
  const bool diag_id = (act.d0 == 1.0 && act.d1 == 1.0 && act.d2 == 1.0);
  const double gam_d = act.gam;
  const double ff_d = dot(act.f, act.f);
  const double kap_d = rank1 ? gam_d / (1.0 - gam_d * ff_d) : 0.0;
  const R qcoef = (R)(-gam_d * (2.0 - gam_d * ff_d));  // T G T^T - G = qcoef psi psi^T
  const double dd[3] = {act.d0, act.d1, act.d2};
  {
    const double t3 = hhalf * theta_q;
    // ---- metrics in FP64 ----
    const V3<double> g0 = add(a1c, scale(t3, a31c));
    const V3<double> g1 = add(a2c, scale(t3, a32c));
    const V3<double> h0v = add(ref.a1, scale(t3, ref.a31));
    const V3<double> h1v = add(ref.a2, scale(t3, ref.a32));
    const V3<double> h2v = ref.a3;
    const Sym<double> gm = gram(g0, g1, a3c);
    const Sym<double> gm0 = gram(h0v, h1v, h2v);
    const Sym<double> Gd = sinv(gm0, (double*)nullptr);
    const Sym<double> gmI = sinv(gm, (double*)nullptr);
    // full-metric difference (gm - gm0) for the first iteration's C^-1 - G
    Sym<double> dg;
    dg.a00 = gm.a00 - gm0.a00; dg.a11 = gm.a11 - gm0.a11; dg.a22 = gm.a22 - gm0.a22;
    dg.a01 = gm.a01 - gm0.a01; dg.a02 = gm.a02 - gm0.a02; dg.a12 = gm.a12 - gm0.a12;
    const M33<R> G = full(Sym<R>{(R)Gd.a00, (R)Gd.a11, (R)Gd.a22, (R)Gd.a01, (R)Gd.a02, (R)Gd.a12});
    const M33<R> g0m = full(Sym<R>{(R)gm0.a00, (R)gm0.a11, (R)gm0.a22, (R)gm0.a01, (R)gm0.a02, (R)gm0.a12});
    // ---- activation: T = Fa.Cont g0, H = Fa_inv.Cov G, Q = T G T^T - G ----
    const double phid[3] = {dot(act.f, h0v), dot(act.f, h1v), dot(act.f, h2v)};
    M33<R> T, H, Q, Ha;
    {
      // psi = G phi (FP64 then rounded)
      double psid[3];
      for (int i = 0; i < 3; i++)
        psid[i] = (i == 0 ? Gd.a00 : i == 1 ? Gd.a01 : Gd.a02) * phid[0] +
                  (i == 0 ? Gd.a01 : i == 1 ? Gd.a11 : Gd.a12) * phid[1] +
                  (i == 0 ? Gd.a02 : i == 1 ? Gd.a12 : Gd.a22) * phid[2];
      // xi_r[i] = g0_i[r]: components of e_r on the reference covariant basis
      const double xi[3][3] = {{h0v.x, h1v.x, h2v.x}, {h0v.y, h1v.y, h2v.y}, {h0v.z, h1v.z, h2v.z}};
      // FaCov = S^T diag(d) S - gam phi phi^T, FiCov = S^T diag(1/d) S + kap phi phi^T
      double FaC[3][3], FiC[3][3];
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
          double a = 0.0, b = 0.0;
          if (diag_id) {
            const double gij = (i == j) ? (i == 0 ? gm0.a00 : i == 1 ? gm0.a11 : gm0.a22)
                                        : ((i + j == 1) ? gm0.a01 : (i + j == 2) ? gm0.a02 : gm0.a12);
            a = b = gij;
          } else {
            for (int r = 0; r < 3; r++) {
              a += dd[r] * xi[r][i] * xi[r][j];
              b += (1.0 / dd[r]) * xi[r][i] * xi[r][j];
            }
          }
          FaC[i][j] = a - gam_d * phid[i] * phid[j];
          FiC[i][j] = b + kap_d * phid[i] * phid[j];
        }
      const double Gf[3][3] = {{Gd.a00, Gd.a01, Gd.a02}, {Gd.a01, Gd.a11, Gd.a12}, {Gd.a02, Gd.a12, Gd.a22}};
      const double g0f[3][3] = {{gm0.a00, gm0.a01, gm0.a02}, {gm0.a01, gm0.a11, gm0.a12}, {gm0.a02, gm0.a12, gm0.a22}};
      // T = G FaCov G g0 = G FaCov (FP64: it multiplies the small C^-1 - G only, but its
      // identity part must be exact for the Q split to hold)
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
          double t = 0.0, hh = 0.0, ha = 0.0;
          for (int k = 0; k < 3; k++) {
            t += Gf[i][k] * FaC[k][j];
            hh += FiC[i][k] * Gf[k][j];
            ha += fabs(FiC[i][k]) * fabs(Gf[k][j]);
          }
          T.m[i][j] = (R)t;
          H.m[i][j] = (R)hh;
          Ha.m[i][j] = (R)ha;
        }
      (void)g0f;
      if (diag_id) {
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) Q.m[i][j] = qcoef * (R)psid[i] * (R)psid[j];
      } else {
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) Q.m[i][j] = 0;
        for (int r = 0; r < 3; r++) {
          double pr[3];
          for (int i = 0; i < 3; i++) pr[i] = Gf[i][0] * xi[r][0] + Gf[i][1] * xi[r][1] + Gf[i][2] * xi[r][2];
          const R c = (R)((dd[r] - 1.0) * (dd[r] + 1.0));
          for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) Q.m[i][j] += c * (R)pr[i] * (R)pr[j];
        }
      }
    }
    const M33<R> H2 = mul(H, H);
    const M33<R> H2a = abs_mul(Ha, Ha);
    // ---- C33 iteration ----
    const double c00 = gm.a00, c01 = gm.a01, c11 = gm.a11;
    double c33 = 1.0;
    // after the first update C.Cov is block-diagonal: its in-plane inverse does not depend
    // on C33
    const double rD = 1.0 / (c00 * c11 - c01 * c01);
    // E = C^-1 - G for the first iteration (C^-1 = gm^-1 in-plane block, 1 at [2][2])
    M33<R> E;
    {
      const M33<double> gmIf = full(gmI), dgf = full(dg), Gfull = full(Gd);
      const M33<double> X = mul(mul(gmIf, dgf), Gfull);  // gm^-1 (gm - gm0) gm0^-1
      E.m[0][0] = (R)(-X.m[0][0]); E.m[0][1] = (R)(-X.m[0][1]); E.m[1][0] = (R)(-X.m[1][0]);
      E.m[1][1] = (R)(-X.m[1][1]);
      E.m[0][2] = E.m[2][0] = (R)(-Gd.a02);
      E.m[1][2] = E.m[2][1] = (R)(-Gd.a12);
      E.m[2][2] = (R)(1.0 - Gd.a22);
    }
    R S00 = 0, S11 = 0, S01 = 0;
    for (int it = 0; it < iters; it++) {
      const bool last = it == iters - 1;
      // dC = T E T^T + Q  (Ce^-1 - G)
      const M33<R> TE = mul(T, E);
      M33<R> dC;
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
          dC.m[i][j] = TE.m[i][0] * T.m[j][0] + TE.m[i][1] * T.m[j][1] + TE.m[i][2] * T.m[j][2] + Q.m[i][j];
      // M = dC g0, d = det(I + M) - 1
      const M33<R> Mm = mul(dC, g0m);
      const R tr = Mm.m[0][0] + Mm.m[1][1] + Mm.m[2][2];
      const R mn2 = (Mm.m[0][0] * Mm.m[1][1] - Mm.m[0][1] * Mm.m[1][0]) +
                    (Mm.m[0][0] * Mm.m[2][2] - Mm.m[0][2] * Mm.m[2][0]) +
                    (Mm.m[1][1] * Mm.m[2][2] - Mm.m[1][2] * Mm.m[2][1]);
      const R dm = Mm.m[0][0] * (Mm.m[1][1] * Mm.m[2][2] - Mm.m[1][2] * Mm.m[2][1]) -
                   Mm.m[0][1] * (Mm.m[1][0] * Mm.m[2][2] - Mm.m[1][2] * Mm.m[2][0]) +
                   Mm.m[0][2] * (Mm.m[1][0] * Mm.m[2][1] - Mm.m[1][1] * Mm.m[2][0]);
      const R dlt = tr + mn2 + dm;
      const R sq = sqrt(R(1) + dlt);
      const R Je = R(1) / sq;
      const R Jm1 = -dlt / (sq * (R(1) + sq));
      const R JJ = Je * Jm1;
      // Ch = G + dC;  Se = -mu dC + K JJ Ch
      M33<R> Ch, Se;
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
          Ch.m[i][j] = G.m[i][j] + dC.m[i][j];
          Se.m[i][j] = -mu * dC.m[i][j] + K * JJ * Ch.m[i][j];
        }
      // S^ij = 1/2 (sum_p H[p][i] SeH[p][j] + sum_p Se[p][i] H2[p][j]), SeH = Se H: S^22 in
      // every iteration, the in-plane entries in the last (the ones kept)
      auto seh = [&](int p, int j) {
        return Se.m[p][0] * H.m[0][j] + Se.m[p][1] * H.m[1][j] + Se.m[p][2] * H.m[2][j];
      };
      auto tot = [&](int i, int j) {
        return R(0.5) * (H.m[0][i] * seh(0, j) + H.m[1][i] * seh(1, j) + H.m[2][i] * seh(2, j) +
                         Se.m[0][i] * H2.m[0][j] + Se.m[1][i] * H2.m[1][j] +
                         Se.m[2][i] * H2.m[2][j]);
      };
      if (last) {
        S00 = tot(0, 0);
        S11 = tot(1, 1);
        S01 = tot(0, 1);
      }
      const R S22 = tot(2, 2);
      // CC^{2222}
      const R alpha = K * JJ - mu;
      const R beta = R(0.5) * K * Je * (R(2) * Je - R(1));
      const R hx = H.m[0][2], hy = H.m[1][2], hz = H.m[2][2];
      const R Px = H2.m[0][2], Py = H2.m[1][2], Pz = H2.m[2][2];
      const R rx = Ch.m[0][2], ry = Ch.m[1][2], rz = Ch.m[2][2];
      const R Chx = Ch.m[0][0] * hx + Ch.m[0][1] * hy + Ch.m[0][2] * hz;
      const R Chy = Ch.m[1][0] * hx + Ch.m[1][1] * hy + Ch.m[1][2] * hz;
      const R Chz = Ch.m[2][0] * hx + Ch.m[2][1] * hy + Ch.m[2][2] * hz;
      const R b = hx * Chx + hy * Chy + hz * Chz;
      const R BhP = Px * Chx + Py * Chy + Pz * Chz;
      const R CPx = Ch.m[0][0] * Px + Ch.m[0][1] * Py + Ch.m[0][2] * Pz;
      const R CPy = Ch.m[1][0] * Px + Ch.m[1][1] * Py + Ch.m[1][2] * Pz;
      const R CPz = Ch.m[2][0] * Px + Ch.m[2][1] * Py + Ch.m[2][2] * Pz;
      const R BPP = Px * CPx + Py * CPy + Pz * CPz;
      const R Pr = Px * rx + Py * ry + Pz * rz;
      const R hr = hx * rx + hy * ry + hz * rz;
      const R CC = R(0.25) * (R(2) * (beta - alpha) * b * b + R(4) * beta * b * Pr -
                              R(4) * alpha * BhP * hr + R(2) * beta * Pr * Pr -
                              alpha * BPP * Ch.m[2][2] - alpha * Pr * Pr);
      const R rCC = R(1) / CC;
      {
        const R nz = c33_update_noise(Ha, H2a, Ch, Se, alpha, beta, S22, fabs(rCC));
        if (!(nz <= R(kC33NoiseMax))) out.status |= kStatIllCond;
        out.noise = fmaxf(out.noise, (float)nz);
      }
      // ModElemC33 in FP64 (the FP32 quotient's rounding is far below that of S^22 itself)
      const bool fail = !(fabs(CC) > R(1e-14));
      if (fail) out.status |= kStatCC3333;
      if (!skip(fail, it)) {
        double dCd = -(double)(S22 * rCC);
        const double cold = c33;
        double trial = cold + dCd;
        if (c00 * (c11 * trial) - c01 * (c01 * trial) <= 0.0) {
          if (stop_on_clamp) {  // the element is handed over; its staged result is unused
            out.status |= kStatClamp;
            return out;
          }
          double lo = 0, hi = 1;
          for (int bk = 0; bk < 16; bk++) {
            const double mid = 0.5 * (lo + hi);
            const double cm = cold + mid * dCd;
            if (c00 * (c11 * cm) - c01 * (c01 * cm) > 0.0) lo = mid; else hi = mid;
          }
          dCd *= lo;
          trial = cold + dCd;
          out.status |= kStatClamp;
        }
        c33 = trial;
      }
      // E = C^-1 - G = -C^-1 (C - g0) G with the block-diagonal C (FP64 inverse), for the
      // next iteration
      if (!last) {
        const double ci22 = 1.0 / c33;
        const double Ci[2][2] = {{c11 * rD, -c01 * rD}, {-c01 * rD, c00 * rD}};
        const double Dm[3][3] = {{dg.a00, dg.a01, -gm0.a02},
                                 {dg.a01, dg.a11, -gm0.a12},
                                 {-gm0.a02, -gm0.a12, c33 - gm0.a22}};
        const double Gf[3][3] = {{Gd.a00, Gd.a01, Gd.a02}, {Gd.a01, Gd.a11, Gd.a12}, {Gd.a02, Gd.a12, Gd.a22}};
        double CD[3][3];
        for (int j = 0; j < 3; j++) {
          CD[0][j] = Ci[0][0] * Dm[0][j] + Ci[0][1] * Dm[1][j];
          CD[1][j] = Ci[1][0] * Dm[0][j] + Ci[1][1] * Dm[1][j];
          CD[2][j] = ci22 * Dm[2][j];
        }
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++)
            E.m[i][j] = (R)(-(CD[i][0] * Gf[0][j] + CD[i][1] * Gf[1][j] + CD[i][2] * Gf[2][j]));
      }
    }
    out.c33 = c33;
    out.S00 = S00;
    out.S11 = S11;
    out.S01 = S01;
  }
  return out;
}

BF_HD Resultants<float> integrate_element_mixed(const V3<double>& a1c, const V3<double>& a2c,
                                                const V3<double>& a3c, const V3<double>& a31c,
                                                const V3<double>& a32c, const RefGeom& ref,
                                                const Activation& act, double h0, double mu_d,
                                                double K_d, int n_qp, int iters,
                                                const double* theta, const double* wq) {
  Resultants<float> out;
  out.N0 = out.N1 = out.N2 = out.M0 = out.M1 = out.M2 = 0.0;
  out.status = 0;
  out.noise = 0.0f;
  SeqGuard guard;
  for (int qp = 0; qp < n_qp; qp++) {
    const StagedQp<float> s = staged_qp_mixed(a1c, a2c, a3c, a31c, a32c, ref, act, h0, mu_d, K_d,
                                              iters, theta[qp], guard);
    out.c33[qp] = s.c33;
    out.status |= s.status;
    out.noise = fmaxf(out.noise, s.noise);
    add_resultants(out, s, wq[qp], theta[qp]);
  }
  return out;
}

// Elements whose staged result must be replaced by the reference-order one: the local C33
// iteration left the range where the result stays within the path's error budget (a clamp
// of ModElemC33's bisection, C33 outside [kC33Lo, kC33Hi] at any qp, an update whose
// reference rounding bound c33_update_noise exceeds kC33NoiseMax, the CC3333 guard, or a
// non-finite value).  There the 3 half-Newton steps of the reference amplify any rounding
// difference by orders of magnitude, so the fast paths hand such elements to the plain
// kernel.  The rounding bound is the sharp test; the range keeps out near-degenerate
// thickness stretches (C33 -> 0), where the bound's first-order model no longer holds and,
// for FP32, where the staged C33 itself loses digits.  On the noisy 256 x 192 membrane the
// FP64 settings hand over 6.3 % of the elements and leave the staged result within 1.2e-14
// of the reference order (force, norm-wise; C33 within 3.4e-13 per element).  The FP32
// chain keeps [0.4, 2.5] (a range of [0.3, 3] there reaches 4e-6); what it cannot hold goes
// to the FP64 chain at the looser kC33NoiseMaxTier2, and only the rest to the reference
// order: 3.7 % of the membrane's elements, the mixed result within 4e-7 (CPU build of this
// header; 1e-6 on noisier states of the same mesh).
constexpr double kC33Lo = 0.1, kC33Hi = 10.0;
constexpr double kC33LoMixed = 0.4, kC33HiMixed = 2.5;
BF_HD int c33_range_bit(double c33, bool mixed) {
  const double lo = mixed ? kC33LoMixed : kC33Lo, hi = mixed ? kC33HiMixed : kC33Hi;
  return (c33 >= lo && c33 <= hi) ? 0 : kStatRange;
}
BF_HD bool needs_exact_bits(int status, bool finite) {
  return !finite || (status & (kStatClamp | kStatCC3333 | kStatIllCond | kStatRange)) != 0;
}
BF_HD bool needs_exact(const double* c33, int n_qp, int status, bool finite, bool mixed) {
  for (int q = 0; q < n_qp; q++) status |= c33_range_bit(c33[q], mixed);
  return needs_exact_bits(status, finite);
}

// Nodal force of one patch node a (ElemUpdFint :1879-1913 / :1934-1983), summed over qps:
//   Fb_a = pref0 [ (Na0 N0 + Na1 N2) a1 + (Na1 N1 + Na0 N2) a2
//                  - (h0/2) (c0 M0 + c1 M1 + 2 c2 M2) a3 ],
//   c_k = Nab_k + (Na0 gc1 + Na1 gc2) . A_k,  A = (Aaa, Abb, Aab).
BF_HD V3<double> node_force(double Na0, double Na1, double Nb0, double Nb1, double Nb2,
                            const MidGeom& G, double pref0, double half_h0, double N0,
                            double N1, double N2, double M0, double M1, double M2) {
  const V3<double> ng = add(scale(Na0, G.gc1), scale(Na1, G.gc2));
  const double c0 = Nb0 + dot(ng, G.Aaa);
  const double c1 = Nb1 + dot(ng, G.Abb);
  const double c2 = Nb2 + dot(ng, G.Aab);
  const double k1 = Na0 * N0 + Na1 * N2;
  const double k2 = Na1 * N1 + Na0 * N2;
  const double k3 = -half_h0 * (c0 * M0 + c1 * M1 + 2.0 * c2 * M2);
  return mk(pref0 * (k1 * G.a1.x + k2 * G.a2.x + k3 * G.a3.x),
            pref0 * (k1 * G.a1.y + k2 * G.a2.y + k3 * G.a3.y),
            pref0 * (k1 * G.a1.z + k2 * G.a2.z + k3 * G.a3.z));
}

}  // namespace bf
