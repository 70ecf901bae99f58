// elem_shim.cpp -- CPU harness for the staged element arithmetic in
// biofem-ia_amd/csrc/bf_elem.h (the same header the gfx950 kernels include).
//
// Test-only: it lets the CPU test suite check the staged constitutive algebra and the
// FP32-mixed error budget against the oracle without a GPU, and measure the reference
// order's own rounding against the staged chain in x87 extended precision (precision 6).
// It is never linked into or called by the product library.
// precision: 0 staged FP64, 1 FP32-mixed, 2 reference order, 3 staged FP32,
//            4 / 5 staged FP64 / FP32-mixed with the hand-over, 6 staged extended.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../biofem-ia_amd/csrc/bf_elem.h"
#include "../biofem-ia_amd/csrc/bf_exact.h"

using namespace bf;

static const double kNa[2][12] = {
    {-0.0247, -0.0309, 0.0000, -0.4815, -0.1852, 0.0247, 0.4815, 0.0000, -0.0062, 0.0309,
     0.1852, 0.0062},
    {-0.0309, -0.0247, -0.1852, -0.4815, 0.0000, -0.0062, 0.0000, 0.4815, 0.0247, 0.0062,
     0.1852, 0.0309}};
static const double kNab[3][12] = {
    {0.1111, 0.2222, -0.2222, -0.2222, 0.4444, 0.1111, -0.2222, -0.8889, 0.0000, 0.2222,
     0.4444, 0.0000},
    {0.2222, 0.1111, 0.4444, -0.2222, -0.2222, 0.0000, -0.8889, -0.2222, 0.1111, 0.0000,
     0.4444, 0.2222},
    {0.1667, 0.1667, -0.1111, 0.2222, -0.1111, -0.0556, -0.4444, -0.4444, -0.0556, 0.0556,
     0.5556, 0.0556}};

extern "C" int shim_finternal(int n_v, int n_nodes, int n_elmt, int stride, const int* patch,
                              const int* ire, const int* val, const double* in_tables,
                              const double* x0, const double* y0, const double* z0,
                              const double* x, const double* y, const double* z,
                              const double* dA0, const double* fib, const double* gam_e,
                              const double* diag /*3*/, double h0, double mu, double K,
                              int n_qp, int iters, const double* theta, const double* wq,
                              int precision, double* fint, double* c33_out) {
  // in_tables: [13][5][18] IN rows per valence (only v=3..12 filled)
  for (int i = 0; i < 3 * n_nodes; i++) fint[i] = 0.0;
  int warn = 0;
  std::vector<double> Fb(3 * kMaxNen);
  for (int e = 0; e < n_elmt; e++) {
    const int* pt = patch + (size_t)stride * e;
    const int kind = ire[e];
    const int nen = kind == 0 ? 12 : val[e] + 6;
    const double *na0, *na1, *nb0, *nb1, *nb2;
    if (kind == 0) {
      na0 = kNa[0]; na1 = kNa[1]; nb0 = kNab[0]; nb1 = kNab[1]; nb2 = kNab[2];
    } else {
      const double* t = in_tables + (size_t)val[e] * 5 * 18;
      na0 = t; na1 = t + 18; nb0 = t + 36; nb1 = t + 54; nb2 = t + 72;
    }
    auto cur = [&](int i) { int n = pt[i]; return mk(x[n], y[n], z[n]); };
    auto ref = [&](int i) { int n = pt[i]; return mk(x0[n], y0[n], z0[n]); };
    MidGeom Gc = mid_geom(nen, na0, na1, nb0, nb1, nb2, cur);
    MidGeom G0 = mid_geom(nen, na0, na1, nb0, nb1, nb2, ref);
    RefGeom R0;
    R0.a1 = G0.a1; R0.a2 = G0.a2; R0.a3 = G0.a3;
    a3_alpha(G0.a1, G0.a2, G0.Aaa, G0.Aab, G0.Abb, R0.a31, R0.a32);
    V3<double> a31c, a32c;
    a3_alpha(Gc.a1, Gc.a2, Gc.Aaa, Gc.Aab, Gc.Abb, a31c, a32c);
    Activation act;
    act.d0 = diag[0]; act.d1 = diag[1]; act.d2 = diag[2];
    act.gam = gam_e[e];
    act.f = mk(fib[3 * e], fib[3 * e + 1], fib[3 * e + 2]);
    double N0, N1, N2, M0, M1, M2;
    int st;
    bool do_exact = precision == 2;
    if (precision == 4 || precision == 5) {  // staged (4: FP64, 5: mixed) with exact fallback
      Resultants<double> rd;
      Resultants<float> rf;
      const double* cc;
      int stt;
      bool fin;
      if (precision == 4) {
        rd = integrate_element<double>(Gc.a1, Gc.a2, Gc.a3, a31c, a32c, R0, act, h0, mu, K, n_qp, iters, theta, wq);
        cc = rd.c33; stt = rd.status; fin = std::isfinite(rd.N0) && std::isfinite(rd.M0);
        N0 = rd.N0; N1 = rd.N1; N2 = rd.N2; M0 = rd.M0; M1 = rd.M1; M2 = rd.M2;
      } else {
        rf = integrate_element_mixed(Gc.a1, Gc.a2, Gc.a3, a31c, a32c, R0, act, h0, mu, K, n_qp, iters, theta, wq);
        cc = rf.c33; stt = rf.status; fin = std::isfinite(rf.N0) && std::isfinite(rf.M0);
        N0 = rf.N0; N1 = rf.N1; N2 = rf.N2; M0 = rf.M0; M1 = rf.M1; M2 = rf.M2;
      }
      bool hand = needs_exact(cc, n_qp, stt, fin, precision == 5);
      if (hand && precision == 5) {  // second tier of the mixed path: the FP64 staged chain
        rd = integrate_element<double>(Gc.a1, Gc.a2, Gc.a3, a31c, a32c, R0, act, h0, mu, K, n_qp,
                                       iters, theta, wq, kC33NoiseMaxTier2);
        cc = rd.c33; stt = rd.status; fin = std::isfinite(rd.N0) && std::isfinite(rd.M0);
        N0 = rd.N0; N1 = rd.N1; N2 = rd.N2; M0 = rd.M0; M1 = rd.M1; M2 = rd.M2;
        hand = needs_exact(cc, n_qp, stt, fin, false);
      }
      if (hand) {
        do_exact = true;
        warn++;
      } else {
        if (c33_out) for (int q = 0; q < n_qp; q++) c33_out[(size_t)e * n_qp + q] = cc[q];
        const double pref0 = 0.5 * h0 * dA0[e];
        for (int a = 0; a < nen; a++) {
          V3<double> f = node_force(na0[a], na1[a], nb0[a], nb1[a], nb2[a], Gc, pref0, 0.5 * h0,
                                    N0, N1, N2, M0, M1, M2);
          const int n = pt[a];
          if (n == kSentinel || n < 0 || n >= n_v) continue;
          fint[3 * n] += f.x;
          fint[3 * n + 1] += f.y;
          fint[3 * n + 2] += f.z;
        }
        continue;
      }
    }
    if (do_exact) {  // reference operation order (bf_exact.h)
      exact::ElemIn in;
      in.cur = Gc;
      in.a31c = a31c;
      in.a32c = a32c;
      in.ref = R0;
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
          in.Fa_cart[i][j] = 0.0;
      const double nf[3] = {fib[3 * e], fib[3 * e + 1], fib[3 * e + 2]};
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) in.Fa_cart[i][j] = (i == j ? 1.0 : 0.0) - gam_e[e] * nf[i] * nf[j];
      in.h0 = h0; in.mu = mu; in.K = K; in.n_qp = n_qp; in.iters = iters; in.theta = theta; in.wq = wq;
      exact::ElemOut o;
      exact::integrate(in, o);
      if (precision == 2 && o.status) warn++;
      if (c33_out) for (int q = 0; q < n_qp; q++) c33_out[(size_t)e * n_qp + q] = o.c33[q];
      const double pref0 = 0.5 * h0 * dA0[e];
      for (int a = 0; a < nen; a++) {
        V3<double> f = exact::node_force_exact(na0[a], na1[a], nb0[a], nb1[a], nb2[a], Gc, pref0,
                                               0.5 * h0, o.Sv, theta, wq, n_qp, kind != 0);
        const int n = pt[a];
        if (n == kSentinel || n < 0 || n >= n_v) continue;
        fint[3 * n] += f.x;
        fint[3 * n + 1] += f.y;
        fint[3 * n + 2] += f.z;
      }
      continue;
    }
    if (precision == 6) {  // the staged chain in x87 extended precision (64-bit mantissa)
      Resultants<long double> r = integrate_element<long double>(
          Gc.a1, Gc.a2, Gc.a3, a31c, a32c, R0, act, h0, mu, K, n_qp, iters, theta, wq);
      N0 = r.N0; N1 = r.N1; N2 = r.N2; M0 = r.M0; M1 = r.M1; M2 = r.M2; st = r.status;
      if (c33_out) for (int q = 0; q < n_qp; q++) c33_out[(size_t)e * n_qp + q] = r.c33[q];
    } else if (precision == 0) {
      Resultants<double> r = integrate_element<double>(Gc.a1, Gc.a2, Gc.a3, a31c, a32c, R0, act,
                                                       h0, mu, K, n_qp, iters, theta, wq);
      N0 = r.N0; N1 = r.N1; N2 = r.N2; M0 = r.M0; M1 = r.M1; M2 = r.M2; st = r.status;
      if (c33_out) for (int q = 0; q < n_qp; q++) c33_out[(size_t)e * n_qp + q] = r.c33[q];
    } else {
      Resultants<float> r = (precision == 3)
          ? integrate_element<float>(Gc.a1, Gc.a2, Gc.a3, a31c, a32c, R0, act, h0, mu, K, n_qp,
                                     iters, theta, wq)
          : integrate_element_mixed(Gc.a1, Gc.a2, Gc.a3, a31c, a32c, R0, act, h0, mu, K, n_qp,
                                    iters, theta, wq);
      N0 = r.N0; N1 = r.N1; N2 = r.N2; M0 = r.M0; M1 = r.M1; M2 = r.M2; st = r.status;
      if (c33_out) for (int q = 0; q < n_qp; q++) c33_out[(size_t)e * n_qp + q] = r.c33[q];
    }
    if (st) warn++;
    const double pref0 = 0.5 * h0 * dA0[e];
    for (int a = 0; a < nen; a++) {
      V3<double> f = node_force(na0[a], na1[a], nb0[a], nb1[a], nb2[a], Gc, pref0, 0.5 * h0,
                                N0, N1, N2, M0, M1, M2);
      const int n = pt[a];
      if (n == kSentinel || n < 0 || n >= n_v) continue;
      fint[3 * n] += f.x;
      fint[3 * n + 1] += f.y;
      fint[3 * n + 2] += f.z;
    }
  }
  return warn;
}

// ---- the single-pass kernel's element tiling (biofem-ia_amd/csrc/bf_tiles.cpp), exported
// for the CPU checks of its index invariants.  Call with null outputs to get the sizes:
// sizes = {n_tiles, local nodes, n_partials, contributions}.
#include "../biofem-ia_amd/csrc/bf_tiles.h"

extern "C" int shim_tiles(int E, int n_v, int n_nodes, int NS, const int* patch, const int* nen,
                          const double* x0, const double* y0, const double* z0, int T,
                          int* sizes, int* elem, unsigned short* lslot, int* node_ptr, int* node,
                          int* out, int* cptr, unsigned short* cent, int* pptr,
                          unsigned char* complete) {
  bf::TileHost th;
  const int rc = bf::build_tiles(E, n_v, n_nodes, NS, patch, nen, x0, y0, z0, T, th);
  if (rc != 0) return rc;
  sizes[0] = th.n_tiles;
  sizes[1] = (int)th.node.size();
  sizes[2] = th.n_partials;
  sizes[3] = (int)th.cent.size();
  if (!elem) return 0;
  std::copy(th.elem.begin(), th.elem.end(), elem);
  std::copy(th.lslot.begin(), th.lslot.end(), lslot);
  std::copy(th.node_ptr.begin(), th.node_ptr.end(), node_ptr);
  std::copy(th.node.begin(), th.node.end(), node);
  std::copy(th.out.begin(), th.out.end(), out);
  std::copy(th.cptr.begin(), th.cptr.end(), cptr);
  std::copy(th.cent.begin(), th.cent.end(), cent);
  std::copy(th.pptr.begin(), th.pptr.end(), pptr);
  std::copy(th.complete.begin(), th.complete.end(), complete);
  return 0;
}
