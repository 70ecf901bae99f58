// bf_exact.hip -- kernels compiled without floating-point contraction: the geom0 cache
// and the plain FP64 element kernel in the reference's operation order (bf_exact.h).
// Build flag: -ffp-contract=off (so a*b+c rounds twice, as in the reference's C).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bf_exact.h"
#include "bf_kernels.h"

namespace bf {

// Reference midsurface cache (ElemUpdateGeom0Subdiv over all cells, dmplex_geom.c:859-865),
// stored as a1, a2, a3, a3_1, a3_2 -- everything ElemUpdateG reads from geom0 (:925-931).
__global__ void __launch_bounds__(256) bf_geom0_kernel(DevMesh m, int E,
                                                       const double* __restrict__ x,
                                                       const double* __restrict__ y,
                                                       const double* __restrict__ z,
                                                       double* __restrict__ ref) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const int meta = m.meta[e];
  const int kind = meta & 3, val = (meta >> 8) & 255, nen = (meta >> 16) & 255;
  MidGeom G;
  auto xyz = [&](int i) { return ld3(x, y, z, m.patch[(size_t)i * E + e]); };
  if (kind == kRegular) {
    G = mid_geom(12, c_Na[0], c_Na[1], c_Nab[0], c_Nab[1], c_Nab[2], xyz);
  } else if (kind == kIrregular) {
    const double(*t)[kMaxNen] = c_IN[val];
    G = mid_geom(nen, t[0], t[1], t[2], t[3], t[4], xyz);
  } else {
    G = cap_geom(ld3(x, y, z, m.corner[e]), ld3(x, y, z, m.corner[(size_t)E + e]),
                 ld3(x, y, z, m.corner[(size_t)2 * E + e]));
  }
  V3<double> a31, a32;
  a3_alpha(G.a1, G.a2, G.Aaa, G.Aab, G.Abb, a31, a32);
  const V3<double> v[5] = {G.a1, G.a2, G.a3, a31, a32};
#pragma unroll
  for (int k = 0; k < 5; k++) {
    ref[(size_t)(3 * k + 0) * E + e] = v[k].x;
    ref[(size_t)(3 * k + 1) * E + e] = v[k].y;
    ref[(size_t)(3 * k + 2) * E + e] = v[k].z;
  }
}


// Plain FP64 element kernel: ElemUpdateGeomSubdiv + ElemUpdateG + ElemActDefGrad +
// ElemCGDefTens + ElemC33Solve + ElemUpdFint of one element per lane, every operation in
// the order of src/active_strain.c.  Writes the same slot-major element-force buffer as the
// staged kernel; the node gather then sums in element order (active_strain.c:2238-2260).
constexpr int kExactBlock = 128;

__device__ __forceinline__ void exact_element(const DevMesh& m, const KParams& kp,
                                              const double* __restrict__ x,
                                              const double* __restrict__ y,
                                              const double* __restrict__ z, int e,
                                              unsigned char extra_flags,
                                              exact::ElemIn& in, exact::ElemOut& o) {
  const int E = kp.E;
  const int meta = m.meta[e];
  const int kind = meta & 3, val = (meta >> 8) & 255, nen = (meta >> 16) & 255;
  auto xyz = [&](int i) { return ld3(x, y, z, m.patch[(size_t)i * E + e]); };
  const double *na0, *na1, *nb0, *nb1, *nb2;
  if (kind == kRegular) {
    na0 = c_Na[0]; na1 = c_Na[1]; nb0 = c_Nab[0]; nb1 = c_Nab[1]; nb2 = c_Nab[2];
    in.cur = mid_geom(12, na0, na1, nb0, nb1, nb2, xyz);
  } else if (kind == kIrregular) {
    const double(*t)[kMaxNen] = c_IN[val];
    na0 = t[0]; na1 = t[1]; nb0 = t[2]; nb1 = t[3]; nb2 = t[4];
    in.cur = mid_geom(nen, na0, na1, nb0, nb1, nb2, xyz);
  } else {
    na0 = na1 = nb0 = nb1 = nb2 = nullptr;
    in.cur = cap_geom(ld3(x, y, z, m.corner[e]), ld3(x, y, z, m.corner[(size_t)E + e]),
                      ld3(x, y, z, m.corner[(size_t)2 * E + e]));
  }
  a3_alpha(in.cur.a1, in.cur.a2, in.cur.Aaa, in.cur.Aab, in.cur.Abb, in.a31c, in.a32c);
  const double* rf = m.ref;
  in.ref.a1 = mk(rf[(size_t)0 * E + e], rf[(size_t)1 * E + e], rf[(size_t)2 * E + e]);
  in.ref.a2 = mk(rf[(size_t)3 * E + e], rf[(size_t)4 * E + e], rf[(size_t)5 * E + e]);
  in.ref.a3 = mk(rf[(size_t)6 * E + e], rf[(size_t)7 * E + e], rf[(size_t)8 * E + e]);
  in.ref.a31 = mk(rf[(size_t)9 * E + e], rf[(size_t)10 * E + e], rf[(size_t)11 * E + e]);
  in.ref.a32 = mk(rf[(size_t)12 * E + e], rf[(size_t)13 * E + e], rf[(size_t)14 * E + e]);
  // ---- Fa_cart (ElemActDefGrad :1208-1256) ----
  double g = kp.has_delay ? ramp_gamma(kp.ramp, m.delay[e]) : kp.gamma_t;
  g *= m.gscale[e];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) in.Fa_cart[i][j] = 0.0;
  if (kp.act_mode == 0) {
    const double nf[3] = {m.fib[e], m.fib[(size_t)E + e], m.fib[(size_t)2 * E + e]};
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) in.Fa_cart[i][j] = (i == j ? 1.0 : 0.0) - g * nf[i] * nf[j];
  } else if (kp.act_mode == 1) {
    in.Fa_cart[0][0] = 1.0 - kp.man_gamma;
    in.Fa_cart[1][1] = 1.0;
    in.Fa_cart[2][2] = 1.0;
  } else {
    in.Fa_cart[0][0] = 1.0 - g;
    in.Fa_cart[1][1] = 1.0 - g;
    in.Fa_cart[2][2] = 1.0 / (in.Fa_cart[0][0] * in.Fa_cart[1][1]);
  }
  in.h0 = kp.h0;
  in.mu = kp.mu;
  in.K = kp.K;
  in.n_qp = kp.n_qp;
  in.iters = kp.iters;
  in.theta = kp.theta;
  in.wq = kp.w;
  exact::integrate(in, o);
  for (int q = 0; q < kp.n_qp; q++) m.c33[(size_t)q * E + e] = o.c33[q];
  // ---- ElemUpdFint :1799-1801 ----
  const double A0 = m.dA0[e];
  const double half_h0 = 0.5 * kp.h0;
  const double pref0 = 0.5 * kp.h0 * A0;
  double* Fe = m.Fe;
  bool finite = true;
  if (kind != kCap) {
    for (int a = 0; a < nen; a++) {
      const V3<double> f = exact::node_force_exact(na0[a], na1[a], nb0[a], nb1[a], nb2[a], in.cur,
                                                   pref0, half_h0, o, kp.theta, kp.w, kp.n_qp,
                                                   kind == kIrregular);
      finite = finite && isfinite(f.x) && isfinite(f.y) && isfinite(f.z);
      Fe[(size_t)(3 * a + 0) * E + e] = f.x;
      Fe[(size_t)(3 * a + 1) * E + e] = f.y;
      Fe[(size_t)(3 * a + 2) * E + e] = f.z;
    }
  } else {
    for (int a = 0; a < nen; a++) {
      Fe[(size_t)(3 * a + 0) * E + e] = 0.0;
      Fe[(size_t)(3 * a + 1) * E + e] = 0.0;
      Fe[(size_t)(3 * a + 2) * E + e] = 0.0;
    }
    const double dN[3][2] = {{-1.0, -1.0}, {1.0, 0.0}, {0.0, 1.0}};
    for (int a = 0; a < 3; a++) {
      const int s = m.corner[(size_t)(3 + a) * E + e];
      if (s < 0) continue;
      const V3<double> f = exact::cap_force_exact(dN[a][0], dN[a][1], in.cur, pref0, o, kp.w, kp.n_qp);
      finite = finite && isfinite(f.x) && isfinite(f.y) && isfinite(f.z);
      Fe[(size_t)(3 * s + 0) * E + e] = f.x;
      Fe[(size_t)(3 * s + 1) * E + e] = f.y;
      Fe[(size_t)(3 * s + 2) * E + e] = f.z;
    }
  }
  if (!finite) atomicMin(&m.status[0], e);
  m.flags[e] = (unsigned char)(o.status | (finite ? 0 : kStatNonFinite) | extra_flags);
  if (o.status & kStatCC3333) {
    atomicMin(&m.status[2], e);
    atomicAdd(&m.status[3], 1);
  }
}


__global__ void __launch_bounds__(kExactBlock) bf_element_kernel_exact(DevMesh m, KParams kp,
                                                               const double* __restrict__ x,
                                                               const double* __restrict__ y,
                                                               const double* __restrict__ z) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= kp.E) return;
  exact::ElemIn in;
  exact::ElemOut o;
  exact_element(m, kp, x, y, z, e, 0, in, o);
}

// Persistent drain of the fallback list: every lane strides over [0, count); count is
// final because the producing kernel completed earlier on the same stream.
__global__ void __launch_bounds__(kExactBlock) bf_element_kernel_exact_list(DevMesh m, KParams kp,
                                                                    const double* __restrict__ x,
                                                                    const double* __restrict__ y,
                                                                    const double* __restrict__ z) {
  const int count = m.status[4];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
    exact::ElemIn in;
    exact::ElemOut o;
    exact_element(m, kp, x, y, z, m.fallback[i], kStatFallback, in, o);
  }
}

hipError_t upload_in_tables_exact(const double* tab, size_t bytes) {
  return hipMemcpyToSymbol(HIP_SYMBOL(c_IN), tab, bytes);
}

hipError_t launch_geom0(const DevMesh& m, int E, const double* x, const double* y,
                        const double* z, double* ref, hipStream_t s) {
  if (E <= 0) return hipSuccess;
  hipLaunchKernelGGL(bf_geom0_kernel, dim3((E + 255) / 256), dim3(256), 0, s, m, E, x, y, z, ref);
  return hipGetLastError();
}

hipError_t launch_element_exact_fallback(const DevMesh& m, const KParams& kp, const double* x,
                                         const double* y, const double* z, hipStream_t s) {
  if (kp.E <= 0) return hipSuccess;
  // one lane per listed element when the list fills the machine (2048 waves of 64)
  const int blocks = std::min((kp.E + 127) / 128, 1024);
  hipLaunchKernelGGL(bf_element_kernel_exact_list, dim3(blocks), dim3(kExactBlock), 0, s, m, kp, x, y, z);
  return hipGetLastError();
}

hipError_t launch_element_exact(const DevMesh& m, const KParams& kp, const double* x,
                                const double* y, const double* z, hipStream_t s) {
  if (kp.E <= 0) return hipSuccess;
  hipLaunchKernelGGL(bf_element_kernel_exact, dim3((kp.E + kExactBlock - 1) / kExactBlock),
                     dim3(kExactBlock), 0, s, m, kp, x, y, z);
  return hipGetLastError();
}

}  // namespace bf
