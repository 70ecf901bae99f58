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
// ElemCGDefTens + ElemC33Solve + ElemUpdFint, every operation in the order of
// src/active_strain.c.  An element occupies a group of P adjacent lanes of one wave (P = the
// largest n_qp of the launch; floor(64 / P) groups per wave), lane q of the group running
// through-thickness point q: each lane forms the element geometry, runs its qp's C33 loop,
// then the group exchanges the qp stresses by cross-lane reads and each lane writes the
// forces of every P-th patch node, summing over qp in the reference's order.  The
// CC3333 guard, which in the reference stops the qp loop of an iteration at the first
// failing qp, becomes a wave ballot per iteration.  Writes the same slot-row element-force
// buffer as the staged kernels; the node gather then sums in element order
// (active_strain.c:2238-2260).
constexpr int kExactBlock = 128;

struct QpGroup {
  int e;      // element of this lane's group, -1 when the lane has no work
  int q;      // through-thickness point of this lane
  int base;   // wave lane of the group's q = 0
};

__device__ __forceinline__ void exact_element(const DevMesh& m, const KParams& kp,
                                              const double* __restrict__ x,
                                              const double* __restrict__ y,
                                              const double* __restrict__ z, const QpGroup& grp,
                                              unsigned char extra_flags) {
  const int E = kp.E;
  const int e = grp.e, q = grp.q;
  const int n_qp = kp.n_qp;
  if (q >= n_qp) return;  // lanes beyond this element's qps (mixed n_qp in one launch)
  const int meta = m.meta[e];
  const int kind = meta & 3, val = (meta >> 8) & 255, nen = (meta >> 16) & 255;
  auto xyz = [&](int i) { return ld3(x, y, z, m.patch[(size_t)i * E + e]); };
  exact::ElemIn in;
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
  in.n_qp = n_qp;
  in.iters = kp.iters;
  in.theta = kp.theta;
  in.wq = kp.w;
  // ---- this lane's qp; the guard of iteration it skips qp q if any qp <= q failed ----
  const unsigned long long upto = (2ull << q) - 1ull;
  const exact::QpOut r = exact::integrate_qp(in, q, [&](bool fail, int) {
    return ((__ballot(fail) >> grp.base) & upto) != 0ull;
  });
  m.c33[(size_t)q * E + e] = r.c33;
  // ---- the group's qp stresses, in qp order ----
  double Sv[kMaxQp][3];
  int status = 0;
#pragma unroll
  for (int j = 0; j < kMaxQp; j++) {
    if (j < n_qp) {
      Sv[j][0] = __shfl(r.S00, grp.base + j);
      Sv[j][1] = __shfl(r.S11, grp.base + j);
      Sv[j][2] = __shfl(r.S01, grp.base + j);
      status |= __shfl(r.status, grp.base + j);
    } else {
      Sv[j][0] = Sv[j][1] = Sv[j][2] = 0.0;
    }
  }
  // ---- ElemUpdFint :1799-1801, nodes q, q + n_qp, ... of the patch ----
  const double A0 = m.dA0[e];
  const double half_h0 = 0.5 * kp.h0;
  const double pref0 = 0.5 * kp.h0 * A0;
  double* Fe = m.Fe;
  bool finite = true;
  if (kind != kCap) {
    for (int a = q; a < nen; a += n_qp) {
      const V3<double> f = exact::node_force_exact(na0[a], na1[a], nb0[a], nb1[a], nb2[a], in.cur,
                                                   pref0, half_h0, Sv, kp.theta, kp.w, n_qp,
                                                   kind == kIrregular);
      finite = finite && isfinite(f.x) && isfinite(f.y) && isfinite(f.z);
      put_force(Fe, e, E, a, f);
    }
  } else if (q == 0) {
    // zeroed element vector, corner contributions added (:1807-1851)
    const double dN[3][2] = {{-1.0, -1.0}, {1.0, 0.0}, {0.0, 1.0}};
    V3<double> fc[3];
    int sl[3];
    for (int k = 0; k < 3; k++) {
      sl[k] = m.corner[(size_t)(3 + k) * E + e];
      fc[k] = exact::cap_force_exact(dN[k][0], dN[k][1], in.cur, pref0, Sv, kp.w, n_qp);
    }
    for (int a = 0; a < nen; a++) {
      V3<double> f = mk(0.0, 0.0, 0.0);
      for (int k = 0; k < 3; k++)
        if (sl[k] == a) {
          f = mk(f.x + fc[k].x, f.y + fc[k].y, f.z + fc[k].z);
          finite = finite && isfinite(fc[k].x) && isfinite(fc[k].y) && isfinite(fc[k].z);
        }
      put_force(Fe, e, E, a, f);
    }
  }
  // ---- element status from the group (non-finite from any lane's nodes) ----
  int bad = finite ? 0 : 1;
#pragma unroll
  for (int j = 0; j < kMaxQp; j++)
    if (j < n_qp) bad |= __shfl(bad, grp.base + j);
  if (q != 0) return;
  if (bad) atomicMin(&m.status[0], e);
  m.flags[e] = (unsigned char)(status | (bad ? kStatNonFinite : 0) | extra_flags);
  if (status & kStatCC3333) {
    atomicMin(&m.status[2], e);
    atomicAdd(&m.status[3], 1);
  }
}

__global__ void __launch_bounds__(kExactBlock) bf_element_kernel_exact(DevMesh m, KParams kp,
                                                               const double* __restrict__ x,
                                                               const double* __restrict__ y,
                                                               const double* __restrict__ z) {
  const int P = kp.n_qp, per_wave = 64 / P;
  QpGroup g;
  int slot;
  lane_group(P, threadIdx.x & 63, slot, g.q, g.base);
  // natural block order: the XCD-contiguous remap used by the memory-bound kernels measured
  // 16 % slower here (this kernel is ALU-bound and gains nothing from L2 locality)
  const int wave = blockIdx.x * (kExactBlock / 64) + (threadIdx.x >> 6);
  g.e = wave * per_wave + slot;
  if (slot >= per_wave || g.e >= kp.E) return;
  exact_element(m, kp, x, y, z, g, 0);
}

// Persistent drain of the hand-over list: groups stride over [0, count); count is final
// because the producing kernel completed earlier on the same stream.
__global__ void __launch_bounds__(kExactBlock) bf_element_kernel_exact_list(DevMesh m, KParams kp,
                                                                    const double* __restrict__ x,
                                                                    const double* __restrict__ y,
                                                                    const double* __restrict__ z) {
  const int count = m.status[4];
  const int P = kp.n_qp, per_wave = 64 / P;
  QpGroup g;
  int slot;
  lane_group(P, threadIdx.x & 63, slot, g.q, g.base);
  if (slot >= per_wave) return;
  const int waves = gridDim.x * (kExactBlock / 64);
  for (int i = (blockIdx.x * (kExactBlock / 64) + (threadIdx.x >> 6)) * per_wave + slot; i < count;
       i += waves * per_wave) {
    g.e = m.fallback[i];
    exact_element(m, kp, x, y, z, g, kStatFallback);
  }
}

// Plain FP64 element phase of a batch over the element tiles (one case per tile).
__global__ void __launch_bounds__(kExactBlock) bf_batch_exact_kernel(BatchDev b) {
  const int P = b.P, per_wave = 64 / P;
  QpGroup g;
  int slot;
  lane_group(P, threadIdx.x & 63, slot, g.q, g.base);
  // natural order over the work list (see bf_element_kernel_exact)
  const int wave = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  const int stride = b.persistent ? (int)(gridDim.x * (blockDim.x >> 6)) : b.n_etiles;
  for (int t = wave; t < b.n_etiles; t += stride) {
    const Tile tl = b.etiles[t];
    const int c = __builtin_amdgcn_readfirstlane(tl.c);
    const int e0 = __builtin_amdgcn_readfirstlane(tl.b);
    const int n = __builtin_amdgcn_readfirstlane(tl.n);
    const CaseDev& cd = b.cases[c];
    for (int k0 = 0; k0 < n; k0 += per_wave) {
      const int i = k0 + slot;
      if (slot < per_wave && i < n) {
        g.e = e0 + i;
        exact_element(cd.m, cd.kp, cd.x, cd.y, cd.z, g, 0);
      }
    }
  }
}

// Hand-over drain of a batch: element groups stride over b.hlist[0 .. *b.htotal).
__global__ void __launch_bounds__(kExactBlock) bf_batch_exact_list_kernel(BatchDev b) {
  const int count = *b.htotal;
  const int P = b.P, per_wave = 64 / P;
  QpGroup g;
  int slot;
  lane_group(P, threadIdx.x & 63, slot, g.q, g.base);
  if (slot >= per_wave) return;
  const int waves = gridDim.x * (kExactBlock / 64);
  for (int i = (blockIdx.x * (kExactBlock / 64) + (threadIdx.x >> 6)) * per_wave + slot; i < count;
       i += waves * per_wave) {
    const int2 v = b.hlist[i];
    const CaseDev& cd = b.cases[v.x];
    g.e = v.y;
    exact_element(cd.m, cd.kp, cd.x, cd.y, cd.z, g, kStatFallback);
  }
}

hipError_t launch_batch_exact(const BatchDev& b, int blocks, hipStream_t s) {
  if (b.n_etiles <= 0) return hipSuccess;
  if (blocks <= 0) blocks = (b.n_etiles + (kExactBlock / 64) - 1) / (kExactBlock / 64);
  hipLaunchKernelGGL(bf_batch_exact_kernel, dim3(blocks), dim3(kExactBlock), 0, s, b);
  return hipGetLastError();
}

hipError_t launch_batch_exact_list(const BatchDev& b, int blocks, hipStream_t s) {
  if (b.n_etiles <= 0) return hipSuccess;
  hipLaunchKernelGGL(bf_batch_exact_list_kernel, dim3(blocks), dim3(kExactBlock), 0, s, b);
  return hipGetLastError();
}

// Node phase of one explicit damped time step (the substitute for the reference's
// implicit step, see StepState): node n's force is gathered as in bf_gather_kernel, then
//   v <- damp * v - (dt / m_n) * f,   x <- x + dt * v      (nodes not clamped)
// in FP64 with every product rounded (no contraction), so a host restatement with the same
// operations reproduces it bit for bit.  Ghost rows of fint are zero; ghost coordinates
// follow from the rules at the next step's GlobalGhost.
__device__ __forceinline__ void step_node(const int* __restrict__ ptr,
                                          const unsigned* __restrict__ ent,
                                          const double* __restrict__ Fe, int E, int n_v,
                                          double* __restrict__ fint, double* x, double* y,
                                          double* z, double* v, const double* winv,
                                          const unsigned char* fixed, double dt, double damp,
                                          int n) {
  gather_node(ptr, ent, Fe, E, n_v, 0, fint, n);
  if (n >= n_v || fixed[n]) return;
  const double w = winv[n];
  double* vn = v + 3 * (size_t)n;
  const double vx = damp * vn[0] - w * fint[3 * (size_t)n + 0];
  const double vy = damp * vn[1] - w * fint[3 * (size_t)n + 1];
  const double vz = damp * vn[2] - w * fint[3 * (size_t)n + 2];
  vn[0] = vx;
  vn[1] = vy;
  vn[2] = vz;
  x[n] = x[n] + dt * vx;
  y[n] = y[n] + dt * vy;
  z[n] = z[n] + dt * vz;
}

__global__ void __launch_bounds__(256) bf_gather_step_kernel(const int* __restrict__ ptr,
                                                             const unsigned* __restrict__ ent,
                                                             const double* __restrict__ Fe,
                                                             int E, int n_v, int n_nodes,
                                                             double* __restrict__ fint,
                                                             double* x, double* y, double* z,
                                                             double* v, const double* winv,
                                                             const unsigned char* fixed,
                                                             double dt, double damp) {
  const int n = logical_block() * blockDim.x + threadIdx.x;
  if (n >= n_nodes) return;
  step_node(ptr, ent, Fe, E, n_v, fint, x, y, z, v, winv, fixed, dt, damp, n);
}

__global__ void __launch_bounds__(256) bf_batch_gather_step_kernel(BatchDev b) {
  const Tile tl = b.ntiles[logical_block()];
  const CaseDev& cd = b.cases[tl.c];
  const int i = threadIdx.x;
  if (i >= tl.n) return;
  const int n = tl.b + i;
  if (cd.v)
    step_node(cd.ptr, cd.ent, cd.m.Fe, cd.kp.E, cd.n_v, cd.fint, cd.x, cd.y, cd.z, cd.v,
              cd.winv, cd.fixed, cd.dt, cd.damp, n);
  else
    gather_node(cd.ptr, cd.ent, cd.m.Fe, cd.kp.E, cd.n_v, 0, cd.fint, n);
}

hipError_t launch_gather_step(const int* ptr, const unsigned* ent, const double* Fe, int E,
                              int n_v, int n_nodes, double* fint, double* x, double* y,
                              double* z, double* v, const double* winv,
                              const unsigned char* fixed, double dt, double damp, hipStream_t s) {
  if (n_nodes <= 0) return hipSuccess;
  hipLaunchKernelGGL(bf_gather_step_kernel, dim3((n_nodes + 255) / 256), dim3(256), 0, s, ptr,
                     ent, Fe, E, n_v, n_nodes, fint, x, y, z, v, winv, fixed, dt, damp);
  return hipGetLastError();
}

hipError_t launch_batch_gather_step(const BatchDev& b, hipStream_t s) {
  if (b.n_ntiles <= 0) return hipSuccess;
  hipLaunchKernelGGL(bf_batch_gather_step_kernel, dim3(b.n_ntiles), dim3(256), 0, s, b);
  return hipGetLastError();
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
  // persistent: one wave per SIMD of the 256-CU part at the kernel's occupancy of one
  // (the count is known on the device only; extra waves of a larger grid would only exit)
  const int per_block = (kExactBlock / 64) * (64 / kp.n_qp);
  const int blocks = std::min((kp.E + per_block - 1) / per_block, 512);
  hipLaunchKernelGGL(bf_element_kernel_exact_list, dim3(blocks), dim3(kExactBlock), 0, s, m, kp, x, y, z);
  return hipGetLastError();
}

hipError_t launch_element_exact(const DevMesh& m, const KParams& kp, const double* x,
                                const double* y, const double* z, hipStream_t s) {
  if (kp.E <= 0) return hipSuccess;
  const int per_block = (kExactBlock / 64) * (64 / kp.n_qp);
  hipLaunchKernelGGL(bf_element_kernel_exact, dim3((kp.E + per_block - 1) / per_block),
                     dim3(kExactBlock), 0, s, m, kp, x, y, z);
  return hipGetLastError();
}

}  // namespace bf
