// bf_force.hip -- gfx950 kernels and C ABI of the element internal-force path.
//
// Pipeline of one force evaluation (replaces FInternalAct, src/active_strain.c:2226-2267):
//   staged paths (FP64, FP32-mixed):
//   1. bf_tile_kernel<R>: one workgroup per tile of 64 spatially close elements, one lane per
//      (element, qp): the tile's node coordinates staged in LDS, the current midsurface
//      geometry once per element (FP64), the staged active-strain chain of each qp in R, the
//      qp resultants summed in qp order, the nodal contributions reduced per node in LDS in
//      ascending element order; nodes whose contributions all lie in the tile are written,
//      the others leave a tile partial.  Elements outside the well-conditioned regime are
//      listed (SegList) instead of added.
//   2. (FP32-mixed) bf_tile_tier2_kernel: the listed elements again with the FP64 staged
//      chain; what it cannot keep is listed as (element, qp) pairs.
//   3. the reference-order solve / finish / redo kernels over the list (bf_exact.hip).
//   4. bf_tile_fixup_kernel (node pass): partial nodes sum their tile partials in tile order,
//      nodes of handed-over elements add those elements' rows in ascending element order.
//   The reference-order path: bf_exact kernels for every element, then bf_gather_kernel (one
//   lane per node, the node's (element, slot) contributions in ascending element order from
//   a CSR list built at upload): the summation order of VecSetValues(ADD_VALUES) on one rank
//   (active_strain.c:2238-2260), deterministic run to run, without atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/biofem_hip.h"
#include "bf_elem.h"
#include "bf_kernels.h"
#include "bf_tiles.h"

using namespace bf;

// Scratch of bf_force_eval_device_noalloc: the per-evaluation arrays of one force evaluation
// (element-force buffer, hand-over segments and list, per-qp stresses and status, redo
// list) carved out of one file-scope device array instead of the context's allocations.
// Each GPU the library runs on holds its own copy until the process ends.  In host code
// `scratch` names a host-side stand-in: its device address comes from hipGetSymbolAddress,
// once per device (noalloc_scratch_init), and only that address is used.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// ------------------------------------------------------------------ kernels ------------
// Outcome of staged_element for one lane's qp.
enum : int { kStagedKept = 0, kStagedQpKept = 1, kStagedQpFail = 2 };

// Whether a qp of an element the FP32-mixed path hands over goes to the reference-order
// solve: its own conditioning bits or non-finite stresses, or forces of the element that are
// not finite.  The other qps keep their staged stresses, from which the reference-order
// finish forms the forces -- within the mixed path's budget (1e-5), which has orders of
// magnitude to spare.  The FP64 path hands over every qp of such an element: a passing qp
// of a handed element can carry staged errors above its rounding bound, and where the
// element's stresses are far above the mesh's typical ones (a clamped neighbouring qp), one
// such qp alone exceeds the path's 1e-12 relative budget on a whole mesh (measured: 1.5e-12
// on the 256 x 192 membrane at z noise 7e-4, from one element).
__device__ __forceinline__ bool qp_fails(int qstatus, double s0, double s1, double s2,
                                         bool elem_finite) {
  return !elem_finite ||
         needs_exact_bits(qstatus, isfinite(s0) && isfinite(s1) && isfinite(s2));
}

// Staged element body for lane q of the P-lane group that holds element e (P >= n_qp):
// every lane of the group forms the element's current midsurface geometry (FP64), lane q
// runs through-thickness point q of the staged constitutive chain in R, the group sums the
// qp resultants in qp order by cross-lane reads, and each lane writes the nodal forces of
// patch nodes q, q + n_qp, ... to the element-force buffer (kStagedKept).  An element
// outside the well-conditioned regime writes no forces: lane q stores its qp's stresses for
// the reference-order finish and answers kStagedQpKept, or answers kStagedQpFail when its
// qp must be solved in the reference order (qp_fails); the caller lists the element and the
// failing pairs.
template <class R>
__device__ __forceinline__ int staged_element(const DevMesh& m, const KParams& kp,
                                               const double* __restrict__ x,
                                               const double* __restrict__ y,
                                               const double* __restrict__ z, int e, int q,
                                               int base, double noise_max = kC33NoiseMax) {
  const int E = kp.E;
  const int n_qp = kp.n_qp;
  if (q >= n_qp) return kStagedKept;
  // Every per-element load that does not depend on another one is issued here, before the
  // first branch, so that they share one memory round trip: kind, the first 12 patch slots,
  // the geom0 cache, activation data and the area.  The coordinate gathers form the second.
  const int meta = m.meta[e];
  int pidx[12];
#pragma unroll
  for (int i = 0; i < 12; i++) pidx[i] = m.patch[(size_t)i * E + e];
  const double* rf = m.ref;
  double rv[15];
#pragma unroll
  for (int k = 0; k < 15; k++) rv[k] = rf[(size_t)k * E + e];
  const double f0 = m.fib[e], f1 = m.fib[(size_t)E + e], f2 = m.fib[(size_t)2 * E + e];
  const double gs = m.gscale ? m.gscale[e] : 1.0;
  const double dly = kp.has_delay ? m.delay[e] : 0.0;
  const double A0 = m.dA0[e];
  const int kind = meta & 3, val = (meta >> 8) & 255, nen = (meta >> 16) & 255;

  // ---- current midsurface geometry (ElemUpdateGeomSubdiv, active_strain.c:727) ----
  MidGeom G;
  const double *na0, *na1, *nb0, *nb1, *nb2;
  if (kind == kRegular) {
    na0 = c_Na[0]; na1 = c_Na[1]; nb0 = c_Nab[0]; nb1 = c_Nab[1]; nb2 = c_Nab[2];
    G = mid_geom_regular([&](int i) { return ld3(x, y, z, pidx[i]); });
  } else if (kind == kIrregular) {
    const double(*t)[kMaxNen] = c_IN[val];
    na0 = t[0]; na1 = t[1]; nb0 = t[2]; nb1 = t[3]; nb2 = t[4];
    G = mid_geom_loaded(nen, na0, na1, nb0, nb1, nb2,
                        [&](int i) { return ld3(x, y, z, m.patch[(size_t)i * E + e]); });
  } else {
    na0 = na1 = nb0 = nb1 = nb2 = nullptr;
    G = cap_geom(ld3(x, y, z, m.corner[e]), ld3(x, y, z, m.corner[(size_t)E + e]),
                 ld3(x, y, z, m.corner[(size_t)2 * E + e]));
  }
  V3<double> a31c, a32c;
  a3_alpha(G.a1, G.a2, G.Aaa, G.Aab, G.Abb, a31c, a32c);

  // ---- reference cache ----
  RefGeom rg;
  rg.a1 = mk(rv[0], rv[1], rv[2]);
  rg.a2 = mk(rv[3], rv[4], rv[5]);
  rg.a3 = mk(rv[6], rv[7], rv[8]);
  rg.a31 = mk(rv[9], rv[10], rv[11]);
  rg.a32 = mk(rv[12], rv[13], rv[14]);

  // ---- activation (ElemActDefGrad, active_strain.c:1208-1257) ----
  double g = kp.has_delay ? ramp_gamma(kp.ramp, dly) : kp.gamma_t;
  g *= gs;
  Activation act;
  act.f = mk(f0, f1, f2);
  if (kp.act_mode == 0) {
    act.d0 = act.d1 = act.d2 = 1.0;
    act.gam = g;
  } else if (kp.act_mode == 1) {
    act.d0 = 1.0 - kp.man_gamma;
    act.d1 = act.d2 = 1.0;
    act.gam = 0.0;
  } else {
    act.d0 = act.d1 = 1.0 - g;
    act.d2 = 1.0 / (act.d0 * act.d1);
    act.gam = 0.0;
  }

  // ---- this lane's qp.  A CC3333 guard failure hands the element over, so the
  //      reference's cross-qp guard rule never reaches a staged result: no coupling. ----
  auto own = [](bool fail, int) { return fail; };
  // theta of this lane's qp by selects over the wave-uniform table (a lane-indexed read of
  // the kernel arguments would be a vector memory load in the middle of the chain)
  double theta_q = kp.theta[0];
#pragma unroll
  for (int j = 1; j < kMaxQp; j++) theta_q = (q == j) ? kp.theta[j] : theta_q;
  StagedQp<R> s;
  if constexpr (sizeof(R) == 4)
    s = staged_qp_mixed(G.a1, G.a2, G.a3, a31c, a32c, rg, act, kp.h0, kp.mu, kp.K, kp.iters,
                        theta_q, own, kp.fallback != 0);
  else
    s = staged_qp<R>(G.a1, G.a2, G.a3, a31c, a32c, rg, act, kp.h0, kp.mu, kp.K, kp.iters,
                     theta_q, own, kp.fallback != 0, noise_max);
  // ---- group: resultants in qp order, status bits ----
  Resultants<R> r;
  r.N0 = r.N1 = r.N2 = r.M0 = r.M1 = r.M2 = 0.0;
  int st = 0;
#pragma unroll
  for (int j = 0; j < kMaxQp; j++) {
    if (j < n_qp) {
      StagedQp<R> o;
      o.S00 = __shfl(s.S00, base + j);
      o.S11 = __shfl(s.S11, base + j);
      o.S01 = __shfl(s.S01, base + j);
      add_resultants(r, o, kp.w[j], kp.theta[j]);
      st |= __shfl(s.status | c33_range_bit(s.c33, sizeof(R) == 4), base + j);
    }
  }
  const bool finite = isfinite(r.N0) && isfinite(r.N1) && isfinite(r.N2) && isfinite(r.M0) &&
                      isfinite(r.M1) && isfinite(r.M2);
  // outside the well-conditioned regime: the reference-order kernel recomputes it (or, on
  // the FP32-mixed path, first the FP64 staged chain of the second tier)
  const bool handed = kp.fallback && needs_exact_bits(st, finite);
  m.c33[(size_t)q * E + e] = s.c33;  // a failing qp's is replaced by the solve's
  if (handed) {
    const int qs = s.status | c33_range_bit(s.c33, sizeof(R) == 4);
    if (qp_fails(qs, (double)s.S00, (double)s.S11, (double)s.S01, finite)) return kStagedQpFail;
    double* o = m.qsv + 3 * ((size_t)q * E + e);
    o[0] = (double)s.S00;
    o[1] = (double)s.S11;
    o[2] = (double)s.S01;
    m.qst[(size_t)q * E + e] = 0;
    return kStagedQpKept;
  }

  // ---- nodal forces (ElemUpdFint, active_strain.c:1762-1995) ----
  const double pref0 = 0.5 * kp.h0 * A0;
  const double hh = 0.5 * kp.h0;
  double* Fe = m.Fe;
  if (kind != kCap) {
    for (int a = q; a < nen; a += n_qp) {
      const V3<double> f = node_force(na0[a], na1[a], nb0[a], nb1[a], nb2[a], G, pref0, hh, r.N0,
                                      r.N1, r.N2, r.M0, r.M1, r.M2);
      put_force(Fe, e, E, a, f);
    }
  } else if (q == 0) {
    // CST membrane into the patch slots holding nv1..nv3 (:1807-1851); other slots zero
    const double dN[3][2] = {{-1.0, -1.0}, {1.0, 0.0}, {0.0, 1.0}};
    V3<double> fc[3];
    int sl[3];
    for (int k = 0; k < 3; k++) {
      sl[k] = m.corner[(size_t)(3 + k) * E + e];
      fc[k] = node_force(dN[k][0], dN[k][1], 0.0, 0.0, 0.0, G, pref0, hh, r.N0, r.N1, r.N2, 0.0,
                         0.0, 0.0);
    }
    for (int a = 0; a < nen; a++) {
      V3<double> f = mk(0.0, 0.0, 0.0);
      for (int k = 0; k < 3; k++)
        if (sl[k] == a) f = add(f, fc[k]);
      put_force(Fe, e, E, a, f);
    }
  }
  // ---- status (group leader) ----
  if (q != 0) return kStagedKept;
  if (!finite) atomicMin(&m.status[0], e);
  m.flags[e] = (unsigned char)((st & (kStatCC3333 | kStatClamp)) | (finite ? 0 : kStatNonFinite));
  if (st & kStatCC3333) {
    atomicMin(&m.status[2], e);
    atomicAdd(&m.status[3], 1);
  }
  return kStagedKept;
}

constexpr int kStagedBlock = 256;

// ---- single-pass staged element phase ----
// One tile of td.T elements (bf_tiles.h) per workgroup of at least td.T * n_qp threads:
//   0. the coordinates of the tile's local nodes are staged in LDS;
//   1. thread i < T forms element i's current midsurface geometry once (FP64, patch
//      coordinates from LDS), its a3,alpha and activation, and leaves them in LDS;
//   2. thread i + q T runs through-thickness point q of element i's staged chain in R
//      (q is uniform over a wave: T is a multiple of 64) and leaves its stresses, C33 and
//      status bits in LDS;
//   3. thread i sums the resultants in qp order, decides the hand-over as the reference's
//      conditioning requires (needs_exact_bits), and folds geometry and resultants into
//      five coefficient vectors: the force on patch slot a is
//        Na0[a] U0 + Na1[a] U1 + (Nab0[a] v0 + Nab1[a] v1 + Nab2[a] v2) a3
//      (ElemUpdFint's row with its sums regrouped), which it leaves in LDS per slot;
//   4. thread l sums local node l's contributions in the tiling's order (lane order, then
//      slot order) and writes a complete node to fint, a partial node to its tile partial.
// A handed-over element contributes zeros here: it is listed in its wave's segment, marks
// its nodes as touched and sets deferred; the node pass adds its force once it is known.
// LDS (doubles, tile_smem_doubles): region A = node coordinates [3][lmax] and the element
// block [rows][T] -- rows kept to phase 3 (a1, a2, a3, the six products gc_k . A_l, pref0),
// then the chain-only inputs, which the qp results [4][P][T] replace once every chain thread
// holds its inputs -- and, once phase 3 has read all of it, the slot forces [3][NS][T];
// then the qp status bits [P][T] (ints), the tile's local slot codes [NS][T] (ushort), its
// contribution lists (cent, ushort, at most NS T entries; cptr, lmax + 1 ints), a touched
// flag per local node (bytes), and each local node's global index and output code (ints).
// Every global load that depends on the tile or element index alone is issued before the
// first barrier, and the barriers order LDS only: the global stores of a phase drain while
// the next one runs.
constexpr int kTileMaxThreads = 512;  // T * n_qp: two waves per SIMD at 256 VGPRs
enum : int { kTbA1 = 0, kTbA2 = 3, kTbA3 = 6, kTbG = 9, kTbPref = 15, kTbKeep = 16,
             kTbA31 = 16, kTbA32 = 19, kTbFib = 22, kTbD = 25, kTbGam = 28, kTbChain = 13 };
__host__ __device__ inline int tile_elem_rows(int P) {
  return kTbKeep + (4 * P > kTbChain ? 4 * P : kTbChain);
}
__host__ __device__ inline size_t tile_region_a(int T, int P, int NS, int lmax) {
  const size_t a = 3 * (size_t)lmax + (size_t)tile_elem_rows(P) * T, f = 3 * (size_t)NS * T;
  return a > f ? a : f;
}
__host__ __device__ inline size_t tile_smem_doubles(int T, int P, int NS, int lmax) {
  return tile_region_a(T, P, NS, lmax) + ((size_t)P * T + 1) / 2 + ((size_t)NS * T + 3) / 4 +
         ((size_t)NS * T + 3) / 4 + ((size_t)lmax + 2) / 2 + ((size_t)lmax + 7) / 8 + (size_t)lmax;
}

// Workgroup barrier that orders LDS only (global stores stay in flight across it).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <class R>
__device__ __forceinline__ void tile_body(const DevMesh& m, const KParams& kp, const TileDev& td,
                                          int t, const double* __restrict__ x,
                                          const double* __restrict__ y,
                                          const double* __restrict__ z,
                                          double* __restrict__ fint, int accumulate, int n_v,
                                          double* smem) {
  const int T = td.T, NS = td.NS, lm = td.lmax;
  const int NT = (int)blockDim.x;
  const int P = kp.n_qp;  // chain threads per element; threads past T * P only take barriers
  const int tid = (int)threadIdx.x;
  double *sx = smem, *sy = smem + lm, *sz = smem + 2 * lm;
  double* eb = smem + 3 * lm;                           // element block [rows][T]
  double* qb = eb + (size_t)kTbKeep * T;                // qp results [4][P][T] (after 2a)
  double* sf = smem;                                    // slot forces [3][NS][T] (phase 3b)
  const size_t ra = tile_region_a(T, P, NS, lm);
  int* qs = (int*)(smem + ra);                          // qp status [P][T]
  unsigned short* sc = (unsigned short*)(smem + ra + ((size_t)P * T + 1) / 2);  // [NS][T]
  unsigned short* ce = sc + (size_t)((NS * T + 3) / 4) * 4;  // contributions of the tile
  int* cp = (int*)(ce + (size_t)((NS * T + 3) / 4) * 4);    // [lmax + 1] ranges in ce
  unsigned char* tf = (unsigned char*)(cp + ((lm + 2) / 2) * 2);  // [lmax] touched flags
  int* sn = (int*)(tf + ((lm + 7) / 8) * 8);            // [lmax] global node
  int* so = sn + lm;                                    // [lmax] output code (TileDev out)
  const int E = kp.E;
  const int q = tid / T, i = tid - q * T;  // q: wave-uniform
  const bool elem_thread = q == 0;
  // ---- 0. loads, in the order they are needed: a wave waits for its loads in issue
  //      order, so the tile's node lists and contributions go first, then the node
  //      coordinates, then what depends on the element index (read in phases 1 and 2) ----
  const int j0 = td.node_ptr[t], L = td.node_ptr[t + 1] - j0;
  const int cb = td.cptr[j0], nc = td.cptr[j0 + L] - cb;
  unsigned short cv[kMaxNen];  // the element threads' slot codes (before the branch-free part)
  if (elem_thread) {
    const unsigned short* ls = td.lslot + (size_t)t * NS * T + i;
#pragma unroll
    for (int a = 0; a < kMaxNen; a++) cv[a] = ls[(size_t)(a < NS ? a : NS - 1) * T];
  }
  const int e = td.elem[(size_t)t * T + i];
  const bool live = e >= 0;
  const bool chain = live && q < P;
  // first pass: local nodes tid and tid + NT, contributions tid + r NT (r < 4) -- all of
  // them when n_qp >= 2 and NS <= 4 n_qp; later passes below.  The loads are branch-free
  // (indices clamped into the tile's ranges, L >= 1 and nc >= 1; dead lanes read element 0),
  // so that each wait covers exactly the loads it needs.
  const int la = tid, lb = tid + NT;
  const bool ha = la < L, hb = lb < L;
  const int ia = j0 + (ha ? la : L - 1), ib = j0 + (hb ? lb : L - 1);
  const int na = td.node[ia], nb = td.node[ib];
  const int ca = td.cptr[ia + 1], cbn = td.cptr[ib + 1];
  const int oa = td.out[ia], ob = td.out[ib];
  const double xa = x[na], ya = y[na], za = z[na];
  const double xb = x[nb], yb = y[nb], zb = z[nb];
  unsigned short cv4[4];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int k = tid + r * NT;
    cv4[r] = td.cent[cb + (k < nc ? k : nc - 1)];
  }
  const int el = live ? e : 0;
  // read in phase 1 by the element threads only, loaded by every thread (no branch)
  int meta = m.meta[el];
  const double f0 = m.fib[el], f1 = m.fib[(size_t)E + el], f2 = m.fib[(size_t)2 * E + el];
  const double gs = m.gscale ? m.gscale[el] : 1.0;
  const double dly = kp.has_delay ? m.delay[el] : 0.0;
  const double A0 = m.dA0[el];
  if (!live || !elem_thread) meta = 0;
  double rv[15];  // read in phase 2
#pragma unroll
  for (int k = 0; k < 15; k++) rv[k] = m.ref[(size_t)k * E + el];
  if (ha) {
    sx[la] = xa;
    sy[la] = ya;
    sz[la] = za;
    cp[la + 1] = ca - cb;
    tf[la] = 0;
    sn[la] = na;
    so[la] = oa;
  }
  if (hb) {
    sx[lb] = xb;
    sy[lb] = yb;
    sz[lb] = zb;
    cp[lb + 1] = cbn - cb;
    tf[lb] = 0;
    sn[lb] = nb;
    so[lb] = ob;
  }
#pragma unroll
  for (int r = 0; r < 4; r++)
    if (tid + r * NT < nc) ce[tid + r * NT] = cv4[r];
  if (elem_thread) {
#pragma unroll
    for (int a = 0; a < kMaxNen; a++)
      if (a < NS) sc[(size_t)a * T + i] = cv[a];
  }
  for (int l = tid + 2 * NT; l < L; l += NT) {
    const int n = td.node[j0 + l];
    sx[l] = x[n];
    sy[l] = y[n];
    sz[l] = z[n];
    cp[l + 1] = td.cptr[j0 + l + 1] - cb;
    tf[l] = 0;
    sn[l] = n;
    so[l] = td.out[j0 + l];
  }
  for (int k = tid + 4 * NT; k < nc; k += NT) ce[k] = td.cent[cb + k];
  int csl[3] = {-1, -1, -1};  // LV cap: the patch slots of the CST corners
  if (elem_thread && live && (meta & 3) == kCap)
    for (int k = 0; k < 3; k++) csl[k] = m.corner[(size_t)(3 + k) * E + e];
  if (tid == 0) cp[0] = 0;
  lds_barrier();
  const int kind = meta & 3, val = (meta >> 8) & 255, nen = (meta >> 16) & 255;
  // ---- 1. geometry and activation, once per element ----
  if (elem_thread && live) {
    auto xyz = [&](int s) {
      const int l = sc[(size_t)s * T + i];
      return mk(sx[l], sy[l], sz[l]);
    };
    MidGeom G;
    if (kind == kRegular) {
      G = mid_geom_regular(xyz);
    } else if (kind == kIrregular) {
      const double(*tb)[kMaxNen] = c_IN[val];
      G = mid_geom(nen, tb[0], tb[1], tb[2], tb[3], tb[4], xyz);
    } else {
      G = cap_geom(xyz(csl[0]), xyz(csl[1]), xyz(csl[2]));
    }
    V3<double> a31c, a32c;
    a3_alpha(G.a1, G.a2, G.Aaa, G.Aab, G.Abb, a31c, a32c);
    double g = kp.has_delay ? ramp_gamma(kp.ramp, dly) : kp.gamma_t;
    g *= gs;
    double d0, d1, d2, gam;
    if (kp.act_mode == 0) {
      d0 = d1 = d2 = 1.0;
      gam = g;
    } else if (kp.act_mode == 1) {
      d0 = 1.0 - kp.man_gamma;
      d1 = d2 = 1.0;
      gam = 0.0;
    } else {
      d0 = d1 = 1.0 - g;
      d2 = 1.0 / (d0 * d1);
      gam = 0.0;
    }
    const double v[kTbKeep + kTbChain] = {
        G.a1.x, G.a1.y, G.a1.z, G.a2.x, G.a2.y, G.a2.z, G.a3.x, G.a3.y, G.a3.z,
        dot(G.gc1, G.Aaa), dot(G.gc1, G.Abb), dot(G.gc1, G.Aab),
        dot(G.gc2, G.Aaa), dot(G.gc2, G.Abb), dot(G.gc2, G.Aab), 0.5 * kp.h0 * A0,
        a31c.x, a31c.y, a31c.z, a32c.x, a32c.y, a32c.z, f0, f1, f2, d0, d1, d2, gam};
#pragma unroll
    for (int k = 0; k < kTbKeep + kTbChain; k++) eb[(size_t)k * T + i] = v[k];
  }
  lds_barrier();
  // ---- 2. one (element, qp) per thread ----
  V3<double> a1c, a2c, a3c, a31c, a32c;
  Activation act;
  auto ld3 = [&](int k) { return mk(eb[(size_t)k * T + i], eb[(size_t)(k + 1) * T + i],
                                    eb[(size_t)(k + 2) * T + i]); };
  if (chain) {
    a1c = ld3(kTbA1);
    a2c = ld3(kTbA2);
    a3c = ld3(kTbA3);
    a31c = ld3(kTbA31);
    a32c = ld3(kTbA32);
    act.f = ld3(kTbFib);
    act.d0 = eb[(size_t)kTbD * T + i];
    act.d1 = eb[(size_t)(kTbD + 1) * T + i];
    act.d2 = eb[(size_t)(kTbD + 2) * T + i];
    act.gam = eb[(size_t)kTbGam * T + i];
  }
  lds_barrier();  // every chain thread holds its inputs: the qp results take their place
  double c33_keep = 0.0;  // stored after phase 4, off the path of the phases between
  if (chain) {
    RefGeom rg;
    rg.a1 = mk(rv[0], rv[1], rv[2]);
    rg.a2 = mk(rv[3], rv[4], rv[5]);
    rg.a3 = mk(rv[6], rv[7], rv[8]);
    rg.a31 = mk(rv[9], rv[10], rv[11]);
    rg.a32 = mk(rv[12], rv[13], rv[14]);
    auto own = [](bool fail, int) { return fail; };
    const double th = kp.theta[q];
    StagedQp<R> sq;
    if constexpr (sizeof(R) == 4)
      sq = staged_qp_mixed(a1c, a2c, a3c, a31c, a32c, rg, act, kp.h0, kp.mu, kp.K, kp.iters, th,
                           own, kp.fallback != 0);
    else
      sq = staged_qp<R>(a1c, a2c, a3c, a31c, a32c, rg, act, kp.h0, kp.mu, kp.K, kp.iters, th, own,
                        kp.fallback != 0, kC33NoiseMax);
    double* o = qb + (size_t)q * T + i;
    o[0] = (double)sq.S00;
    o[(size_t)P * T] = (double)sq.S11;
    o[(size_t)2 * P * T] = (double)sq.S01;
    o[(size_t)3 * P * T] = sq.c33;
    qs[q * T + i] = sq.status | c33_range_bit(sq.c33, sizeof(R) == 4);
    c33_keep = sq.c33;
  }
  lds_barrier();
  // ---- 3. resultants, hand-over, force coefficients ----
  bool handed = false;
  int st_keep = 0;           // the element's status bits and finiteness, stored after
  bool finite_keep = true;   // phase 4 with the hand-over mark
  unsigned qfail = 0u;  // qps of a handed-over element that the reference order solves
  V3<double> U0 = mk(0.0, 0.0, 0.0), U1 = U0, A3 = U0;
  double v0 = 0.0, v1 = 0.0, v2 = 0.0;
  if (elem_thread && live) {
    Resultants<R> r;
    r.N0 = r.N1 = r.N2 = r.M0 = r.M1 = r.M2 = 0.0;
    int st = 0;
    for (int k = 0; k < P; k++) {
      const double* o = qb + (size_t)k * T + i;
      StagedQp<double> s;
      s.S00 = o[0];
      s.S11 = o[(size_t)P * T];
      s.S01 = o[(size_t)2 * P * T];
      add_resultants(r, s, kp.w[k], kp.theta[k]);
      st |= qs[k * T + i];
    }
    const bool finite = isfinite(r.N0) && isfinite(r.N1) && isfinite(r.N2) &&
                        isfinite(r.M0) && isfinite(r.M1) && isfinite(r.M2);
    handed = kp.fallback && needs_exact_bits(st, finite);
    st_keep = st;
    finite_keep = finite;
    if (handed) {  // its nodes are marked in LDS; phase 4 writes the marks
      unsigned short code[kMaxNen];  // read branch-free (rows clamped), then masked
#pragma unroll
      for (int a = 0; a < kMaxNen; a++) code[a] = sc[(size_t)(a < NS ? a : NS - 1) * T + i];
#pragma unroll
      for (int a = 0; a < kMaxNen; a++)
        if (a < nen && code[a] != kTileIdle) tf[code[a]] = 1;  // kTileIdle: a slot without force (LV cap)
      // FP64: every qp goes to the reference-order solve (no pairs are listed: the solve
      // takes the element list).  FP32-mixed (the second tier re-evaluates the element in
      // FP64 first): the failing qps are listed, the others keep their staged stresses, from
      // which the reference-order finish forms the forces.
      if (sizeof(R) == 4) for (int k = 0; k < P; k++) {
        const double* o = qb + (size_t)k * T + i;
        const double s0 = o[0], s1 = o[(size_t)P * T], s2 = o[(size_t)2 * P * T];
        if (qp_fails(qs[k * T + i], s0, s1, s2, finite)) {
          qfail |= 1u << k;
        } else {
          double* d = m.qsv + 3 * ((size_t)k * E + e);
          d[0] = s0;
          d[1] = s1;
          d[2] = s2;
          m.qst[(size_t)k * E + e] = 0;
        }
      }
    }
    // the force coefficients, formed by every lane of the wave alike (no divergent branch)
    // and zero for a handed-over element, whose force comes later; the kept rows: the qp
    // results took the chain-input rows only
    const V3<double> a1 = ld3(kTbA1), a2 = ld3(kTbA2);
    A3 = ld3(kTbA3);
    const double* gg = eb + (size_t)kTbG * T + i;
    const double pref0 = eb[(size_t)kTbPref * T + i];
    const double hh = 0.5 * kp.h0;
    // LV cap: CST membrane only (:1807-1851)
    const double M0 = kind == kCap ? 0.0 : r.M0, M1 = kind == kCap ? 0.0 : r.M1,
                 M2 = kind == kCap ? 0.0 : r.M2;
    const double b0 = gg[0] * M0 + gg[(size_t)T] * M1 + 2.0 * gg[(size_t)2 * T] * M2;
    const double b1 = gg[(size_t)3 * T] * M0 + gg[(size_t)4 * T] * M1 +
                      2.0 * gg[(size_t)5 * T] * M2;
    U0 = scale(pref0, add(add(scale(r.N0, a1), scale(r.N2, a2)), scale(-hh * b0, A3)));
    U1 = scale(pref0, add(add(scale(r.N2, a1), scale(r.N1, a2)), scale(-hh * b1, A3)));
    v0 = -pref0 * hh * M0;
    v1 = -pref0 * hh * M1;
    v2 = -2.0 * pref0 * hh * M2;
    if (handed) {
      U0 = U1 = A3 = mk(0.0, 0.0, 0.0);
      v0 = v1 = v2 = 0.0;
    }
  }
  // FP64, and the single-context FP32-mixed path: straight into a list (the hand-over
  // list; the second tier's input), no compaction pass; the list atomic is issued here and
  // its offset used after phase 4.  Batches on the FP32-mixed path: per-wave segments.
  const bool direct = sizeof(R) == 8 || td.single;
  int* hcount = td.list_count;
  int* hlist = td.list_e;
  if (sizeof(R) == 4 && td.single) {  // this tile's region of the list (SegList)
    const int j = t / td.t1_tiles;
    hcount += kCntStride * j;
    hlist += (size_t)j * td.t1_tiles * T;
  }
  unsigned long long hbal = 0ull;
  int hbase = 0;
  if (elem_thread) {
    if (direct) {
      hbal = __ballot(live && handed);
      hbase = hand_list_reserve(hcount, hbal);
    } else {  // per-wave segments, which the second tier re-evaluates
      const int w = t * (T >> 6) + (tid >> 6);
      HandSink{td.wlist, td.wqlist, td.wcount, w, td.any}.push(live && handed, qfail, e, P);
    }
  }
  lds_barrier();  // region A is read: the slot forces take its place
  if (elem_thread) {
    const bool adds = live && !handed;
    auto put = [&](int a, const V3<double>& f) {
      sf[(size_t)a * T + i] = f.x;
      sf[(size_t)(NS + a) * T + i] = f.y;
      sf[(size_t)(2 * NS + a) * T + i] = f.z;
    };
    // a handed-over or empty lane takes its kind's branch with zero coefficients (its slot
    // forces are signed zeros, which leave every node sum as it is)
    if (kind == kRegular) {  // the Loop tables as immediates
#pragma unroll
      for (int a = 0; a < 12; a++) {
        const double s = kLoopNab[0][a] * v0 + kLoopNab[1][a] * v1 + kLoopNab[2][a] * v2;
        put(a, add(add(scale(kLoopNa[0][a], U0), scale(kLoopNa[1][a], U1)), scale(s, A3)));
      }
      for (int a = 12; a < NS; a++) put(a, mk(0.0, 0.0, 0.0));
    } else if (kind == kIrregular) {
      const double(*tb)[kMaxNen] = c_IN[val];
      for (int a = 0; a < NS; a++) {
        V3<double> f = mk(0.0, 0.0, 0.0);
        if (a < nen) {
          const double s = tb[2][a] * v0 + tb[3][a] * v1 + tb[4][a] * v2;
          f = add(add(scale(tb[0][a], U0), scale(tb[1][a], U1)), scale(s, A3));
        }
        put(a, f);
      }
    } else {
      for (int a = 0; a < NS; a++) {
        V3<double> f = mk(0.0, 0.0, 0.0);
        if (adds) {  // LV cap, corner k of the CST: dN = (-1, -1), (1, 0), (0, 1)
          if (csl[0] == a) f = add(f, scale(-1.0, add(U0, U1)));
          if (csl[1] == a) f = add(f, U0);
          if (csl[2] == a) f = add(f, U1);
        }
        put(a, f);
      }
    }
  }
  lds_barrier();
  // ---- 4. node sums in the tiling's order ----
  for (int l = tid; l < L; l += NT) {
    const int c0 = cp[l], c1 = cp[l + 1];
    double ax = 0.0, ay = 0.0, az = 0.0;
    // contributions in chunks: the chunk's entries, then their slot forces, each read
    // together; added in entry order
    constexpr int kCh = 8;
    for (int cc = c0; cc < c1; cc += kCh) {
      unsigned v[kCh];
#pragma unroll
      for (int r = 0; r < kCh; r++) v[r] = cc + r < c1 ? ce[cc + r] : 0u;
      double fx[kCh], fy[kCh], fz[kCh];
#pragma unroll
      for (int r = 0; r < kCh; r++) {
        fx[r] = fy[r] = fz[r] = 0.0;
        if (cc + r < c1) {
          const size_t k = (size_t)(v[r] >> 8) * T + (v[r] & 0xFFu);
          fx[r] = sf[k];
          fy[r] = sf[(size_t)NS * T + k];
          fz[r] = sf[(size_t)2 * NS * T + k];
        }
      }
#pragma unroll
      for (int r = 0; r < kCh; r++)
        if (cc + r < c1) {
          ax += fx[r];
          ay += fy[r];
          az += fz[r];
        }
    }
    const int o = so[l];
    const int n = sn[l];
    if (tf[l] && n < n_v) td.touched[n] = 1;
    if (o == -1) {
      double* fo = fint + 3 * (size_t)n;
      const double bx = accumulate ? fo[0] : 0.0, by = accumulate ? fo[1] : 0.0,
                   bz = accumulate ? fo[2] : 0.0;
      fo[0] = bx + ax;
      fo[1] = by + ay;
      fo[2] = bz + az;
    } else if (o >= 0) {
      double* pr = td.partial + 3 * (size_t)o;
      pr[0] = ax;
      pr[1] = ay;
      pr[2] = az;
    }
  }
  if (direct && elem_thread) hand_list_write(td, hlist, hbal, hbase, live && handed, e);
  if (elem_thread && live) {
    td.deferred[e] = handed ? 1 : 0;
    if (!handed) {
      if (!finite_keep) atomicMin(&m.status[0], e);
      m.flags[e] = (unsigned char)((st_keep & (kStatCC3333 | kStatClamp)) |
                                   (finite_keep ? 0 : kStatNonFinite));
      if (st_keep & kStatCC3333) {
        atomicMin(&m.status[2], e);
        atomicAdd(&m.status[3], 1);
      }
    }
  }
  // handed-over elements get their C33 from the reference-order solve later
  if (chain) m.c33[(size_t)q * E + e] = c33_keep;
}

template <class R>
__global__ void __launch_bounds__(kTileMaxThreads) __attribute__((amdgpu_waves_per_eu(3)))
bf_tile_kernel(DevMesh m, KParams kp, TileDev td, const double* __restrict__ x,
               const double* __restrict__ y, const double* __restrict__ z,
               double* __restrict__ fint, int accumulate, int n_v) {
  extern __shared__ double tile_smem[];
  const int t = logical_block();
  if (t >= td.n_tiles) return;
  tile_body<R>(m, kp, td, t, x, y, z, fint, accumulate, n_v, tile_smem);
}

// Appends the pair (e, q) of every lane with `fail` to its qp's segment of wave w (see
// HandSink); pushed[k]: wave-uniform counts of what earlier calls appended per qp.  Every
// lane active here was active in the earlier calls of the wave.
__device__ __forceinline__ unsigned long long hand_counts(int n_elem,
                                                         const int (&pairs)[kMaxQp]) {
  unsigned long long w = hand_count(n_elem);
#pragma unroll
  for (int k = 0; k < kMaxQp; k++) w |= hand_count_pairs(k, pairs[k]);
  return w;
}

__device__ __forceinline__ void push_failing_pair(int* list, int w, bool fail, int e, int q,
                                                  int P, int (&pushed)[kMaxQp]) {
  const int lane = (int)(threadIdx.x & 63u);
  const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
  for (int k = 0; k < kMaxQp; k++) {
    if (k >= P) break;
    const bool take = fail && q == k;
    const unsigned long long bal = __ballot(take);
    if (take) list[pair_segment(w, k) + pushed[k] + __popcll(bal & below)] = qp_pair(e, q);
    pushed[k] += __popcll(bal);
  }
}

// Tiles per region of the tile kernel's list, and its entries in all.
inline int t1_region_tiles(int n_tiles) { return std::max(1, (n_tiles + kRegions - 1) / kRegions); }
inline size_t t1_list_entries(size_t n_tiles, size_t T) {
  return (size_t)kRegions * t1_region_tiles((int)n_tiles) * T;
}
// The second tier's grid: two waves per tile wave, at most 32 elements per wave on average
// over a list as long as the mesh, so a wave takes at most ceil(32 / per_wave) passes of
// per_wave = floor(64 / n_qp) elements: 64 or fewer (n_qp <= 5).  Waves per region of its
// element list (64 entries per wave), and that list's entries in all.
inline int tier2_waves(int n_waves) { return 2 * std::max(n_waves, 1); }
inline int t2_region_waves(int n_waves) { return (tier2_waves(n_waves) + kRegions - 1) / kRegions; }
inline size_t t2_list_entries(int n_waves) { return (size_t)kRegions * t2_region_waves(n_waves) * 64; }

// Second tier of the single-context FP32-mixed path: the elements the tile kernel handed
// over (its SegList, TileDev::t1_tiles) are re-evaluated with the FP64 staged chain,
// floor(64 / n_qp) lane groups per pass, and what that chain cannot hold within the mixed
// budget either is listed again (the FP32 chain keeps the well-conditioned elements:
// rounding bound 1e-11, C33 in [0.4, 2.5]; the FP64 chain the moderately conditioned ones:
// bound 1e-8, C33 in [0.1, 10]): the element in region k / region_waves of the second
// tier's element SegList (t2_list), its failing (element, qp) pairs in the class of their
// qp rank of m.qlist (E entries per class), at offsets from one atomic per wave and list.
// Wave k takes passes k, k + gridDim.x, ..: with tier2_waves no wave takes more than 64
// elements, and a short list is one pass's latency.  The forces of the elements it keeps go
// to the element-force buffer (they stay deferred: the node pass adds them).
__global__ void __launch_bounds__(64) bf_tile_tier2_kernel(DevMesh m, KParams kp, TileDev td,
                                                           int* __restrict__ t2_list,
                                                           int region_waves,
                                                           const double* __restrict__ x,
                                                           const double* __restrict__ y,
                                                           const double* __restrict__ z) {
  const int P = kp.n_qp, per_wave = 64 / P, E = kp.E;
  int slot, q, base;
  lane_group(P, (int)(threadIdx.x & 63u), slot, q, base);
  const int k = (int)blockIdx.x;
  const SegList in{td.list_e, td.list_count, td.t1_tiles * td.T, kRegions};
  const SegCount sc = seg_count(in);
  const int j2 = k / region_waves;
  int* ecount = m.status + kT2Base + kCntStride * j2;
  int* elist = t2_list + (size_t)j2 * region_waves * 64;
  for (int p = k; p * per_wave < sc.total; p += (int)gridDim.x) {
    const int i = p * per_wave + slot;
    const bool act = slot < per_wave && i < sc.total;
    int r = kStagedKept, e = 0;
    if (act) {
      e = seg_at(in, sc, i);
      r = staged_element<double>(m, kp, x, y, z, e, q, base, kC33NoiseMaxTier2);
    }
    // the atomics of every list are issued before any offset is used
    const bool take_e = act && q == 0 && r != kStagedKept;
    const unsigned long long be = __ballot(take_e);
    const int eb = hand_list_reserve(ecount, be);
    unsigned long long bq[kMaxQp];
    int qb[kMaxQp];
#pragma unroll
    for (int rk = 0; rk < kMaxQp; rk++) {
      const bool tq = rk < P && act && r == kStagedQpFail && q == qp_of_rank(rk, P);
      bq[rk] = __ballot(tq);
      qb[rk] = hand_list_reserve(m.status + kClassBase + kCntStride * rk, bq[rk]);
    }
    hand_list_write(td, elist, be, eb, take_e, e);
#pragma unroll
    for (int rk = 0; rk < kMaxQp; rk++)
      hand_list_write(td, m.qlist + (size_t)rk * E, bq[rk], qb[rk],
                      ((bq[rk] >> (threadIdx.x & 63u)) & 1ull) != 0ull, qp_pair(e, q));
  }
}

// Batches on the FP32-mixed path: concatenates the tile waves' hand-over segments (in wave
// order) into one element list and their pair segments into one pair list, qp by qp in
// qp_of_rank order, and stores the totals and the mid-surface pairs (the last class) at the
// end (set_totals).  One thread per wave, blocks of kCompactBlock waves: every block
// reads all the packed count words (loads in flight together) and forms the class totals
// and the sums of the waves before its own, then a block-wide scan gives each wave its
// offsets, and each thread copies its wave's segments with the loads of every class in
// flight together.  The segments were just written by waves on every XCD, so each dependent
// load is a round trip to memory; the copy takes one, whatever the number of segments.
constexpr int kCompactBlock = 256;
constexpr int kCompactWaves = kCompactBlock / 64;
constexpr int kCompactClasses = 1 + kMaxQp;  // elements, then pairs by qp rank
constexpr int kCompactChunk = 8;             // entries per class loaded together

// Count of class c (0: elements, 1 + r: pairs of qp rank r) in a packed word.
__device__ __forceinline__ int class_count(unsigned long long w, int c, int P) {
  return c == 0 ? count_elems(w) : count_pairs(w, qp_of_rank(c - 1, P));
}

// Block-wide sums (red) and exclusive scan (scan) of v[c], c < kCompactClasses;
// s_wave: [2][kCompactWaves][kCompactClasses].
__device__ __forceinline__ void compact_block_sums(int* s_wave, const int (&v)[kCompactClasses],
                                                   const int (&u)[kCompactClasses],
                                                   int (&scan)[kCompactClasses],
                                                   int (&red)[kCompactClasses]) {
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  int inc[kCompactClasses], tot[kCompactClasses];
#pragma unroll
  for (int c = 0; c < kCompactClasses; c++) {
    inc[c] = v[c];
    tot[c] = u[c];
  }
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
#pragma unroll
    for (int c = 0; c < kCompactClasses; c++) {
      const int x = __shfl_up(inc[c], d);
      if (lane >= d) inc[c] += x;
      tot[c] += __shfl_xor(tot[c], d);
    }
  }
  if (lane == 63) {
#pragma unroll
    for (int c = 0; c < kCompactClasses; c++) {
      s_wave[wv * kCompactClasses + c] = inc[c];
      s_wave[(kCompactWaves + wv) * kCompactClasses + c] = tot[c];
    }
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < kCompactClasses; c++) {
    int before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < kCompactWaves; w++) {
      before += w < wv ? s_wave[w * kCompactClasses + c] : 0;
      all += s_wave[(kCompactWaves + w) * kCompactClasses + c];
    }
    scan[c] = before + inc[c] - v[c];
    red[c] = all;
  }
}

// The element segments (class 0, 64 entries per wave, elist) and the pair segments of the P
// qps in qp_of_rank order (classes 1 .. P, plist, pair_segment; none for P = 0) of waves
// [0, n_waves) with packed counts cnt: each class concatenated in wave order, the pair
// classes one after another.  emit_elem(at, w, v), emit_pair(at, w, v); totals to
// set_totals(elements, pairs, mid-surface pairs) by one thread.  Launched with
// ceil(n_waves / kCompactBlock) blocks (at least one).
template <class EmitE, class EmitP, class Totals>
__device__ __forceinline__ void compact_all(int* s_wave, const unsigned long long* cnt,
                                            const int* elist, const int* plist, int n_waves,
                                            int P, EmitE&& emit_elem, EmitP&& emit_pair,
                                            Totals&& set_totals) {
  // P = 0: elements only
  const int t = threadIdx.x;
  const int w0 = blockIdx.x * kCompactBlock, w = w0 + t;
  const int nc = 1 + P;
  const unsigned long long mine = w < n_waves ? cnt[w] : 0ull;
  // class sums of every wave (all) and of the waves before this block (before)
  int before[kCompactClasses] = {}, all[kCompactClasses] = {};
  constexpr int kR = 16;  // count words in flight per thread
  for (int g0 = t; g0 < n_waves; g0 += kR * kCompactBlock) {
    unsigned long long wd[kR];
#pragma unroll
    for (int r = 0; r < kR; r++) {
      const int g = g0 + r * kCompactBlock;
      wd[r] = cnt[g < n_waves ? g : n_waves - 1];
      if (g >= n_waves) wd[r] = 0ull;
    }
#pragma unroll
    for (int r = 0; r < kR; r++) {
      const bool pre = g0 + r * kCompactBlock < w0;
#pragma unroll
      for (int c = 0; c < kCompactClasses; c++) {
        const int k = c < nc ? class_count(wd[r], c, P) : 0;
        all[c] += k;
        before[c] += pre ? k : 0;
      }
    }
  }
  int own[kCompactClasses];
#pragma unroll
  for (int c = 0; c < kCompactClasses; c++) own[c] = c < nc ? class_count(mine, c, P) : 0;
  int scan[kCompactClasses], sb[kCompactClasses], total[kCompactClasses];
  compact_block_sums(s_wave, own, all, scan, total);
  __syncthreads();
  compact_block_sums(s_wave, before, before, sb, sb);  // sb: the block's sum of `before`
  int base[kCompactClasses];  // start of each class in its output list
  base[0] = 0;
  int acc = 0;
#pragma unroll
  for (int c = 1; c < kCompactClasses; c++) {
    base[c] = acc;
    acc += c < nc ? total[c] : 0;
  }
  if (mine != 0ull) {  // most waves hold no hand-over
    int v[kCompactClasses][kCompactChunk];
#pragma unroll
    for (int c = 0; c < kCompactClasses; c++) {
      const int n = own[c];
      const int* src = c == 0 ? elist + (size_t)w * 64
                              : plist + pair_segment(w, c < nc ? qp_of_rank(c - 1, P) : 0);
#pragma unroll
      for (int k = 0; k < kCompactChunk; k++) v[c][k] = src[k < n ? k : (n > 0 ? n - 1 : 0)];
    }
#pragma unroll
    for (int c = 0; c < kCompactClasses; c++) {
      const int n = own[c];
      const int at = base[c] + sb[c] + scan[c];
#pragma unroll
      for (int k = 0; k < kCompactChunk; k++) {
        if (k < n) {
          if (c == 0) emit_elem(at + k, w, v[c][k]);
          else emit_pair(at + k, w, v[c][k]);
        }
      }
      if (n > kCompactChunk) {  // rare: a long segment
        const int* src = c == 0 ? elist + (size_t)w * 64
                                : plist + pair_segment(w, qp_of_rank(c - 1, P));
        for (int k = kCompactChunk; k < n; k++) {
          if (c == 0) emit_elem(at + k, w, src[k]);
          else emit_pair(at + k, w, src[k]);
        }
      }
    }
  }
  if (blockIdx.x == 0 && t == 0) set_totals(total[0], acc, has_mid_qp(P) ? total[P] : 0);
}

__global__ void __launch_bounds__(256) bf_gather_kernel(const int* __restrict__ ptr,
                                                        const unsigned* __restrict__ ent,
                                                        const double* __restrict__ Fe, int E,
                                                        int n_v, int n_nodes, int accumulate,
                                                        double* __restrict__ fint) {
  const int n = logical_block() * blockDim.x + threadIdx.x;
  if (n >= n_nodes) return;
  gather_node(ptr, ent, Fe, E, n_v, accumulate, fint, n);
}

// ------------------------------------------------------------------ batch kernels ------
// Prologue of a batched step: per-case status reset, then (ghosts != 0) GlobalGhost over
// the ghost tiles, one block per tile stride.
__global__ void __launch_bounds__(256) bf_batch_prologue(BatchDev b, int ghosts) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < b.n_cases; c += gridDim.x * blockDim.x) {
    int* st = b.cases[c].m.status;
    st[0] = INT_MAX;
    st[1] = 0;
    st[2] = INT_MAX;
    st[3] = 0;
    st[4] = 0;
    st[5] = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    b.rtotal[0] = 0;
    b.rtotal[1] = 0;
    *b.htotal = 0;  // the FP64 tile kernel appends to hlist
    b.ptotal[2] = 0;  // TileDev::any
  }
  if (!ghosts) return;
  for (int t = blockIdx.x; t < b.n_gtiles; t += gridDim.x) {
    const Tile tl = b.gtiles[t];
    const CaseDev& cd = b.cases[tl.c];
    for (int i = threadIdx.x; i < tl.n; i += blockDim.x) {
      const int g = tl.b + i;
      const int rb = cd.rules[3 * g], ra = cd.rules[3 * g + 1], rc = cd.rules[3 * g + 2];
      cd.x[cd.n_v + g] = cd.x[rb] + cd.x[ra] - cd.x[rc];
      cd.y[cd.n_v + g] = cd.y[rb] + cd.y[ra] - cd.y[rc];
      cd.z[cd.n_v + g] = cd.z[rb] + cd.z[ra] - cd.z[rc];
    }
  }
}

// Single-pass staged element phase of a batch: one tile (case, tile) of the work list per
// workgroup.  (A persistent loop over the list held ~70 more VGPRs live across tiles: two
// waves per SIMD instead of three.)
template <class R>
__global__ void __launch_bounds__(kTileMaxThreads) __attribute__((amdgpu_waves_per_eu(3)))
bf_batch_tile_kernel(BatchDev b) {
  extern __shared__ double tile_smem[];
  const int i = logical_block();
  if (i >= b.n_ttiles) return;
  const Tile tl = b.ttiles[i];
  const int c = __builtin_amdgcn_readfirstlane(tl.c);
  const int t = __builtin_amdgcn_readfirstlane(tl.b);
  const CaseDev& cd = b.cases[c];
  tile_body<R>(cd.m, cd.kp, cd.td, t, cd.x, cd.y, cd.z, cd.fint, 0, cd.n_v, tile_smem);
}

// Second tier of the FP32-mixed batch path (as bf_tier2_kernel): the wave of tile t
// re-evaluates the tile's hand-over segment with the FP64 staged chain in passes and
// rewrites the segment in place with what remains for the reference-order kernel (a pass
// reads its entries before it appends, and appends never pass the read position).
__global__ void __launch_bounds__(kStagedBlock) bf_batch_tier2_kernel(BatchDev b) {
  if (b.ptotal[2] == 0) return;  // no tile wave handed anything over
  const int P = b.P, per_wave = 64 / P;
  int slot, q, base;
  lane_group(P, threadIdx.x & 63, slot, q, base);
  const int step = (int)gridDim.x * (kStagedBlock / 64);
  const int t0 = __builtin_amdgcn_readfirstlane(logical_block() * (kStagedBlock / 64) +
                                                (int)(threadIdx.x >> 6));
  for (int t = t0; t < b.n_etiles; t += step) {
    const int n = count_elems(b.tcount[t]);
    if (n == 0) continue;
    const Tile tl = b.etiles[t];
    const int c = __builtin_amdgcn_readfirstlane(tl.c);
    const CaseDev& cd = b.cases[c];
    int* seg = b.tlist + (size_t)t * b.tile_cap;
    int pushed = 0;
    int qpushed[kMaxQp] = {};
    for (int k0 = 0; k0 < n; k0 += per_wave) {
      const int i = k0 + slot;
      if (slot < per_wave && i < n) {
        const int e = seg[i];
        const int r = staged_element<double>(cd.m, cd.kp, cd.x, cd.y, cd.z, e, q, base,
                                             kC33NoiseMaxTier2);
        TileSink{seg, &pushed}.push(r != kStagedKept && q == 0, e);
        push_failing_pair(b.tqlist, t, r == kStagedQpFail, e, q, P, qpushed);
      }
    }
    if ((threadIdx.x & 63) == 0) b.tcount[t] = hand_counts(pushed, qpushed);
  }
}

// Concatenates the tiles' hand-over segments into (case, element) pairs, *htotal = count,
// and (pairs: the FP32-mixed path) their pair segments into (case, qp_pair) entries by qp
// rank, ptotal[0] = count, ptotal[1] = mid-surface pairs at the end.
__global__ void __launch_bounds__(kCompactBlock) bf_batch_compact_kernel(BatchDev b, int pairs) {
  __shared__ int s_wave[2 * kCompactWaves * kCompactClasses];
  if (b.ptotal[2] == 0) {  // no tile wave handed anything over: empty lists
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      *b.htotal = 0;
      b.ptotal[0] = 0;
      b.ptotal[1] = 0;
    }
    return;
  }
  compact_all(s_wave, b.tcount, b.tlist, b.tqlist, b.n_etiles, pairs ? b.P : 0,
              [&](int at, int w, int v) { b.hlist[at] = make_int2(b.etiles[w].c, v); },
              [&](int at, int w, int v) { b.plist[at] = make_int2(b.etiles[w].c, v); },
              [&](int ne, int np, int nm) {
                *b.htotal = ne;
                b.ptotal[0] = np;
                b.ptotal[1] = nm;
              });
}

// Node gather of a batch: one block per node tile.
__global__ void __launch_bounds__(256) bf_batch_gather_kernel(BatchDev b) {
  const Tile tl = b.ntiles[logical_block()];
  const CaseDev& cd = b.cases[tl.c];
  const int i = threadIdx.x;
  if (i >= tl.n) return;
  gather_node(cd.ptr, cd.ent, cd.m.Fe, cd.kp.E, cd.n_v, 0, cd.fint, tl.b + i);
}

// GlobalGhost (src/bending.c:1763-1814): ghost n_v+g = x[b] + x[a] - x[c] per rule (b,a,c).
__global__ void __launch_bounds__(256) bf_ghost_kernel(const int* __restrict__ rules, int G,
                                                       int n_v, double* __restrict__ x,
                                                       double* __restrict__ y,
                                                       double* __restrict__ z) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  const int b = rules[3 * g], a = rules[3 * g + 1], c = rules[3 * g + 2];
  x[n_v + g] = x[b] + x[a] - x[c];
  y[n_v + g] = y[b] + y[a] - y[c];
  z[n_v + g] = z[b] + z[a] - z[c];
}

// Status words of every case of a batch: first non-finite element and first CC3333 trip.
__global__ void __launch_bounds__(256) bf_batch_case_status_kernel(const CaseDev* __restrict__ cases,
                                                                   int n, int* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int* st = cases[c].m.status;
  out[2 * c] = st[0];
  out[2 * c + 1] = st[2];
}

__global__ void __launch_bounds__(256) bf_fill_u8(unsigned char* p, int n, unsigned char v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void bf_status_reset(int* st) {
  if (threadIdx.x == 0) {
    st[0] = INT_MAX;
    st[1] = 0;
    st[2] = INT_MAX;
    st[3] = 0;
    st[4] = 0;
    st[5] = 0;
    st[6] = 0;
    st[7] = 0;
    st[8] = 0;
    st[9] = 0;
  }
  // the SegList counters: kRegions tile regions, kRegions second-tier regions, kMaxQp classes
  if (threadIdx.x < 2 * kRegions + kMaxQp) st[kT1Base + kCntStride * threadIdx.x] = 0;
}

// ------------------------------------------------------------------ host side ----------
struct bf_ctx {
  int device = 0;
  int n_cu = 256;  // compute units of the device (persistent grid sizes)
  hipStream_t stream = nullptr;
  // per evaluation: before the element kernel, after it, after the hand-over kernels,
  // after the gather; before and after the hand-over solve kernels
  static constexpr int kEv = 6;
  hipEvent_t ev[kEv] = {};
  bool timing = false;
  hipEvent_t own_ev[kEv] = {};
  // timing ring: kEv events per evaluation, harvested when full or on bf_timing_read
  static constexpr int kRing = 64;
  std::vector<hipEvent_t> ring;
  int ring_n = 0;
  bool ring_on = false;
  bool timed_last = false;    // the last evaluation recorded every timing event
  bool ev3_recorded = false;  // ... or at least ev[3] (timing, or a caller's stream)
  long long t_count = 0;
  double t_elem = 0.0, t_handover = 0.0, t_gather = 0.0, t_solve = 0.0;
  // mesh
  int n_v = 0, n_nodes = 0, E = 0, NS = 12, n_elmt_base = 0, nnz = 0, n_ghost = 0;
  bool has_mesh = false, has_params = false, has_delay = false, has_cap = false;
  int* d_patch = nullptr;
  int* d_meta = nullptr;
  int* d_corner = nullptr;
  double* d_ref = nullptr;
  double* d_dA0 = nullptr;
  double* d_fib = nullptr;
  double* d_gscale = nullptr;
  double* d_delay = nullptr;
  double* d_Fe = nullptr;
  double* d_c33 = nullptr;
  unsigned char* d_flags = nullptr;
  int* d_fallback = nullptr;
  int* d_qlist = nullptr;
  SegList last_list{};  // hand-over list of the last staged evaluation (own or
                              // scratch): pairs (FP32-mixed) or elements (FP64)
  double* d_qsv = nullptr;
  int* d_qst = nullptr;
  int* d_redo = nullptr;
  // explicit time-stepping state (bf_state_init)
  double* d_v = nullptr;
  double* d_winv = nullptr;
  unsigned char* d_fixed = nullptr;
  double dt = 0.0, damp = 0.0;
  bool has_state = false;
  // Launch of the hand-over solve per staged precision (index BF_PREC_*): whether the list
  // length last read back for it fitted the wide split (a short list), so the wide solve
  // kernel runs beside the narrow one.  False until a status read follows an evaluation of
  // that precision; the choice changes the time, never the result.
  bool wide_hint[3] = {false, false, false};
  int fb_prec = -1;  // precision of the last evaluation if it ran the hand-over, else -1
  int* d_ptr = nullptr;
  unsigned* d_ent = nullptr;
  int* d_rules = nullptr;
  int* d_status = nullptr;    // [kStatusAlloc] the step's status words and list counters
  int* d_status_vis = nullptr;  // [kStatusAlloc] their copy after a staged step (host reads)
  int* st_read = nullptr;       // where the host reads the last step's words
  bool st_fresh = false;        // d_status is reset for the next step (status_handoff)
  int* h_status = nullptr;  // pinned
  double *d_x = nullptr, *d_y = nullptr, *d_z = nullptr, *d_fint = nullptr;
  // params
  bf_params p{};
  KParams kp{};
  // single-pass staged tiling (bf_tiles.h): host copy of what it is built from, the device
  // arrays and the per-evaluation arrays
  std::vector<int> h_patch;  // [NS][E] as uploaded (parked slots)
  std::vector<int> h_nen;    // [E]
  std::vector<double> h_x0, h_y0, h_z0;
  int tileT = 0;
  int n_tiles = 0, n_partials = 0, tile_lmax = 0;
  int* d_telem = nullptr;
  unsigned short* d_tlslot = nullptr;
  int* d_tcptr = nullptr;
  unsigned short* d_tcent = nullptr;
  int* d_tnode_ptr = nullptr;
  int* d_tnode = nullptr;
  int* d_tout = nullptr;
  int* d_tpptr = nullptr;
  unsigned char* d_tcomplete = nullptr;
  double* d_tpartial = nullptr;
  unsigned char* d_touched = nullptr;
  unsigned char* d_deferred = nullptr;
  int* d_twlist = nullptr;
};

namespace {

int g_tables_device = -1;

#define BF_HIP(call)                                                                      \
  do {                                                                                   \
    hipError_t _e = (call);                                                              \
    if (_e != hipSuccess) {                                                              \
      fprintf(stderr, "biofem_hip: %s failed: %s (%s:%d)\n", #call, hipGetErrorString(_e), \
              __FILE__, __LINE__);                                                       \
      return (_e == hipErrorOutOfMemory) ? BF_ERR_NOMEM : BF_ERR_DEVICE;                 \
    }                                                                                    \
  } while (0)

template <class T>
int dalloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  BF_HIP(hipMalloc((void**)p, n * sizeof(T)));
  return BF_OK;
}
template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

void free_tiles(bf_ctx* c) {
  dfree(c->d_telem); dfree(c->d_tlslot); dfree(c->d_tcptr); dfree(c->d_tcent);
  dfree(c->d_tnode_ptr); dfree(c->d_tnode); dfree(c->d_tout); dfree(c->d_tpptr);
  dfree(c->d_tcomplete); dfree(c->d_tpartial); dfree(c->d_touched); dfree(c->d_deferred);
  dfree(c->d_twlist);
  c->n_tiles = c->n_partials = c->tileT = 0;
}

// LDS of the single-pass staged kernel for tiles of T elements and P chain threads each.
static size_t tile_smem_bytes(int T, int P, int NS) {
  return sizeof(double) * tile_smem_doubles(T, P, NS, tile_node_cap(T));
}
constexpr size_t kTileMaxLds = size_t(64) << 10;  // LDS of one workgroup

// Tile size for a request of T elements per tile and P points per element: 128 or 64, the
// larger whose workgroup (T * P threads) and LDS fit.
static int tile_fit(int T, int P, int NS) {
  T = T >= 128 ? 128 : 64;
  while (T > 64 && (T * P > kTileMaxThreads || tile_smem_bytes(T, P, NS) > kTileMaxLds)) T /= 2;
  return T;
}

// (Re)builds the single-pass tiling of a context's mesh with T elements per tile and
// uploads it (synchronous on the context stream).
int build_ctx_tiles(bf_ctx* c, int T) {
  TileHost th;
  if (build_tiles(c->E, c->n_v, c->n_nodes, c->NS, c->h_patch.data(), c->h_nen.data(),
                  c->h_x0.data(), c->h_y0.data(), c->h_z0.data(), T, th) != 0) {
    return BF_ERR_ARG;
  }
  BF_HIP(hipSetDevice(c->device));
  BF_HIP(hipStreamSynchronize(c->stream));
  free_tiles(c);
  const size_t nt = (size_t)std::max(th.n_tiles, 1);
  const size_t nloc = std::max<size_t>(th.node.size(), 1);
  int rc;
#define TALLOC(p, n) \
  if ((rc = dalloc(&(p), (n))) != BF_OK) { free_tiles(c); return rc; }
  TALLOC(c->d_telem, nt * T);
  TALLOC(c->d_tlslot, nt * (size_t)c->NS * T);
  TALLOC(c->d_tcptr, nloc + 1);
  TALLOC(c->d_tcent, std::max<size_t>(th.cent.size(), 1));
  TALLOC(c->d_tnode_ptr, nt + 1);
  TALLOC(c->d_tnode, nloc);
  TALLOC(c->d_tout, nloc);
  TALLOC(c->d_tpptr, (size_t)c->n_v + 1);
  TALLOC(c->d_tcomplete, (size_t)std::max(c->n_v, 1));
  TALLOC(c->d_tpartial, 3 * (size_t)std::max(th.n_partials, 1));
  TALLOC(c->d_touched, (size_t)std::max(c->n_nodes, 1));
  TALLOC(c->d_deferred, (size_t)std::max(c->E, 1));
  // the tile kernel's list (the FP64 hand-over list, the FP32-mixed second tier's input),
  // then the second tier's element list
  TALLOC(c->d_twlist, t1_list_entries(nt, T) + t2_list_entries((int)(nt * T / 64)));
#undef TALLOC
  hipStream_t s = c->stream;
  if (th.n_tiles > 0) {
    BF_HIP(hipMemcpyAsync(c->d_telem, th.elem.data(), sizeof(int) * th.elem.size(), hipMemcpyHostToDevice, s));
    BF_HIP(hipMemcpyAsync(c->d_tlslot, th.lslot.data(), sizeof(unsigned short) * th.lslot.size(), hipMemcpyHostToDevice, s));
    BF_HIP(hipMemcpyAsync(c->d_tcent, th.cent.data(), sizeof(unsigned short) * th.cent.size(), hipMemcpyHostToDevice, s));
    BF_HIP(hipMemcpyAsync(c->d_tnode, th.node.data(), sizeof(int) * th.node.size(), hipMemcpyHostToDevice, s));
    BF_HIP(hipMemcpyAsync(c->d_tout, th.out.data(), sizeof(int) * th.out.size(), hipMemcpyHostToDevice, s));
  }
  BF_HIP(hipMemcpyAsync(c->d_tnode_ptr, th.node_ptr.data(), sizeof(int) * th.node_ptr.size(), hipMemcpyHostToDevice, s));
  BF_HIP(hipMemcpyAsync(c->d_tpptr, th.pptr.data(), sizeof(int) * th.pptr.size(), hipMemcpyHostToDevice, s));
  BF_HIP(hipMemcpyAsync(c->d_tcptr, th.cptr.data(), sizeof(int) * th.cptr.size(), hipMemcpyHostToDevice, s));
  if (c->n_v > 0)
    BF_HIP(hipMemcpyAsync(c->d_tcomplete, th.complete.data(), th.complete.size(), hipMemcpyHostToDevice, s));
  BF_HIP(hipMemsetAsync(c->d_touched, 0, (size_t)std::max(c->n_nodes, 1), s));
  BF_HIP(hipStreamSynchronize(s));
  c->tileT = T;
  c->n_tiles = th.n_tiles;
  c->n_partials = th.n_partials;
  c->tile_lmax = tile_node_cap(T);
  return BF_OK;
}

void free_mesh(bf_ctx* c) {
  dfree(c->d_patch); dfree(c->d_meta); dfree(c->d_corner); dfree(c->d_ref); dfree(c->d_dA0);
  dfree(c->d_fib); dfree(c->d_gscale); dfree(c->d_delay); dfree(c->d_Fe); dfree(c->d_c33);
  dfree(c->d_flags); dfree(c->d_fallback); dfree(c->d_qlist);
  dfree(c->d_qsv); dfree(c->d_qst); dfree(c->d_redo);
  dfree(c->d_v); dfree(c->d_winv); dfree(c->d_fixed);
  c->has_state = false;
  dfree(c->d_ptr); dfree(c->d_ent); dfree(c->d_rules); dfree(c->d_x); dfree(c->d_y);
  dfree(c->d_z); dfree(c->d_fint);
  free_tiles(c);
  c->has_mesh = false;
}

int upload_tables(int device) {
  if (g_tables_device == device) return BF_OK;
  static double tab[13][5][kMaxNen];
  memset(tab, 0, sizeof(tab));
  std::vector<double> buf(5 * kMaxNen);
  for (int v = 3; v <= 12; v++) {
    if (bf_loop_tables(v, buf.data()) != BF_OK) return BF_ERR_ARG;
    const int nen = v + 6;
    for (int r = 0; r < 5; r++)
      for (int j = 0; j < nen; j++) tab[v][r][j] = buf[(size_t)r * nen + j];
  }
  BF_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_IN), tab, sizeof(tab)));
  BF_HIP(upload_in_tables_exact(&tab[0][0][0], sizeof(tab)));
  g_tables_device = device;
  return BF_OK;
}

int harvest_ring(bf_ctx* c) {
  for (int i = 0; i < c->ring_n; i++) {
    hipEvent_t* r = &c->ring[bf_ctx::kEv * i];
    BF_HIP(hipEventSynchronize(r[3]));
    float a = 0.f, h = 0.f, b = 0.f, v = 0.f;
    BF_HIP(hipEventElapsedTime(&a, r[0], r[1]));
    BF_HIP(hipEventElapsedTime(&h, r[1], r[2]));
    BF_HIP(hipEventElapsedTime(&b, r[2], r[3]));
    BF_HIP(hipEventElapsedTime(&v, r[4], r[5]));
    c->t_elem += a;
    c->t_handover += h;
    c->t_gather += b;
    c->t_solve += v;
    c->t_count++;
  }
  c->ring_n = 0;
  return BF_OK;
}

// Device address of `scratch` per device (hipGetSymbolAddress with that device current),
// filled once per device by the first context created on it.
constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

int noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return BF_ERR_ARG;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return BF_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    fprintf(stderr, "biofem_hip: hipGetSymbolAddress(scratch) failed on device %d: %s\n", device,
            hipGetErrorString(e));
    return BF_ERR_DEVICE;
  }
  g_scratch_ptr[device] = ptr;
  return BF_OK;
}

// Byte offsets of the per-evaluation arrays of a mesh (E elements, NS patch slots) in the
// static scratch, 256-byte aligned; returns the total.
struct ScratchLayout {
  size_t Fe, fallback, qlist, qsv, qst, redo, partial, touched, deferred, twlist,
      twqlist, total;
};
ScratchLayout noalloc_layout(size_t E, size_t NS, size_t n_nodes, size_t n_tiles, size_t T,
                             size_t n_partials) {
  auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
  ScratchLayout L{};
  const size_t En = std::max<size_t>(E, 1), nt = std::max<size_t>(n_tiles, 1);
  size_t o = 0;
  L.Fe = o;       o += up(sizeof(double) * 3 * NS * En);
  L.fallback = o; o += up(sizeof(int) * En);
  L.qlist = o;    o += up(sizeof(int) * kMaxQp * En);
  L.qsv = o;      o += up(sizeof(double) * 3 * kMaxQp * En);
  L.qst = o;      o += up(sizeof(int) * kMaxQp * En);
  L.redo = o;     o += up(sizeof(int) * En);
  L.partial = o;  o += up(sizeof(double) * 3 * std::max<size_t>(n_partials, 1));
  L.touched = o;  o += up(std::max<size_t>(n_nodes, 1));
  L.deferred = o; o += up(En);
  L.twlist = o;   o += up(sizeof(int) * (t1_list_entries(nt, T) + t2_list_entries((int)(nt * T / 64))));
  L.total = o;
  return L;
}

int kernel_launch_check() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    fprintf(stderr, "biofem_hip: kernel launch failed: %s\n", hipGetErrorString(e));
    return BF_ERR_DEVICE;
  }
  return BF_OK;
}

}  // namespace

extern "C" {

const char* bf_version(void) { return "biofem-hip 0.1 gfx950"; }

int bf_ctx_create(int device, bf_ctx** out) {
  if (!out) return BF_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  BF_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return BF_ERR_ARG;
  BF_HIP(hipSetDevice(device));
  bf_ctx* c = new (std::nothrow) bf_ctx();
  if (!c) return BF_ERR_NOMEM;
  c->device = device;
  {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
      c->n_cu = prop.multiProcessorCount;
  }
  int rc = upload_tables(device);
  if (rc == BF_OK) rc = noalloc_scratch_init(device);
  if (rc != BF_OK) { delete c; return rc; }
  bool events = true;
  for (auto& e : c->own_ev) events = events && hipEventCreate(&e) == hipSuccess;
  if (!events || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc((void**)&c->d_status, 2 * kStatusAlloc * sizeof(int)) != hipSuccess ||
      hipHostMalloc((void**)&c->h_status, kStatusWords * sizeof(int), hipHostMallocDefault) != hipSuccess) {
    bf_ctx_destroy(c);
    return BF_ERR_DEVICE;
  }
  for (int i = 0; i < bf_ctx::kEv; i++) c->ev[i] = c->own_ev[i];
  c->d_status_vis = c->d_status + kStatusAlloc;
  c->st_read = c->d_status;
  c->kp.fallback = 1;
  *out = c;
  return BF_OK;
}

int bf_ctx_destroy(bf_ctx* c) {
  if (!c) return BF_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_mesh(c);
  dfree(c->d_status);
  if (c->h_status) (void)hipHostFree(c->h_status);
  for (auto& e : c->own_ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : c->ring)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return BF_OK;
}

int bf_mesh_upload(bf_ctx* c, const bf_mesh_desc* m) {
  if (!c || !m) return BF_ERR_ARG;
  if (m->n_v < 0 || m->n_nodes < m->n_v || m->n_elmt < 0 || !m->patch || !m->ire || !m->val ||
      !m->x0 || !m->y0 || !m->z0 || !m->dA0 || !m->n_fib || m->patch_stride < 12 ||
      m->n_nodes >= BF_PATCH_SENTINEL || m->n_elmt >= (1 << 26))
    return BF_ERR_ARG;
  if (m->n_elmt_base > 0 && (!m->nv1 || !m->nv2 || !m->nv3)) return BF_ERR_ARG;
  if (m->n_ghost_rules < 0 || (m->n_ghost_rules > 0 && !m->ghost_rules) ||
      m->n_v + m->n_ghost_rules > m->n_nodes)
    return BF_ERR_ARG;
  BF_HIP(hipSetDevice(c->device));
  const int E = m->n_elmt, S = m->patch_stride;
  // ---- validate and pack topology (the per-residual checks of active_strain.c:498, :577,
  //      :599, :721 are done once here; they depend on topology only) ----
  std::vector<int> meta((size_t)E), nenv((size_t)E);
  int NS = 12;
  bool cap = false;
  for (int e = 0; e < E; e++) {
    const bool is_cap = m->n_elmt_base > 0 && e >= m->n_elmt_base;
    int kind, nen, v = m->val[e];
    if (is_cap) {
      kind = kCap;
      nen = (m->ire[e] == 0) ? 12 : v + 6;  // assembly width of FInternalAct (:2245)
      cap = true;
    } else if (m->ire[e] == 0) {
      kind = kRegular;
      nen = 12;
    } else if (m->ire[e] == 1) {
      kind = kIrregular;
      nen = v + 6;
      if (v < 3 || v > 12) {
        fprintf(stderr, "biofem_hip: element %d valence %d outside 3..12\n", e, v);
        return (v + 6 > BF_MAX_NEN) ? BF_ERR_UNSUPPORTED : BF_ERR_ARG;
      }
    } else {
      fprintf(stderr, "biofem_hip: element %d has ire=%d\n", e, m->ire[e]);
      return BF_ERR_ARG;
    }
    if (nen > S || nen > BF_MAX_NEN) {
      fprintf(stderr, "biofem_hip: element %d needs %d patch slots (stride %d)\n", e, nen, S);
      return BF_ERR_UNSUPPORTED;
    }
    if (!is_cap)
      for (int i = 0; i < nen; i++) {
        const int n = m->patch[(size_t)S * e + i];
        if (n == BF_PATCH_SENTINEL || n < 0 || n >= m->n_nodes) {
          fprintf(stderr, "biofem_hip: element %d patch slot %d incomplete (node %d)\n", e, i, n);
          return BF_ERR_ELEMENT;
        }
      }
    NS = std::max(NS, nen);
    nenv[e] = nen;
    meta[e] = kind | ((v & 255) << 8) | (nen << 16);
  }
  std::vector<int> patch((size_t)NS * std::max(E, 1), 0);
  for (int e = 0; e < E; e++)
    for (int i = 0; i < NS; i++) {
      int n = (i < S) ? m->patch[(size_t)S * e + i] : BF_PATCH_SENTINEL;
      // slots beyond nen are never read; park them on node 0 so any index is in range
      if (i >= nenv[e] || n == BF_PATCH_SENTINEL || n < 0 || n >= m->n_nodes) n = 0;
      patch[(size_t)i * E + e] = n;
    }
  std::vector<int> corner;
  if (cap) {
    corner.assign((size_t)6 * E, 0);
    for (int e = 0; e < E; e++) {
      const int tri[3] = {m->nv1[e], m->nv2[e], m->nv3[e]};
      for (int a = 0; a < 3; a++) {
        corner[(size_t)a * E + e] = tri[a];
        int slot = -1;
        if (m->n_elmt_base > 0 && e >= m->n_elmt_base) {
          const int nslot = m->val[e] + 6;  // :1811-1819
          for (int i = 0; i < nslot && i < S; i++)
            if (m->patch[(size_t)S * e + i] == tri[a]) { slot = i; break; }
          if (slot < 0 || slot >= nenv[e]) {
            fprintf(stderr, "biofem_hip: cap element %d corner %d missing from patch\n", e, tri[a]);
            return BF_ERR_ELEMENT;
          }
        }
        corner[(size_t)(3 + a) * E + e] = slot;
      }
    }
  }
  // ---- node -> (element, slot) CSR in ascending (element, slot) order (:2238-2260);
  //      ghost nodes get no list (their rows are dropped, :2254) ----
  std::vector<int> ptr((size_t)m->n_v + 1, 0);
  for (int e = 0; e < E; e++)
    for (int i = 0; i < nenv[e]; i++) {
      const int n = m->patch[(size_t)S * e + i];
      if (n >= 0 && n < m->n_v) ptr[n + 1]++;
    }
  for (int n = 0; n < m->n_v; n++) ptr[n + 1] += ptr[n];
  const int nnz = ptr[m->n_v];
  std::vector<unsigned> ent((size_t)std::max(nnz, 1));
  {
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int e = 0; e < E; e++)
      for (int i = 0; i < nenv[e]; i++) {
        const int n = m->patch[(size_t)S * e + i];
        if (n >= 0 && n < m->n_v) ent[fill[n]++] = ((unsigned)e << 5) | (unsigned)i;
      }
  }
  // ---- per-element SoA data ----
  std::vector<double> fib((size_t)3 * std::max(E, 1)), gs((size_t)std::max(E, 1), 1.0);
  for (int e = 0; e < E; e++) {
    fib[e] = m->n_fib[3 * (size_t)e];
    fib[(size_t)E + e] = m->n_fib[3 * (size_t)e + 1];
    fib[(size_t)2 * E + e] = m->n_fib[3 * (size_t)e + 2];
    if (m->gamma_scale) gs[e] = m->gamma_scale[e];
  }
  free_mesh(c);
  c->n_v = m->n_v;
  c->n_nodes = m->n_nodes;
  c->E = E;
  c->NS = NS;
  c->n_elmt_base = m->n_elmt_base;
  c->nnz = nnz;
  c->n_ghost = m->n_ghost_rules;
  c->has_delay = m->activation_delay != nullptr;
  c->has_cap = cap;
  const size_t En = (size_t)std::max(E, 1), Nn = (size_t)std::max(m->n_nodes, 1);
  int rc;
#define ALLOC(p, n) \
  if ((rc = dalloc(&(p), (n))) != BF_OK) { free_mesh(c); return rc; }
  ALLOC(c->d_patch, (size_t)NS * En);
  ALLOC(c->d_meta, En);
  ALLOC(c->d_ref, 15 * En);
  ALLOC(c->d_dA0, En);
  ALLOC(c->d_fib, 3 * En);
  ALLOC(c->d_gscale, En);
  ALLOC(c->d_Fe, (size_t)3 * NS * En);
  ALLOC(c->d_c33, (size_t)kMaxQp * En);
  ALLOC(c->d_flags, En);
  ALLOC(c->d_fallback, En);
  ALLOC(c->d_qlist, (size_t)kMaxQp * En);
  ALLOC(c->d_qsv, (size_t)3 * kMaxQp * En);
  ALLOC(c->d_qst, (size_t)kMaxQp * En);
  ALLOC(c->d_redo, En);
  ALLOC(c->d_ptr, (size_t)m->n_v + 1);
  ALLOC(c->d_ent, (size_t)std::max(nnz, 1));
  ALLOC(c->d_x, Nn);
  ALLOC(c->d_y, Nn);
  ALLOC(c->d_z, Nn);
  ALLOC(c->d_fint, 3 * Nn);
  if (c->has_delay) ALLOC(c->d_delay, En);
  if (cap) ALLOC(c->d_corner, 6 * En);
  if (c->n_ghost > 0) ALLOC(c->d_rules, (size_t)3 * c->n_ghost);
#undef ALLOC
  hipStream_t s = c->stream;
  BF_HIP(hipMemcpyAsync(c->d_patch, patch.data(), sizeof(int) * (size_t)NS * En, hipMemcpyHostToDevice, s));
  BF_HIP(hipMemcpyAsync(c->d_meta, meta.data(), sizeof(int) * (size_t)E, hipMemcpyHostToDevice, s));
  BF_HIP(hipMemcpyAsync(c->d_dA0, m->dA0, sizeof(double) * (size_t)E, hipMemcpyHostToDevice, s));
  BF_HIP(hipMemcpyAsync(c->d_fib, fib.data(), sizeof(double) * 3 * (size_t)E, hipMemcpyHostToDevice, s));
  BF_HIP(hipMemcpyAsync(c->d_gscale, gs.data(), sizeof(double) * (size_t)E, hipMemcpyHostToDevice, s));
  BF_HIP(hipMemcpyAsync(c->d_ptr, ptr.data(), sizeof(int) * ((size_t)m->n_v + 1), hipMemcpyHostToDevice, s));
  if (nnz > 0)
    BF_HIP(hipMemcpyAsync(c->d_ent, ent.data(), sizeof(unsigned) * (size_t)nnz, hipMemcpyHostToDevice, s));
  if (c->has_delay)
    BF_HIP(hipMemcpyAsync(c->d_delay, m->activation_delay, sizeof(double) * (size_t)E, hipMemcpyHostToDevice, s));
  if (cap)
    BF_HIP(hipMemcpyAsync(c->d_corner, corner.data(), sizeof(int) * 6 * (size_t)E, hipMemcpyHostToDevice, s));
  if (c->n_ghost > 0)
    BF_HIP(hipMemcpyAsync(c->d_rules, m->ghost_rules, sizeof(int) * 3 * (size_t)c->n_ghost, hipMemcpyHostToDevice, s));
  // reference coordinates -> geom0 cache (scratch: the resident coordinate buffers)
  BF_HIP(hipMemcpyAsync(c->d_x, m->x0, sizeof(double) * (size_t)m->n_nodes, hipMemcpyHostToDevice, s));
  BF_HIP(hipMemcpyAsync(c->d_y, m->y0, sizeof(double) * (size_t)m->n_nodes, hipMemcpyHostToDevice, s));
  BF_HIP(hipMemcpyAsync(c->d_z, m->z0, sizeof(double) * (size_t)m->n_nodes, hipMemcpyHostToDevice, s));
  if (E > 0) {
    DevMesh dm{};
    dm.patch = c->d_patch;
    dm.meta = c->d_meta;
    dm.corner = c->d_corner;
    BF_HIP(launch_geom0(dm, E, c->d_x, c->d_y, c->d_z, c->d_ref, s));
  }
  BF_HIP(hipStreamSynchronize(s));
  // single-pass staged tiling: an LV apex-cap element adds force to its three corner slots
  // only, the other slots are left out of the tiles
  if (cap)
    for (int e = m->n_elmt_base; e < E; e++)
      for (int i = 0; i < nenv[e]; i++)
        if (i != corner[(size_t)3 * E + e] && i != corner[(size_t)4 * E + e] &&
            i != corner[(size_t)5 * E + e])
          patch[(size_t)i * E + e] = -1;
  c->h_patch.swap(patch);
  c->h_nen.swap(nenv);
  c->h_x0.assign(m->x0, m->x0 + m->n_nodes);
  c->h_y0.assign(m->y0, m->y0 + m->n_nodes);
  c->h_z0.assign(m->z0, m->z0 + m->n_nodes);
  if ((rc = build_ctx_tiles(c, tile_fit(kTileElemsDefault, c->has_params ? c->kp.n_qp : 1, c->NS))) !=
      BF_OK) {
    free_mesh(c);
    return rc;
  }
  c->has_mesh = true;
  return BF_OK;
}

int bf_params_set(bf_ctx* c, const bf_params* p) {
  if (!c || !p) return BF_ERR_ARG;
  if (p->n_qp < 1 || p->n_qp > 5 || p->c33_iters < 1 || p->c33_iters > 31 || !(p->h0 > 0.0))
    return BF_ERR_ARG;
  if (p->ramp_type < 0 || p->ramp_type > 4) return BF_ERR_ARG;
  if (p->ramp_type == 4 && !(p->pulse_sigma > 0.0)) return BF_ERR_ARG;
  c->p = *p;
  KParams& k = c->kp;
  k.n_qp = p->n_qp;
  k.iters = p->c33_iters;
  // SetGaussianQuadrature (active_strain.c:184-222)
  static const double th[5][5] = {{0.0},
                                  {-0.5773502691896257, 0.5773502691896257},
                                  {-0.7745966692414834, 0.0, 0.7745966692414834},
                                  {-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526},
                                  {-0.9061798459386640, -0.5384693101056831, 0.0, 0.5384693101056831, 0.9061798459386640}};
  static const double wt[5][5] = {{2.0},
                                  {1.0, 1.0},
                                  {0.5555555555555556, 0.8888888888888888, 0.5555555555555556},
                                  {0.3478548451374538, 0.6521451548625461, 0.6521451548625461, 0.3478548451374538},
                                  {0.2369268850561891, 0.4786286704993665, 0.5688888888888889, 0.4786286704993665, 0.2369268850561891}};
  for (int i = 0; i < kMaxQp; i++) {
    k.theta[i] = i < p->n_qp ? th[p->n_qp - 1][i] : 0.0;
    k.w[i] = i < p->n_qp ? wt[p->n_qp - 1][i] : 0.0;
  }
  k.h0 = p->h0;
  k.mu = p->mu;
  k.K = p->K;
  k.act_mode = p->cart_fib_act ? 0 : (p->manufactured ? 1 : 2);
  RampParams& r = k.ramp;
  r.gamma = p->gamma;
  r.gamma_T = p->gamma_T;
  r.tau_up = p->tau_up;
  r.tau_down = p->tau_down;
  r.apd = p->apd;
  r.pulse_t0 = p->pulse_t0;
  r.pulse_sigma = p->pulse_sigma;
  r.t = (double)p->ti * p->dt;
  r.T_default = (double)p->tisteps * p->dt;
  r.type = p->ramp_type;
  k.gamma_t = ramp_gamma(r, 0.0);
  {
    const double t = p->ti * p->dt;  // active_strain.c:1241-1244
    const double s = sin(M_PI * t / p->manufactured_T);
    k.man_gamma = p->manufactured ? p->manufactured_gamma0 * pow(s, 2.0) : 0.0;
  }
  // a workgroup of the single-pass staged kernel holds T * n_qp threads
  if (c->has_mesh && tile_fit(c->tileT, p->n_qp, c->NS) != c->tileT) {
    const int rc = build_ctx_tiles(c, tile_fit(c->tileT, p->n_qp, c->NS));
    if (rc != BF_OK) return rc;
  }
  c->has_params = true;
  return BF_OK;
}

// The kernels' view of a context: its device arrays and its options.
static DevMesh dev_mesh(const bf_ctx* c) {
  DevMesh dm{};
  dm.patch = c->d_patch;
  dm.meta = c->d_meta;
  dm.corner = c->d_corner;
  dm.ref = c->d_ref;
  dm.dA0 = c->d_dA0;
  dm.fib = c->d_fib;
  dm.gscale = c->d_gscale;
  dm.delay = c->d_delay;
  dm.Fe = c->d_Fe;
  dm.c33 = c->d_c33;
  dm.status = c->d_status;
  dm.flags = c->d_flags;
  dm.fallback = c->d_fallback;
  dm.qlist = c->d_qlist;
  dm.qsv = c->d_qsv;
  dm.qst = c->d_qst;
  dm.redo = c->d_redo;
  return dm;
}
static KParams dev_params(const bf_ctx* c) {
  KParams kp = c->kp;
  kp.E = c->E;
  kp.NS = c->NS;
  kp.has_delay = c->has_delay ? 1 : 0;
  return kp;
}

// The single-pass tiling of a context as the kernels see it (its own per-evaluation arrays).
static TileDev tile_dev(const bf_ctx* c) {
  TileDev td{};
  td.n_tiles = c->n_tiles;
  td.NS = c->NS;
  td.T = c->tileT;
  td.lmax = c->tile_lmax;
  td.elem = c->d_telem;
  td.lslot = c->d_tlslot;
  td.cptr = c->d_tcptr;
  td.cent = c->d_tcent;
  td.node_ptr = c->d_tnode_ptr;
  td.node = c->d_tnode;
  td.out = c->d_tout;
  td.pptr = c->d_tpptr;
  td.complete = c->d_tcomplete;
  td.partial = c->d_tpartial;
  td.touched = c->d_touched;
  td.deferred = c->d_deferred;
  td.wlist = c->d_twlist;  // per-wave segments in batches; here the tile kernel's list
  td.wcount = nullptr;
  td.wqlist = nullptr;
  td.t1_tiles = t1_region_tiles(c->n_tiles);
  td.single = 1;
  return td;
}


// One force evaluation on stream s.  `view` / `tview` (optional) replace the context's own
// arrays: bf_force_eval_device_noalloc points the per-evaluation scratch into the static
// array.  The plain FP64 path evaluates every element in the reference's operation order and
// gathers node sums in ascending element order; the staged paths run the single-pass tile
// kernel, hand ill-conditioned elements to the reference order and finish in the node pass.
static int launch_eval(bf_ctx* c, const double* x, const double* y, const double* z,
                       double* fint, int accumulate, int precision, hipStream_t s,
                       bool step = false, const DevMesh* view = nullptr,
                       const TileDev* tview = nullptr) {
  if (!c->has_mesh || !c->has_params) return BF_ERR_STATE;
  if (precision != BF_PREC_FP64 && precision != BF_PREC_MIXED && precision != BF_PREC_FP64_FAST)
    return BF_ERR_ARG;
  const KParams kp = dev_params(c);
  const DevMesh dm = view ? *view : dev_mesh(c);
  TileDev td = tview ? *tview : tile_dev(c);
  // the tile kernel's list (counters zeroed by bf_status_reset): FP64, the hand-over list in
  // m.fallback (one counter, status[4]: regions spread the tiles' atomics, but the solve over
  // a region-ordered list measured 3.5 us slower than over the atomics' order); FP32-mixed,
  // the second tier's input (SegList of kRegions regions)
  const bool mixed = precision == BF_PREC_MIXED;
  td.list_count = dm.status + (mixed ? kT1Base : 4);
  td.list_e = mixed ? td.wlist : dm.fallback;
  td.list_ce = nullptr;
  td.list_case = 0;
  const bool staged = precision != BF_PREC_FP64;
  // the status words start reset: by the last staged step's node pass, or here
  if (!c->st_fresh) hipLaunchKernelGGL(bf_status_reset, dim3(1), dim3(64), 0, s, c->d_status);
  c->st_fresh = false;
  c->st_read = c->d_status;
  hipEvent_t* ev = c->own_ev;
  if (c->ring_on) {
    if (c->ring_n == bf_ctx::kRing) {
      int rc = harvest_ring(c);
      if (rc != BF_OK) return rc;
    }
    ev = &c->ring[bf_ctx::kEv * c->ring_n];
    c->ring_n++;
  }
  for (int i = 0; i < bf_ctx::kEv; i++) c->ev[i] = ev[i];  // bf_last_timing reads the latest set
  // Timing events only while timing is on: each record on the stream costs ~3.7 us of device
  // time between the kernels (five of them: 18 us per evaluation of config 2).  ev[3], after
  // the last kernel, is also recorded for an evaluation on a caller's stream: the status
  // reads on the context stream are ordered after it (on the context stream they follow
  // anyway).
  const bool timed = c->ring_on;
  c->timed_last = timed;
  c->ev3_recorded = timed || s != c->stream;
  hipEvent_t e0 = ev[0], e1 = ev[1], e2 = ev[2], e3 = ev[3];
  double* fout = step ? c->d_fint : fint;
  const int acc = step ? 0 : accumulate;
  if (timed) BF_HIP(hipEventRecord(e0, s));
  if (c->E > 0) {
    if (!staged) {
      BF_HIP(launch_element_exact(dm, kp, x, y, z, c->n_cu, s));
    } else {
      const int P = kp.n_qp;
      if (td.T * P > kTileMaxThreads) return BF_ERR_STATE;  // tiles not refitted to n_qp
      const size_t smem = tile_smem_bytes(td.T, P, td.NS);
      if (precision == BF_PREC_FP64_FAST)
        hipLaunchKernelGGL(bf_tile_kernel<double>, dim3(td.n_tiles), dim3(td.T * P), smem, s, dm,
                           kp, td, x, y, z, fout, acc, c->n_v);
      else
        hipLaunchKernelGGL(bf_tile_kernel<float>, dim3(td.n_tiles), dim3(td.T * P), smem, s, dm,
                           kp, td, x, y, z, fout, acc, c->n_v);
    }
  }
  if (timed) BF_HIP(hipEventRecord(e1, s));
  c->fb_prec = -1;
  if (c->E > 0 && staged && kp.fallback) {
    const int n_waves = td.n_tiles * (td.T / 64);
    const bool pairs = precision == BF_PREC_MIXED;
    const SegList t1 = pairs ? SegList{td.wlist, dm.status + kT1Base, td.t1_tiles * td.T, kRegions}
                             : SegList{dm.fallback, dm.status + 4, c->E, 1};
    int* t2_list = td.wlist + t1_list_entries(td.n_tiles, td.T);
    const SegList t2{t2_list, dm.status + kT2Base, t2_region_waves(n_waves) * 64, kRegions};
    const SegList cls{dm.qlist, dm.status + kClassBase, c->E, kp.n_qp};
    if (pairs)
      hipLaunchKernelGGL(bf_tile_tier2_kernel, dim3(tier2_waves(n_waves)), dim3(64), 0, s, dm, kp,
                         td, t2_list, t2_region_waves(n_waves), x, y, z);
    BF_HIP(launch_element_exact_fallback(dm, kp, x, y, z, c->n_cu, c->wide_hint[precision],
                                         pairs, pairs ? cls : t1, pairs ? t2 : t1,
                                         timed ? ev[4] : nullptr, timed ? ev[5] : nullptr, s));
    c->fb_prec = precision;
    c->last_list = pairs ? cls : t1;
  }
  if (c->fb_prec < 0 && timed) {  // no hand-over solve: an empty interval
    BF_HIP(hipEventRecord(ev[4], s));
    BF_HIP(hipEventRecord(ev[5], s));
  }
  if (timed) BF_HIP(hipEventRecord(e2, s));
  if (staged) {
    const bool handoff = c->n_nodes > 0;
    BF_HIP(launch_tile_fixup(td, c->d_ptr, c->d_ent, dm.Fe, c->E, c->n_v, c->n_nodes, acc, fout,
                             step, c->d_x, c->d_y, c->d_z, c->d_v, c->d_winv, c->d_fixed, c->dt,
                             c->damp, handoff ? c->d_status : nullptr, c->d_status_vis, s));
    if (handoff) {
      c->st_fresh = true;
      c->st_read = c->d_status_vis;
    }
  } else if (step) {
    BF_HIP(launch_gather_step(c->d_ptr, c->d_ent, dm.Fe, c->E, c->n_v, c->n_nodes, c->d_fint,
                              c->d_x, c->d_y, c->d_z, c->d_v, c->d_winv, c->d_fixed, c->dt,
                              c->damp, s));
  } else if (c->n_nodes > 0) {
    hipLaunchKernelGGL(bf_gather_kernel, dim3((c->n_nodes + 255) / 256), dim3(256), 0, s,
                       c->d_ptr, c->d_ent, dm.Fe, c->E, c->n_v, c->n_nodes, accumulate, fint);
  }
  if (c->ev3_recorded) BF_HIP(hipEventRecord(e3, s));
  c->timing = true;
  return kernel_launch_check();
}

// Orders the context's stream after the last evaluation, which may have run on a caller's
// stream (bf_force_eval_device): ev[3] is recorded on the evaluation's stream after its last
// kernel.  Every read of a per-evaluation result on the context's stream goes through here.
static int after_last_eval(bf_ctx* c) {
  if (c->timing && c->ev3_recorded) BF_HIP(hipStreamWaitEvent(c->stream, c->ev[3], 0));
  return BF_OK;
}

static int collect_status(bf_ctx* c, int* elem, int* code, int* n_warn) {
  if (int rc = after_last_eval(c)) return rc;
  BF_HIP(hipMemcpyAsync(c->h_status, c->st_read, kStatusWords * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  BF_HIP(hipStreamSynchronize(c->stream));
  const int* st = c->h_status;
  if (c->fb_prec >= 0)
    c->wide_hint[c->fb_prec] =
        c->fb_prec == BF_PREC_MIXED
            ? st[8] > 0 && split_fits(st[8], st[9], c->kp.n_qp, true, kWideSplitMin, 4 * c->n_cu)
            : st[4] > 0 && split_fits(st[4], 0, c->kp.n_qp, false, kWideSplitMin, 4 * c->n_cu);
  int el = -1, cd = 0;
  if (st[0] != INT_MAX) { el = st[0]; cd = BF_ELEM_NONFINITE; }
  else if (st[2] != INT_MAX) { el = st[2]; cd = BF_ELEM_CC3333; }
  if (elem) *elem = el;
  if (code) *code = cd;
  if (n_warn) *n_warn = st[3];
  return (st[0] != INT_MAX) ? BF_ERR_ELEMENT : BF_OK;
}

int bf_force_eval(bf_ctx* c, const double* x, const double* y, const double* z, double* fint,
                  int accumulate, int precision) {
  if (!c || !x || !y || !z || !fint) return BF_ERR_ARG;
  if (!c->has_mesh) return BF_ERR_STATE;
  BF_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const size_t N = (size_t)c->n_nodes;
  BF_HIP(hipMemcpyAsync(c->d_x, x, sizeof(double) * N, hipMemcpyHostToDevice, s));
  BF_HIP(hipMemcpyAsync(c->d_y, y, sizeof(double) * N, hipMemcpyHostToDevice, s));
  BF_HIP(hipMemcpyAsync(c->d_z, z, sizeof(double) * N, hipMemcpyHostToDevice, s));
  if (accumulate)
    BF_HIP(hipMemcpyAsync(c->d_fint, fint, sizeof(double) * 3 * N, hipMemcpyHostToDevice, s));
  int rc = launch_eval(c, c->d_x, c->d_y, c->d_z, c->d_fint, accumulate, precision, s);
  if (rc != BF_OK) return rc;
  BF_HIP(hipMemcpyAsync(fint, c->d_fint, sizeof(double) * 3 * N, hipMemcpyDeviceToHost, s));
  return collect_status(c, nullptr, nullptr, nullptr);
}

int bf_force_eval_device(bf_ctx* c, const double* d_x, const double* d_y, const double* d_z,
                         double* d_fint, int accumulate, int precision, void* stream) {
  if (!c || !d_x || !d_y || !d_z || !d_fint) return BF_ERR_ARG;
  BF_HIP(hipSetDevice(c->device));
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  return launch_eval(c, d_x, d_y, d_z, d_fint, accumulate, precision, s);
}

int bf_force_eval_device_noalloc(bf_ctx* c, const double* d_x, const double* d_y,
                                 const double* d_z, double* d_fint, int accumulate,
                                 int precision, void* stream) {
  if (!c || !d_x || !d_y || !d_z || !d_fint) return BF_ERR_ARG;
  if (!c->has_mesh || !c->has_params) return BF_ERR_STATE;
  if (c->device < 0 || c->device >= kMaxDevices || !g_scratch_ptr[c->device]) return BF_ERR_STATE;
  const ScratchLayout L = noalloc_layout((size_t)c->E, (size_t)c->NS, (size_t)c->n_nodes,
                                        (size_t)c->n_tiles, (size_t)c->tileT,
                                        (size_t)c->n_partials);
  if (L.total > kNoAllocScratchBytes) {
    fprintf(stderr, "biofem_hip: bf_force_eval_device_noalloc: the mesh needs %zu bytes of "
            "scratch (%d elements, %d patch slots), above the static maximum of %zu\n", L.total,
            c->E, c->NS, kNoAllocScratchBytes);
    return BF_ERR_UNSUPPORTED;
  }
  BF_HIP(hipSetDevice(c->device));
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  unsigned char* base = (unsigned char*)g_scratch_ptr[c->device];
  DevMesh dm = dev_mesh(c);
  dm.Fe = (double*)(base + L.Fe);
  dm.fallback = (int*)(base + L.fallback);
  dm.qlist = (int*)(base + L.qlist);
  dm.qsv = (double*)(base + L.qsv);
  dm.qst = (int*)(base + L.qst);
  dm.redo = (int*)(base + L.redo);
  TileDev td = tile_dev(c);
  td.partial = (double*)(base + L.partial);
  td.touched = (unsigned char*)(base + L.touched);
  td.deferred = (unsigned char*)(base + L.deferred);
  td.wlist = (int*)(base + L.twlist);
  // the touched marks start at zero (the node pass clears what the tile kernel sets; the
  // array is shared with other meshes' calls, so it is cleared here, by a kernel)
  if (c->n_nodes > 0)
    hipLaunchKernelGGL(bf_fill_u8, dim3((c->n_nodes + 255) / 256), dim3(256), 0, s, td.touched,
                       c->n_nodes, (unsigned char)0);
  return launch_eval(c, d_x, d_y, d_z, d_fint, accumulate, precision, s, false, &dm, &td);
}

int bf_ghost_update_device(bf_ctx* c, double* d_x, double* d_y, double* d_z, void* stream) {
  if (!c || !d_x || !d_y || !d_z) return BF_ERR_ARG;
  if (!c->has_mesh) return BF_ERR_STATE;
  if (c->n_ghost == 0) return BF_OK;
  BF_HIP(hipSetDevice(c->device));
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  hipLaunchKernelGGL(bf_ghost_kernel, dim3((c->n_ghost + 255) / 256), dim3(256), 0, s,
                     c->d_rules, c->n_ghost, c->n_v, d_x, d_y, d_z);
  return kernel_launch_check();
}

int bf_ghost_update(bf_ctx* c, double* x, double* y, double* z) {
  if (!c || !x || !y || !z) return BF_ERR_ARG;
  if (!c->has_mesh) return BF_ERR_STATE;
  BF_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const size_t N = (size_t)c->n_nodes;
  BF_HIP(hipMemcpyAsync(c->d_x, x, sizeof(double) * N, hipMemcpyHostToDevice, s));
  BF_HIP(hipMemcpyAsync(c->d_y, y, sizeof(double) * N, hipMemcpyHostToDevice, s));
  BF_HIP(hipMemcpyAsync(c->d_z, z, sizeof(double) * N, hipMemcpyHostToDevice, s));
  int rc = bf_ghost_update_device(c, c->d_x, c->d_y, c->d_z, s);
  if (rc != BF_OK) return rc;
  BF_HIP(hipMemcpyAsync(x, c->d_x, sizeof(double) * N, hipMemcpyDeviceToHost, s));
  BF_HIP(hipMemcpyAsync(y, c->d_y, sizeof(double) * N, hipMemcpyDeviceToHost, s));
  BF_HIP(hipMemcpyAsync(z, c->d_z, sizeof(double) * N, hipMemcpyDeviceToHost, s));
  BF_HIP(hipStreamSynchronize(s));
  return BF_OK;
}

int bf_c33_download(bf_ctx* c, double* c33) {
  if (!c || !c33) return BF_ERR_ARG;
  if (!c->has_mesh || !c->has_params) return BF_ERR_STATE;
  BF_HIP(hipSetDevice(c->device));
  const int q = c->kp.n_qp, E = c->E;
  std::vector<double> buf((size_t)q * std::max(E, 1));
  if (int rc = after_last_eval(c)) return rc;
  BF_HIP(hipMemcpyAsync(buf.data(), c->d_c33, sizeof(double) * (size_t)q * E, hipMemcpyDeviceToHost, c->stream));
  BF_HIP(hipStreamSynchronize(c->stream));
  for (int e = 0; e < E; e++)
    for (int k = 0; k < q; k++) c33[(size_t)e * q + k] = buf[(size_t)k * E + e];
  return BF_OK;
}

int bf_elem_flags_download(bf_ctx* c, unsigned char* flags) {
  if (!c || !flags) return BF_ERR_ARG;
  if (!c->has_mesh) return BF_ERR_STATE;
  BF_HIP(hipSetDevice(c->device));
  if (int rc = after_last_eval(c)) return rc;
  BF_HIP(hipMemcpyAsync(flags, c->d_flags, (size_t)c->E, hipMemcpyDeviceToHost, c->stream));
  BF_HIP(hipStreamSynchronize(c->stream));
  return BF_OK;
}

int bf_update_thickness(bf_ctx* c, double* thickness) {
  if (!c || !thickness) return BF_ERR_ARG;
  if (!c->has_mesh || !c->has_params) return BF_ERR_STATE;
  const int q = c->kp.n_qp, E = c->E;
  std::vector<double> c33((size_t)q * std::max(E, 1));
  int rc = bf_c33_download(c, c33.data());
  if (rc != BF_OK) return rc;
  for (int e = 0; e < E; e++) {  // active_strain.c:2211-2216
    double sum = 0.0;
    for (int k = 0; k < q; k++) sum += c33[(size_t)e * q + k];
    const double C33 = (q > 0) ? sum / (double)q : 1.0;
    const double lam = (C33 > 0.0) ? std::sqrt(C33) : 1.0;
    thickness[e] = c->p.h0 * lam;
  }
  return BF_OK;
}

int bf_last_error(bf_ctx* c, int* elem, int* code, int* n_warn) {
  if (!c) return BF_ERR_ARG;
  BF_HIP(hipSetDevice(c->device));
  const int rc = collect_status(c, elem, code, n_warn);
  return rc == BF_ERR_ELEMENT ? BF_OK : rc;  // the element failure is reported in *code
}

int bf_last_fallback_count(bf_ctx* c, int* n) {
  if (!c || !n) return BF_ERR_ARG;
  BF_HIP(hipSetDevice(c->device));
  const int rc = collect_status(c, nullptr, nullptr, nullptr);
  if (rc != BF_OK && rc != BF_ERR_ELEMENT) return rc;
  *n = c->h_status[4];
  return BF_OK;
}

int bf_last_handed_qps(bf_ctx* c, unsigned char* mask, int* n_pairs) {
  if (!c || !mask) return BF_ERR_ARG;
  if (!c->has_mesh) return BF_ERR_STATE;
  BF_HIP(hipSetDevice(c->device));
  const int rc = collect_status(c, nullptr, nullptr, nullptr);
  if (rc != BF_OK && rc != BF_ERR_ELEMENT) return rc;
  std::memset(mask, 0, (size_t)c->E);
  if (n_pairs) *n_pairs = 0;
  if (c->fb_prec < 0 || !c->last_list.v) return BF_OK;
  // FP32-mixed: a pair list; FP64: an element list, every qp solved.  The list's regions
  // (SegList) are read in order, their counts from the counters after the status words.
  const bool pairs = c->fb_prec == BF_PREC_MIXED;
  const SegList& L = c->last_list;
  std::vector<int> cnt(kStatusAlloc);
  BF_HIP(hipMemcpyAsync(cnt.data(), c->st_read, sizeof(int) * kStatusAlloc, hipMemcpyDeviceToHost, c->stream));
  BF_HIP(hipStreamSynchronize(c->stream));
  const size_t at = (size_t)(L.cnt - c->d_status);
  std::vector<int> list;
  for (int j = 0; j < L.n; j++) {
    const int n = cnt[at + (size_t)kCntStride * j];
    if (n < 0 || n > L.cap) return BF_ERR_DEVICE;
    const size_t o = list.size();
    list.resize(o + (size_t)n);
    if (n > 0)
      BF_HIP(hipMemcpyAsync(list.data() + o, L.v + (size_t)j * L.cap, sizeof(int) * (size_t)n,
                            hipMemcpyDeviceToHost, c->stream));
  }
  BF_HIP(hipStreamSynchronize(c->stream));
  const unsigned all = (1u << c->kp.n_qp) - 1u;
  int count = 0;
  for (int v : list) {
    const int e = pairs ? pair_elem(v) : v;
    const int q = pairs ? pair_qp(v) : 0;
    if (e < 0 || e >= c->E || q >= c->kp.n_qp) return BF_ERR_DEVICE;
    const unsigned add = pairs ? (1u << q) : all;
    count += __builtin_popcount(add & ~(unsigned)mask[e]);
    mask[e] |= (unsigned char)add;
  }
  if (n_pairs) *n_pairs = count;
  return BF_OK;
}

int bf_set_fallback(bf_ctx* c, int on) {
  if (!c) return BF_ERR_ARG;
  c->kp.fallback = on ? 1 : 0;
  return BF_OK;
}

int bf_last_timing(bf_ctx* c, float* ms_element, float* ms_gather) {
  if (!c) return BF_ERR_ARG;
  if (!c->timing || !c->timed_last) return BF_ERR_STATE;
  BF_HIP(hipEventSynchronize(c->ev[3]));
  float a = 0.f, b = 0.f;
  BF_HIP(hipEventElapsedTime(&a, c->ev[0], c->ev[2]));
  BF_HIP(hipEventElapsedTime(&b, c->ev[2], c->ev[3]));
  if (ms_element) *ms_element = a;
  if (ms_gather) *ms_gather = b;
  return BF_OK;
}

int bf_timing_enable(bf_ctx* c, int on) {
  if (!c) return BF_ERR_ARG;
  BF_HIP(hipSetDevice(c->device));
  if (on && c->ring.empty()) {
    c->ring.assign(bf_ctx::kEv * bf_ctx::kRing, nullptr);
    for (auto& e : c->ring) BF_HIP(hipEventCreate(&e));
  }
  if (!on && c->ring_on) {
    int rc = harvest_ring(c);
    if (rc != BF_OK) return rc;
  }
  c->ring_on = on != 0;
  c->ring_n = 0;
  c->t_count = 0;
  c->t_elem = c->t_handover = c->t_gather = c->t_solve = 0.0;
  return BF_OK;
}

int bf_timing_read(bf_ctx* c, long long* count, double* ms_element, double* ms_gather) {
  if (!c) return BF_ERR_ARG;
  BF_HIP(hipSetDevice(c->device));
  int rc = harvest_ring(c);
  if (rc != BF_OK) return rc;
  if (count) *count = c->t_count;
  if (ms_element) *ms_element = c->t_elem + c->t_handover;
  if (ms_gather) *ms_gather = c->t_gather;
  return BF_OK;
}

int bf_timing_read_split(bf_ctx* c, long long* count, double* ms_element,
                         double* ms_handover, double* ms_gather) {
  if (!c) return BF_ERR_ARG;
  BF_HIP(hipSetDevice(c->device));
  int rc = harvest_ring(c);
  if (rc != BF_OK) return rc;
  if (count) *count = c->t_count;
  if (ms_element) *ms_element = c->t_elem;
  if (ms_handover) *ms_handover = c->t_handover;
  if (ms_gather) *ms_gather = c->t_gather;
  return BF_OK;
}

int bf_timing_read_solve(bf_ctx* c, long long* count, double* ms_solve) {
  if (!c) return BF_ERR_ARG;
  BF_HIP(hipSetDevice(c->device));
  int rc = harvest_ring(c);
  if (rc != BF_OK) return rc;
  if (count) *count = c->t_count;
  if (ms_solve) *ms_solve = c->t_solve;
  return BF_OK;
}

int bf_device_buffers(bf_ctx* c, double** d_x, double** d_y, double** d_z, double** d_fint) {
  if (!c) return BF_ERR_ARG;
  if (!c->has_mesh) return BF_ERR_STATE;
  if (d_x) *d_x = c->d_x;
  if (d_y) *d_y = c->d_y;
  if (d_z) *d_z = c->d_z;
  if (d_fint) *d_fint = c->d_fint;
  return BF_OK;
}

int bf_sync(bf_ctx* c) {
  if (!c) return BF_ERR_ARG;
  BF_HIP(hipSetDevice(c->device));
  BF_HIP(hipStreamSynchronize(c->stream));
  return BF_OK;
}

int bf_state_init(bf_ctx* c, const double* mass, const unsigned char* fixed, const double* v0,
                  double dt, double damping) {
  if (!c || !mass || !(dt > 0.0) || !(damping >= 0.0 && damping <= 1.0)) return BF_ERR_ARG;
  if (!c->has_mesh) return BF_ERR_STATE;
  BF_HIP(hipSetDevice(c->device));
  const size_t N = (size_t)c->n_nodes;
  std::vector<double> winv(N, 0.0);
  std::vector<unsigned char> fx(N, 0);
  for (size_t n = 0; n < N; n++) {
    fx[n] = fixed ? (fixed[n] ? 1 : 0) : 0;
    if (n < (size_t)c->n_v && !fx[n]) {
      if (!(mass[n] > 0.0)) return BF_ERR_ARG;
      winv[n] = dt / mass[n];
    }
  }
  dfree(c->d_v); dfree(c->d_winv); dfree(c->d_fixed);
  c->has_state = false;
  int rc;
  if ((rc = dalloc(&c->d_v, 3 * N)) != BF_OK) return rc;
  if ((rc = dalloc(&c->d_winv, N)) != BF_OK) return rc;
  if ((rc = dalloc(&c->d_fixed, N)) != BF_OK) return rc;
  hipStream_t s = c->stream;
  if (v0)
    BF_HIP(hipMemcpyAsync(c->d_v, v0, sizeof(double) * 3 * N, hipMemcpyHostToDevice, s));
  else
    BF_HIP(hipMemsetAsync(c->d_v, 0, sizeof(double) * 3 * N, s));
  BF_HIP(hipMemcpyAsync(c->d_winv, winv.data(), sizeof(double) * N, hipMemcpyHostToDevice, s));
  BF_HIP(hipMemcpyAsync(c->d_fixed, fx.data(), N, hipMemcpyHostToDevice, s));
  BF_HIP(hipStreamSynchronize(s));
  c->dt = dt;
  c->damp = damping;
  c->has_state = true;
  return BF_OK;
}

int bf_step(bf_ctx* c, int precision, int n_steps) {
  if (!c || n_steps < 0) return BF_ERR_ARG;
  if (!c->has_mesh || !c->has_params || !c->has_state) return BF_ERR_STATE;
  BF_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  for (int k = 0; k < n_steps; k++) {
    if (c->n_ghost > 0) {
      hipLaunchKernelGGL(bf_ghost_kernel, dim3((c->n_ghost + 255) / 256), dim3(256), 0, s,
                         c->d_rules, c->n_ghost, c->n_v, c->d_x, c->d_y, c->d_z);
    }
    const int rc = launch_eval(c, c->d_x, c->d_y, c->d_z, c->d_fint, 0, precision, s, true);
    if (rc != BF_OK) return rc;
  }
  return BF_OK;
}

int bf_state_download(bf_ctx* c, double* x, double* y, double* z, double* v) {
  if (!c) return BF_ERR_ARG;
  if (!c->has_mesh) return BF_ERR_STATE;
  BF_HIP(hipSetDevice(c->device));
  const size_t N = (size_t)c->n_nodes;
  hipStream_t s = c->stream;
  if (x) BF_HIP(hipMemcpyAsync(x, c->d_x, sizeof(double) * N, hipMemcpyDeviceToHost, s));
  if (y) BF_HIP(hipMemcpyAsync(y, c->d_y, sizeof(double) * N, hipMemcpyDeviceToHost, s));
  if (z) BF_HIP(hipMemcpyAsync(z, c->d_z, sizeof(double) * N, hipMemcpyDeviceToHost, s));
  if (v) {
    if (!c->has_state) return BF_ERR_STATE;
    BF_HIP(hipMemcpyAsync(v, c->d_v, sizeof(double) * 3 * N, hipMemcpyDeviceToHost, s));
  }
  BF_HIP(hipStreamSynchronize(s));
  return BF_OK;
}

int bf_coords_upload(bf_ctx* c, const double* x, const double* y, const double* z) {
  if (!c || !x || !y || !z) return BF_ERR_ARG;
  if (!c->has_mesh) return BF_ERR_STATE;
  BF_HIP(hipSetDevice(c->device));
  const size_t N = (size_t)c->n_nodes;
  BF_HIP(hipMemcpyAsync(c->d_x, x, sizeof(double) * N, hipMemcpyHostToDevice, c->stream));
  BF_HIP(hipMemcpyAsync(c->d_y, y, sizeof(double) * N, hipMemcpyHostToDevice, c->stream));
  BF_HIP(hipMemcpyAsync(c->d_z, z, sizeof(double) * N, hipMemcpyHostToDevice, c->stream));
  BF_HIP(hipStreamSynchronize(c->stream));
  return BF_OK;
}

int bf_fint_download(bf_ctx* c, double* fint) {
  if (!c || !fint) return BF_ERR_ARG;
  if (!c->has_mesh) return BF_ERR_STATE;
  BF_HIP(hipSetDevice(c->device));
  BF_HIP(hipMemcpyAsync(fint, c->d_fint, sizeof(double) * 3 * (size_t)c->n_nodes,
                        hipMemcpyDeviceToHost, c->stream));
  BF_HIP(hipStreamSynchronize(c->stream));
  return BF_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ batches --------------
struct bf_batch {
  int device = 0;
  int n_cu = 256;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool timing_on = false;   // ev0 / ev1 around every bf_batch_eval / bf_batch_step call
  bool timed_last = false;  // the last call recorded them
  std::vector<bf_ctx*> cases;
  int P = 1;
  // launch shape of the element phase
  int passes = 1;       // kept for bf_batch_get_launch
  int launch_set = 0;   // 0: default tile size, 1: set or tuned
  int block = kTileElemsDefault;  // elements per tile requested for every case (64 or 128)
  int blocks = 0;       // kept for bf_batch_get_launch
  int persistent = 0;   // kept for bf_batch_get_launch
  int tile_threads = 0;   // threads per workgroup of the staged kernel: max over cases T n_qp
  size_t tile_smem = 0;   // its LDS: max over cases
  // device state
  CaseDev* d_cases = nullptr;
  Tile *d_etiles = nullptr, *d_ntiles = nullptr, *d_gtiles = nullptr, *d_xtiles = nullptr,
       *d_ttiles = nullptr;
  int *d_tlist = nullptr, *d_htotal = nullptr, *d_rtotal = nullptr;
  unsigned long long* d_tcount = nullptr;
  int *d_tqlist = nullptr, *d_ptotal = nullptr;
  int2 *d_hlist = nullptr, *d_rlist = nullptr, *d_plist = nullptr;
  int* d_cstat = nullptr;  // [2 n_cases]: per-case status words read by bf_batch_case_status
  int n_etiles = 0, n_ntiles = 0, n_gtiles = 0, n_xtiles = 0, n_ttiles = 0, tile_cap = 0;
  long long E_total = 0;
  int* h_total = nullptr;  // pinned
  // captured step
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int g_prec = -1, g_ghosts = -1, g_step = -1;
  // served rounds in flight (bf_batch_step_async): slot r & 1 holds round r's per-case
  // status words on the device and in pinned host memory, and its start / end events
  int* d_rstat[2] = {nullptr, nullptr};
  int* h_rstat[2] = {nullptr, nullptr};
  int rstat_cap = 0;
  hipEvent_t rev0[2] = {nullptr, nullptr}, rev1[2] = {nullptr, nullptr};
  long long r_round[2] = {-1, -1};
  int r_cases[2] = {0, 0};
};

namespace {

void batch_free_lists(bf_batch* b) {
  dfree(b->d_cases); dfree(b->d_etiles); dfree(b->d_ntiles); dfree(b->d_gtiles);
  dfree(b->d_xtiles); dfree(b->d_ttiles);
  dfree(b->d_tlist); dfree(b->d_tcount); dfree(b->d_hlist); dfree(b->d_htotal);
  dfree(b->d_tqlist); dfree(b->d_plist); dfree(b->d_ptotal);
  dfree(b->d_rlist); dfree(b->d_rtotal); dfree(b->d_cstat);
  b->n_etiles = b->n_ntiles = b->n_gtiles = b->n_xtiles = b->n_ttiles = 0;
}

void batch_drop_graph(bf_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->g_prec = b->g_ghosts = b->g_step = -1;
}

// Rebuilds the case table and the element / node / ghost work lists.
int batch_build(bf_batch* b) {
  batch_drop_graph(b);
  batch_free_lists(b);
  const int nc = (int)b->cases.size();
  if (nc == 0) return BF_OK;
  int P = 1;
  long long E_all = 0;
  for (bf_ctx* c : b->cases) {
    P = std::max(P, c->kp.n_qp);
    E_all += c->E;
  }
  b->P = P;
  (void)E_all;
  // Single-pass staged kernel: every case tiled with the requested tile size as a single
  // context of its n_qp would be (tile_fit), so a case's node sums have the same order in
  // the batch and alone; work items (case, tile), one per workgroup of the largest T n_qp
  // threads; hand-over segments per tile wave (64 entries), which the FP32-mixed second
  // tier and the compaction walk.
  const int Treq = (b->launch_set ? b->block : kTileElemsDefault) >= 128 ? 128 : 64;
  b->block = Treq;
  b->tile_threads = 64;
  b->tile_smem = 0;
  long long n_tw = 0;
  for (bf_ctx* c : b->cases) {
    const int T = tile_fit(Treq, c->kp.n_qp, c->NS);
    if (c->tileT != T) {
      const int rc = build_ctx_tiles(c, T);
      if (rc != BF_OK) return rc;
    }
    b->tile_threads = std::max(b->tile_threads, T * c->kp.n_qp);
    b->tile_smem = std::max(b->tile_smem, tile_smem_bytes(T, c->kp.n_qp, c->NS));
    n_tw += (long long)c->n_tiles * (T / 64);
  }
  b->tile_cap = 64;
  int rc;
  if ((rc = dalloc(&b->d_tlist, (size_t)std::max<long long>(n_tw, 1) * 64)) != BF_OK) return rc;
  if ((rc = dalloc(&b->d_tqlist, (size_t)std::max<long long>(n_tw, 1) * 64 * kMaxQp)) != BF_OK) {
    batch_free_lists(b);
    return rc;
  }
  std::vector<CaseDev> cases((size_t)nc);
  std::vector<Tile> et, nt, gt, xt, tt;
  std::vector<long long> case_wave((size_t)nc, 0);  // first tile wave of each case
  long long Et = 0, E_pairs = 0, wave_base = 0;
  for (int i = 0; i < nc; i++) {
    bf_ctx* c = b->cases[i];
    CaseDev& cd = cases[i];
    cd.m = dev_mesh(c);
    cd.kp = dev_params(c);
    cd.td = tile_dev(c);
    cd.td.wlist = b->d_tlist + wave_base * 64;
    cd.td.wcount = nullptr;  // set below, once the count arrays exist
    cd.td.wqlist = b->d_tqlist + wave_base * 64 * kMaxQp;
    cd.td.single = 0;  // per-wave segments (FP32-mixed), one (case, element) list (FP64)
    for (int t = 0; t < c->n_tiles; t++) {
      tt.push_back(Tile{i, t, 0, 0});
      for (int w = 0; w < c->tileT / 64; w++) et.push_back(Tile{i, t, w, 0});
    }
    case_wave[i] = wave_base;
    wave_base += (long long)c->n_tiles * (c->tileT / 64);
    cd.x = c->d_x;
    cd.y = c->d_y;
    cd.z = c->d_z;
    cd.fint = c->d_fint;
    cd.ptr = c->d_ptr;
    cd.ent = c->d_ent;
    cd.rules = c->d_rules;
    cd.n_v = c->n_v;
    cd.n_nodes = c->n_nodes;
    cd.n_ghost = c->n_ghost;
    cd.pad = 0;
    cd.v = c->has_state ? c->d_v : nullptr;
    cd.winv = c->d_winv;
    cd.fixed = c->d_fixed;
    cd.dt = c->dt;
    cd.damp = c->damp;
    for (int n0 = 0; n0 < c->n_nodes; n0 += 256) nt.push_back(Tile{i, n0, std::min(256, c->n_nodes - n0), 0});
    for (int g0 = 0; g0 < c->n_ghost; g0 += 256) gt.push_back(Tile{i, g0, std::min(256, c->n_ghost - g0), 0});
    for (int e0 = 0; e0 < c->E; e0 += 64) xt.push_back(Tile{i, e0, std::min(64, c->E - e0), 0});
    Et += c->E;
    E_pairs += (long long)c->E * c->kp.n_qp;
  }
  b->E_total = Et;
  b->n_etiles = (int)et.size();
  b->n_ntiles = (int)nt.size();
  b->n_gtiles = (int)gt.size();
  b->n_xtiles = (int)xt.size();
  b->n_ttiles = (int)tt.size();
#define BALLOC(p, n) \
  if ((rc = dalloc(&(p), (n))) != BF_OK) { batch_free_lists(b); return rc; }
  BALLOC(b->d_cases, (size_t)nc);
  BALLOC(b->d_etiles, std::max<size_t>(et.size(), 1));
  BALLOC(b->d_ntiles, std::max<size_t>(nt.size(), 1));
  BALLOC(b->d_gtiles, std::max<size_t>(gt.size(), 1));
  BALLOC(b->d_xtiles, std::max<size_t>(xt.size(), 1));
  BALLOC(b->d_ttiles, std::max<size_t>(tt.size(), 1));
  BALLOC(b->d_tcount, (size_t)std::max(b->n_etiles, 1));
  BALLOC(b->d_hlist, (size_t)std::max<long long>(Et, 1));
  BALLOC(b->d_htotal, 1);
  BALLOC(b->d_plist, (size_t)std::max<long long>(E_pairs, 1));
  BALLOC(b->d_ptotal, 3);
  BALLOC(b->d_rlist, (size_t)std::max<long long>(Et, 1));
  BALLOC(b->d_rtotal, 2);
  BALLOC(b->d_cstat, 2 * cases.size());
#undef BALLOC
  for (int i = 0; i < nc; i++) {
    cases[i].td.wcount = b->d_tcount + case_wave[i];
    cases[i].td.list_count = b->d_htotal;
    cases[i].td.list_e = nullptr;
    cases[i].td.list_ce = b->d_hlist;
    cases[i].td.list_case = i;
    cases[i].td.any = b->d_ptotal + 2;
  }
  hipStream_t s = b->stream;
  BF_HIP(hipMemcpyAsync(b->d_cases, cases.data(), sizeof(CaseDev) * cases.size(), hipMemcpyHostToDevice, s));
  if (!et.empty())
    BF_HIP(hipMemcpyAsync(b->d_etiles, et.data(), sizeof(Tile) * et.size(), hipMemcpyHostToDevice, s));
  if (!tt.empty())
    BF_HIP(hipMemcpyAsync(b->d_ttiles, tt.data(), sizeof(Tile) * tt.size(), hipMemcpyHostToDevice, s));
  BF_HIP(hipMemcpyAsync(b->d_ntiles, nt.data(), sizeof(Tile) * nt.size(), hipMemcpyHostToDevice, s));
  if (!gt.empty())
    BF_HIP(hipMemcpyAsync(b->d_gtiles, gt.data(), sizeof(Tile) * gt.size(), hipMemcpyHostToDevice, s));
  if (!xt.empty())
    BF_HIP(hipMemcpyAsync(b->d_xtiles, xt.data(), sizeof(Tile) * xt.size(), hipMemcpyHostToDevice, s));
  BF_HIP(hipStreamSynchronize(s));
  return BF_OK;
}

BatchDev batch_dev(const bf_batch* b) {
  BatchDev d{};
  d.cases = b->d_cases;
  d.n_cases = (int)b->cases.size();
  d.P = b->P;
  d.etiles = b->d_etiles;
  d.n_etiles = b->n_etiles;
  d.tile_cap = b->tile_cap;
  d.persistent = b->persistent;
  d.ntiles = b->d_ntiles;
  d.n_ntiles = b->n_ntiles;
  d.gtiles = b->d_gtiles;
  d.n_gtiles = b->n_gtiles;
  d.xtiles = b->d_xtiles;
  d.n_xtiles = b->n_xtiles;
  d.ttiles = b->d_ttiles;
  d.n_ttiles = b->n_ttiles;
  d.tlist = b->d_tlist;
  d.tcount = b->d_tcount;
  d.hlist = b->d_hlist;
  d.htotal = b->d_htotal;
  d.tqlist = b->d_tqlist;
  d.plist = b->d_plist;
  d.ptotal = b->d_ptotal;
  d.rlist = b->d_rlist;
  d.rtotal = b->d_rtotal;
  return d;
}

// The kernels of one batched step, enqueued on the batch stream (captured or eager).
// step: the node phase also advances the cases that hold a time-stepping state.
int batch_enqueue(bf_batch* b, int precision, int ghosts, bool step = false) {
  hipStream_t s = b->stream;
  for (bf_ctx* c : b->cases) {  // the batch resets and writes the cases' own status words
    c->st_fresh = false;
    c->st_read = c->d_status;
  }
  const BatchDev d = batch_dev(b);
  const int pblocks = std::max(1, std::min(b->n_cu * 2, std::max(b->n_gtiles, (d.n_cases + 255) / 256)));
  hipLaunchKernelGGL(bf_batch_prologue, dim3(pblocks), dim3(256), 0, s, d, ghosts);
  if (precision == BF_PREC_FP64) {
    BF_HIP(launch_batch_exact(d, b->n_cu, s));
    if (step)
      BF_HIP(launch_batch_gather_step(d, s));
    else
      hipLaunchKernelGGL(bf_batch_gather_kernel, dim3(b->n_ntiles), dim3(256), 0, s, d);
    return kernel_launch_check();
  }
  if (b->n_ttiles > 0) {
    const int nt = b->tile_threads, blocks = b->n_ttiles;
    const size_t smem = b->tile_smem;
    if (precision == BF_PREC_FP64_FAST) {
      hipLaunchKernelGGL(bf_batch_tile_kernel<double>, dim3(blocks), dim3(nt), smem, s, d);
    } else {
      hipLaunchKernelGGL(bf_batch_tile_kernel<float>, dim3(blocks), dim3(nt), smem, s, d);
      // one wave per tile wave (its segment), persistent over the work list
      hipLaunchKernelGGL(bf_batch_tier2_kernel, dim3(b->n_cu * 2), dim3(kStagedBlock), 0, s, d);
    }
    const bool pairs = precision == BF_PREC_MIXED;
    if (pairs)  // the FP64 tile kernel listed its elements itself
      hipLaunchKernelGGL(bf_batch_compact_kernel,
                         dim3(std::max(1, (b->n_etiles + kCompactBlock - 1) / kCompactBlock)),
                         dim3(kCompactBlock), 0, s, d, 1);
    BF_HIP(launch_batch_exact_list(d, b->n_cu, pairs, s));
  }
  BF_HIP(launch_batch_tile_fixup(d, step, s));
  return kernel_launch_check();
}

}  // namespace

extern "C" {

int bf_batch_create(int device, bf_batch** out) {
  if (!out) return BF_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  BF_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return BF_ERR_ARG;
  BF_HIP(hipSetDevice(device));
  bf_batch* b = new (std::nothrow) bf_batch();
  if (!b) return BF_ERR_NOMEM;
  b->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
    b->n_cu = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess ||
      hipHostMalloc((void**)&b->h_total, sizeof(int), hipHostMallocDefault) != hipSuccess) {
    bf_batch_destroy(b);
    return BF_ERR_DEVICE;
  }
  *out = b;
  return BF_OK;
}

int bf_batch_destroy(bf_batch* b) {
  if (!b) return BF_OK;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_drop_graph(b);
  batch_free_lists(b);
  for (int k = 0; k < 2; k++) {
    dfree(b->d_rstat[k]);
    if (b->h_rstat[k]) (void)hipHostFree(b->h_rstat[k]);
    if (b->rev0[k]) (void)hipEventDestroy(b->rev0[k]);
    if (b->rev1[k]) (void)hipEventDestroy(b->rev1[k]);
  }
  if (b->h_total) (void)hipHostFree(b->h_total);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return BF_OK;
}

int bf_batch_set_cases(bf_batch* b, bf_ctx* const* ctxs, int n) {
  if (!b || n < 0 || (n > 0 && !ctxs)) return BF_ERR_ARG;
  for (int i = 0; i < n; i++) {
    bf_ctx* c = ctxs[i];
    if (!c || c->device != b->device) return BF_ERR_ARG;
    if (!c->has_mesh || !c->has_params) return BF_ERR_STATE;
  }
  if (b->r_round[0] >= 0 || b->r_round[1] >= 0) return BF_ERR_STATE;  // rounds in flight
  BF_HIP(hipSetDevice(b->device));
  BF_HIP(hipStreamSynchronize(b->stream));
  b->cases.assign(ctxs, ctxs + n);
  return batch_build(b);
}

int bf_batch_refresh(bf_batch* b) {
  if (!b) return BF_ERR_ARG;
  BF_HIP(hipSetDevice(b->device));
  BF_HIP(hipStreamSynchronize(b->stream));
  return batch_build(b);
}

int bf_batch_set_launch(bf_batch* b, int passes, int block, int blocks, int persistent) {
  if (!b || passes < 1 || passes > 64 || (block != 64 && block != 128 && block != 256) || blocks < 0)
    return BF_ERR_ARG;
  BF_HIP(hipSetDevice(b->device));
  BF_HIP(hipStreamSynchronize(b->stream));
  b->passes = passes;
  b->block = block;
  b->blocks = blocks;
  b->persistent = persistent ? 1 : 0;
  b->launch_set = 1;
  return batch_build(b);
}

int bf_batch_get_launch(bf_batch* b, int* passes, int* block, int* blocks, int* persistent) {
  if (!b) return BF_ERR_ARG;
  if (passes) *passes = b->passes;
  if (block) *block = b->block;
  if (blocks) *blocks = b->blocks;
  if (persistent) *persistent = b->persistent;
  return BF_OK;
}

static int batch_run(bf_batch* b, int precision, int ghosts, bool step, int n, int use_graph) {
  if (!b || n < 0) return BF_ERR_ARG;
  if (precision != BF_PREC_FP64 && precision != BF_PREC_MIXED && precision != BF_PREC_FP64_FAST)
    return BF_ERR_ARG;
  if (b->cases.empty()) return BF_OK;
  BF_HIP(hipSetDevice(b->device));
  for (bf_ctx* c : b->cases) {  // also for graph replays, which skip batch_enqueue
    c->st_fresh = false;
    c->st_read = c->d_status;
  }
  // each event record costs a few microseconds of device time between the kernels
  if (b->timing_on) BF_HIP(hipEventRecord(b->ev0, b->stream));
  if (use_graph) {
    if (!b->exec || b->g_prec != precision || b->g_ghosts != (ghosts ? 1 : 0) ||
        b->g_step != (step ? 1 : 0)) {
      batch_drop_graph(b);
      BF_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      const int rc = batch_enqueue(b, precision, ghosts, step);
      hipGraph_t g = nullptr;
      const hipError_t ec = hipStreamEndCapture(b->stream, &g);
      if (rc != BF_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      BF_HIP(ec);
      b->graph = g;
      BF_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
      b->g_prec = precision;
      b->g_ghosts = ghosts ? 1 : 0;
      b->g_step = step ? 1 : 0;
    }
    for (int k = 0; k < n; k++) BF_HIP(hipGraphLaunch(b->exec, b->stream));
  } else {
    for (int k = 0; k < n; k++) {
      const int rc = batch_enqueue(b, precision, ghosts, step);
      if (rc != BF_OK) return rc;
    }
  }
  if (b->timing_on) BF_HIP(hipEventRecord(b->ev1, b->stream));
  b->timed_last = b->timing_on;
  return BF_OK;
}

int bf_batch_eval(bf_batch* b, int precision, int ghosts, int use_graph) {
  return batch_run(b, precision, ghosts, false, 1, use_graph);
}

int bf_batch_step(bf_batch* b, int precision, int n_steps, int use_graph) {
  return batch_run(b, precision, 1, true, n_steps, use_graph);
}

int bf_batch_sync(bf_batch* b) {
  if (!b) return BF_ERR_ARG;
  BF_HIP(hipSetDevice(b->device));
  BF_HIP(hipStreamSynchronize(b->stream));
  return BF_OK;
}

int bf_batch_timing_enable(bf_batch* b, int on) {
  if (!b) return BF_ERR_ARG;
  b->timing_on = on != 0;
  return BF_OK;
}

int bf_batch_last_ms(bf_batch* b, float* ms) {
  if (!b || !ms) return BF_ERR_ARG;
  if (!b->timed_last) return BF_ERR_STATE;
  BF_HIP(hipEventSynchronize(b->ev1));
  BF_HIP(hipEventElapsedTime(ms, b->ev0, b->ev1));
  return BF_OK;
}

int bf_batch_last_fallback_count(bf_batch* b, int* n) {
  if (!b || !n) return BF_ERR_ARG;
  *n = 0;
  if (b->cases.empty() || !b->d_htotal) return BF_OK;
  BF_HIP(hipSetDevice(b->device));
  BF_HIP(hipMemcpyAsync(b->h_total, b->d_htotal, sizeof(int), hipMemcpyDeviceToHost, b->stream));
  BF_HIP(hipStreamSynchronize(b->stream));
  *n = *b->h_total;
  return BF_OK;
}

int bf_batch_case_status(bf_batch* b, int* first_nonfinite, int* first_cc3333) {
  if (!b) return BF_ERR_ARG;
  const int n = (int)b->cases.size();
  if (n == 0) return BF_OK;
  if (!b->d_cstat) return BF_ERR_STATE;
  BF_HIP(hipSetDevice(b->device));
  std::vector<int> h((size_t)2 * n);
  hipLaunchKernelGGL(bf_batch_case_status_kernel, dim3((n + 255) / 256), dim3(256), 0, b->stream,
                     (const CaseDev*)b->d_cases, n, b->d_cstat);
  BF_HIP(hipGetLastError());
  BF_HIP(hipMemcpyAsync(h.data(), b->d_cstat, sizeof(int) * 2 * n, hipMemcpyDeviceToHost,
                        b->stream));
  BF_HIP(hipStreamSynchronize(b->stream));
  for (int c = 0; c < n; c++) {
    if (first_nonfinite) first_nonfinite[c] = h[2 * c] == INT_MAX ? -1 : h[2 * c];
    if (first_cc3333) first_cc3333[c] = h[2 * c + 1] == INT_MAX ? -1 : h[2 * c + 1];
  }
  return BF_OK;
}

int bf_batch_step_async(bf_batch* b, int precision, int n_steps, int use_graph, long long round) {
  if (!b || round < 0 || n_steps < 0) return BF_ERR_ARG;
  const int slot = (int)(round & 1);
  if (b->r_round[slot] >= 0) return BF_ERR_STATE;  // round - 2 not waited for
  const int n = (int)b->cases.size();
  BF_HIP(hipSetDevice(b->device));
  if (n > b->rstat_cap) {  // grow both slots; only with no round in flight
    if (b->r_round[slot ^ 1] >= 0) return BF_ERR_STATE;
    for (int k = 0; k < 2; k++) {
      dfree(b->d_rstat[k]);
      if (b->h_rstat[k]) (void)hipHostFree(b->h_rstat[k]);
      b->h_rstat[k] = nullptr;
    }
    b->rstat_cap = 0;
    for (int k = 0; k < 2; k++) {
      int rc = dalloc(&b->d_rstat[k], (size_t)2 * n);
      if (rc != BF_OK) return rc;
      BF_HIP(hipHostMalloc((void**)&b->h_rstat[k], sizeof(int) * 2 * n, hipHostMallocDefault));
    }
    b->rstat_cap = n;
  }
  for (int k = 0; k < 2; k++) {
    if (!b->rev0[k]) BF_HIP(hipEventCreate(&b->rev0[k]));
    if (!b->rev1[k]) BF_HIP(hipEventCreate(&b->rev1[k]));
  }
  BF_HIP(hipEventRecord(b->rev0[slot], b->stream));
  int rc = batch_run(b, precision, 1, true, n_steps, use_graph);
  if (rc != BF_OK) return rc;
  if (n > 0) {
    hipLaunchKernelGGL(bf_batch_case_status_kernel, dim3((n + 255) / 256), dim3(256), 0,
                       b->stream, (const CaseDev*)b->d_cases, n, b->d_rstat[slot]);
    BF_HIP(hipGetLastError());
    BF_HIP(hipMemcpyAsync(b->h_rstat[slot], b->d_rstat[slot], sizeof(int) * 2 * n,
                          hipMemcpyDeviceToHost, b->stream));
  }
  BF_HIP(hipEventRecord(b->rev1[slot], b->stream));
  b->r_round[slot] = round;
  b->r_cases[slot] = n;
  return BF_OK;
}

int bf_batch_step_wait(bf_batch* b, long long round, float* ms, int* first_nonfinite,
                       int* first_cc3333) {
  if (!b || round < 0) return BF_ERR_ARG;
  const int slot = (int)(round & 1);
  if (b->r_round[slot] != round) return BF_ERR_STATE;
  BF_HIP(hipSetDevice(b->device));
  BF_HIP(hipEventSynchronize(b->rev1[slot]));
  float t = 0.f;
  BF_HIP(hipEventElapsedTime(&t, b->rev0[slot], b->rev1[slot]));
  if (ms) *ms = t;
  const int n = b->r_cases[slot];
  const int* h = b->h_rstat[slot];
  for (int c = 0; c < n; c++) {
    if (first_nonfinite) first_nonfinite[c] = h[2 * c] == INT_MAX ? -1 : h[2 * c];
    if (first_cc3333) first_cc3333[c] = h[2 * c + 1] == INT_MAX ? -1 : h[2 * c + 1];
  }
  b->r_round[slot] = -1;
  return BF_OK;
}

int bf_batch_info(bf_batch* b, int* n_cases, long long* n_elmt, int* n_tiles, int* tile_cap) {
  if (!b) return BF_ERR_ARG;
  if (n_cases) *n_cases = (int)b->cases.size();
  if (n_elmt) *n_elmt = b->E_total;
  if (n_tiles) *n_tiles = b->n_etiles;
  if (tile_cap) *tile_cap = b->tile_cap;
  return BF_OK;
}

int bf_batch_tune(bf_batch* b, int precision, int reps, int* passes, int* block, int* blocks,
                  int* persistent, float* best_ms) {
  if (!b || reps < 1) return BF_ERR_ARG;
  if (b->cases.empty()) return BF_ERR_STATE;
  BF_HIP(hipSetDevice(b->device));
  // tile size of the single-pass staged kernel (elements per tile; one tile per workgroup)
  struct Cand { int passes, block, blocks, persistent; };
  std::vector<Cand> cands;
  const int block_opts[] = {64, 128};
  for (int bl : block_opts) cands.push_back({1, bl, 0, 0});
  float best = 1e30f;
  Cand bc{b->passes, b->block, b->blocks, b->persistent};
  b->launch_set = 1;
  for (const Cand& c : cands) {
    b->passes = c.passes;
    b->block = c.block;
    b->blocks = c.blocks;
    b->persistent = c.persistent;
    int rc = batch_build(b);
    if (rc != BF_OK) return rc;
    rc = batch_enqueue(b, precision, 0);  // warm-up
    if (rc != BF_OK) return rc;
    BF_HIP(hipEventRecord(b->ev0, b->stream));
    for (int r = 0; r < reps; r++) {
      rc = batch_enqueue(b, precision, 0);
      if (rc != BF_OK) return rc;
    }
    BF_HIP(hipEventRecord(b->ev1, b->stream));
    BF_HIP(hipEventSynchronize(b->ev1));
    float ms = 0.f;
    BF_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    ms /= (float)reps;
    if (ms < best) {
      best = ms;
      bc = c;
    }
  }
  b->passes = bc.passes;
  b->block = bc.block;
  b->blocks = bc.blocks;
  b->persistent = bc.persistent;
  const int rc = batch_build(b);
  if (rc != BF_OK) return rc;
  if (passes) *passes = bc.passes;
  if (block) *block = bc.block;
  if (blocks) *blocks = bc.blocks;
  if (persistent) *persistent = bc.persistent;
  if (best_ms) *best_ms = best;
  return BF_OK;
}

}  // extern "C"
