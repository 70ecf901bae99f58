// bf_exact.hip -- kernels compiled without floating-point contraction: the geom0 cache
// and the plain FP64 element kernel in the reference's operation order (bf_exact.h).
// Build flag: -ffp-contract=off (so a*b+c rounds twice, as in the reference's C).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

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
  if (kind == kRegular) {  // the same sums (c_Na / c_Nab hold the same values)
    G = mid_geom_regular(xyz);
  } else if (kind == kIrregular) {
    const double(*t)[kMaxNen] = c_IN[val];
    G = mid_geom_loaded(nen, t[0], t[1], t[2], t[3], t[4], xyz);
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


// ---- reference-order element path ---------------------------------------------------------
// ElemUpdateGeomSubdiv + ElemUpdateG + ElemActDefGrad + ElemCGDefTens + ElemC33Solve +
// ElemUpdFint, every operation in the order of src/active_strain.c, in three kernels:
//   1. solve (qp-major): each wave runs ONE through-thickness point q for 64 elements and
//      stores that qp's stresses, C33 and status.  Waves are qp-uniform, so structural
//      zeros of the contraction (a flat reference shell at its mid-surface point) are
//      skipped by whole waves, and cheap and expensive waves interleave on the machine.
//      The qps run independently: the CC3333 guard's cross-qp rule (ModElemC33 returns at
//      the first failing qp, :1674-1681) is not applied here, a failing element is only
//      flagged.
//   2. finish: one lane per element gathers the qp stresses and writes the nodal forces
//      (node sums over qp in the reference's order), flags and status; flagged elements are
//      listed for
//   3. redo: the element again with its qps in one lane group, the guard's cross-qp rule
//      as a wave ballot per iteration (exact_element).
// The forces of every element are then bit for bit the reference's.
constexpr int kExactBlock = 128;

// Midsurface geometry, shape-function rows and all inputs of the C33 loop of element e.
struct ExactRows {
  const double *na0, *na1, *nb0, *nb1, *nb2;
  int kind, nen;
};
__device__ __forceinline__ ExactRows exact_inputs(const DevMesh& m, const KParams& kp,
                                                  const double* __restrict__ x,
                                                  const double* __restrict__ y,
                                                  const double* __restrict__ z, int e,
                                                  exact::ElemIn& in, bool full) {
  const int E = kp.E;
  const int meta = m.meta[e];
  ExactRows r;
  r.kind = meta & 3;
  const int val = (meta >> 8) & 255;
  r.nen = (meta >> 16) & 255;
  auto xyz = [&](int i) { return ld3(x, y, z, m.patch[(size_t)i * E + e]); };
  // full (the solve kernels, where registers are scarce): the slot loop; otherwise every
  // slot's coordinates are loaded first (the same sums: c_Na / c_Nab hold kLoopNa / kLoopNab)
  if (r.kind == kRegular) {
    r.na0 = c_Na[0]; r.na1 = c_Na[1]; r.nb0 = c_Nab[0]; r.nb1 = c_Nab[1]; r.nb2 = c_Nab[2];
    in.cur = full ? mid_geom(12, r.na0, r.na1, r.nb0, r.nb1, r.nb2, xyz) : mid_geom_regular(xyz);
  } else if (r.kind == kIrregular) {
    const double(*t)[kMaxNen] = c_IN[val];
    r.na0 = t[0]; r.na1 = t[1]; r.nb0 = t[2]; r.nb1 = t[3]; r.nb2 = t[4];
    in.cur = full ? mid_geom(r.nen, r.na0, r.na1, r.nb0, r.nb1, r.nb2, xyz)
                  : mid_geom_loaded(r.nen, r.na0, r.na1, r.nb0, r.nb1, r.nb2, xyz);
  } else {
    r.na0 = r.na1 = r.nb0 = r.nb1 = r.nb2 = nullptr;
    in.cur = cap_geom(ld3(x, y, z, m.corner[e]), ld3(x, y, z, m.corner[(size_t)E + e]),
                      ld3(x, y, z, m.corner[(size_t)2 * E + e]));
  }
  if (!full) return r;
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
  return r;
}

// ElemUpdFint of element e from its qp stresses Sv (:1799-1801, :1807-1851 for the cap):
// nodal forces into the element-force buffer; returns false if any is not finite.
__device__ __forceinline__ bool exact_forces(const DevMesh& m, const KParams& kp, int e,
                                             const ExactRows& r, const MidGeom& cur,
                                             const double (*Sv)[3], int a0, int astep) {
  const int E = kp.E;
  const double A0 = m.dA0[e];
  const double half_h0 = 0.5 * kp.h0;
  const double pref0 = 0.5 * kp.h0 * A0;
  double* Fe = m.Fe;
  bool finite = true;
  if (r.kind != kCap) {
    for (int a = a0; a < r.nen; a += astep) {
      const V3<double> f = exact::node_force_exact(r.na0[a], r.na1[a], r.nb0[a], r.nb1[a],
                                                   r.nb2[a], cur, pref0, half_h0, Sv, kp.theta,
                                                   kp.w, kp.n_qp, r.kind == kIrregular);
      finite = finite && isfinite(f.x) && isfinite(f.y) && isfinite(f.z);
      put_force(Fe, e, E, a, f);
    }
  } else if (a0 == 0) {
    // zeroed element vector, corner contributions added
    const double dN[3][2] = {{-1.0, -1.0}, {1.0, 0.0}, {0.0, 1.0}};
    V3<double> fc[3];
    int sl[3];
    for (int k = 0; k < 3; k++) {
      sl[k] = m.corner[(size_t)(3 + k) * E + e];
      fc[k] = exact::cap_force_exact(dN[k][0], dN[k][1], cur, pref0, Sv, kp.w, kp.n_qp);
    }
    for (int a = 0; a < r.nen; a++) {
      V3<double> f = mk(0.0, 0.0, 0.0);
      for (int k = 0; k < 3; k++)
        if (sl[k] == a) {
          f = mk(f.x + fc[k].x, f.y + fc[k].y, f.z + fc[k].z);
          finite = finite && isfinite(fc[k].x) && isfinite(fc[k].y) && isfinite(fc[k].z);
        }
      put_force(Fe, e, E, a, f);
    }
  }
  return finite;
}

// Final per-element bookkeeping (flags, first non-finite element, CC3333 count).
__device__ __forceinline__ void exact_status(const DevMesh& m, int e, int status, bool finite,
                                             unsigned char extra_flags) {
  if (!finite) atomicMin(&m.status[0], e);
  m.flags[e] = (unsigned char)(status | (finite ? 0 : kStatNonFinite) | extra_flags);
  if (status & kStatCC3333) {
    atomicMin(&m.status[2], e);
    atomicAdd(&m.status[3], 1);
  }
}

// 1. solve: qp q of element e, qps independent.
__device__ __forceinline__ void exact_solve(const DevMesh& m, const KParams& kp,
                                            const double* __restrict__ x,
                                            const double* __restrict__ y,
                                            const double* __restrict__ z, int e, int q) {
  const int E = kp.E;
  exact::ElemIn in;
  exact_inputs(m, kp, x, y, z, e, in, true);
  // an element whose guard trips is redone with coupled qps: its solve can stop there
  const exact::QpOut r = exact::integrate_qp(in, q, [](bool fail, int) { return fail; },
                                             exact::CcSerial{}, true);
  m.c33[(size_t)q * E + e] = r.c33;
  double* o = m.qsv + 3 * ((size_t)q * E + e);
  o[0] = r.S00;
  o[1] = r.S11;
  o[2] = r.S01;
  m.qst[(size_t)q * E + e] = r.status;
}

// ---- the hand-over solve with K lanes per (element, qp) ----
// The hand-over list is short (a few percent of the elements), so its solve is bound by
// the time of one wave, not by the device's throughput: with one (element, qp) per lane
// most SIMDs idle while the waves that hold elements add their 6561 terms one by one.
// Here K lanes share one solve.  In round r lane j forms the 81 terms of block r K + j of
// the CC^{2222} sum into its own registers (all lanes at once); then the running sum
// travels through the group: lane 0 adds its block's terms in order and hands the sum to
// lane 1 (a DPP shift by one lane), which adds its own, and so on to lane K - 1, which hands
// it back to lane 0 for the next round.  These are the additions of cc_sum_serial in its
// order, so the sum is bit for bit the serial one, and every addition reads its term from
// the adding lane's registers (a dependent FP64 add is ~5 cycles there, ~16 when its term
// comes from another lane through LDS).  The nine Ce_inv entries of the iteration are divided
// among the K lanes the same way (ce_inv); everything else is formed by all K lanes alike.
// A wave holds split_groups(K) = 64 / K groups.

// Lane i receives lane i - 1's v (lane 0 keeps its own): DPP wave_shr:1, the whole-wave
// shift of GFX9's data-parallel primitives.
__device__ __forceinline__ double from_lane_below(double v) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(v), __double2loint(v), 0x138, 0xF,
                                             0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(v), __double2hiint(v), 0x138, 0xF,
                                             0xF, false);
  return __hiloint2double(hi, lo);
}

// Lane lead of each K-lane group receives lane lead + K - 1's v (other lanes receive
// anything): when K divides 16 a DPP row_ror:(17 - K), which rotates each row of 16 lanes so
// that lane i receives lane (i + K - 1) mod 16, else a cross-lane permute.
template <int K>
__device__ __forceinline__ double from_group_last(double v, int lead) {
  static_assert(K > 1 && K <= 64, "a group holds 2 to 64 lanes");
  if constexpr (16 % K == 0) {
    constexpr int ctrl = 0x120 + (17 - K);
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(v), __double2loint(v), ctrl, 0xF,
                                               0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(v), __double2hiint(v), ctrl, 0xF,
                                               0xF, false);
    return __hiloint2double(hi, lo);
  } else {
    return __shfl(v, lead + K - 1);
  }
}

// LDS copy of the factor matrices Fi, G, Ch (27 doubles) of each split group of a one-wave
// block (wide splits: at most 4 groups): a round's lane reads the factors of its block by
// computed address instead of selecting them from registers with lane-varying indices.
// Measured on the 256 x 192 membrane's 3 676 FP32-mixed pairs (K = 15): 136 us with register
// selects, 117 us with this array; sized for 32 groups, although 4 are used, because the
// 4-group size measured 131 us on the same device (the 184k-triangle ellipsoid's 11 pairs,
// K = 60: 9 us faster with 4).  At K <= 8 the register selects are faster (200 against
// 215 us for the FP64 list of the same membrane).
constexpr int kCcLdsGroups = 32;
template <int K>
constexpr bool cc_factors_from_lds() { return K >= kWideSplitMin; }
__device__ __forceinline__ double* cc_factor_lds() {
  __shared__ double s_fac[kCcLdsGroups * 27];
  return s_fac;
}

template <int K>
struct CcSplit {
  int j;  // lane within the group
  __device__ double operator()(const exact::CcIn& c) const {
    constexpr bool lds = cc_factors_from_lds<K>();
    static_assert(!lds || 64 / K <= kCcLdsGroups, "one LDS slot per group");
    const int lead = (threadIdx.x & 63) - j;
    const double* Fl = nullptr;
    const double* Gl = nullptr;
    const double* Cl = nullptr;
    if constexpr (lds) {
      double* fac = cc_factor_lds() + (size_t)((threadIdx.x & 63) / K) * 27;
      if (j == 0) {
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
          for (int col = 0; col < 3; col++) {
            fac[3 * r + col] = c.Fi.m[r][col];
            fac[9 + 3 * r + col] = c.G.m[r][col];
            fac[18 + 3 * r + col] = c.Ch.m[r][col];
          }
      }
      __syncthreads();  // one wave per block: orders the group leaders' writes before the reads
      Fl = fac;
      Gl = fac + 9;
      Cl = fac + 18;
    }
    double CC = 0.0;  // the running sum, valid in the lane whose block comes next
#pragma unroll 1
    for (int b0 = 0; b0 < 81; b0 += K) {
      const int b = b0 + j;
      exact::CcBlock k;
      bool act = false;
      if (b < 81) {
        if constexpr (lds) {  // the factors of block (g, h, m, n), as exact::cc_block_at forms them
          const int gg = b / 27, h = (b / 9) % 3, mm = (b / 3) % 3, n = b % 3;
          k.Figh = Fl[3 * gg + h];
          k.Gm2 = Gl[3 * mm + 2];
          k.Finm = Fl[3 * n + mm];
          k.Gg2 = Gl[3 * gg + 2];
          k.Ggn = Gl[3 * gg + n];
          k.Chhn = Cl[3 * h + n];
#pragma unroll
          for (int r = 0; r < 3; r++) {
            k.Chr[r] = Cl[3 * h + r];
            k.Chc[r] = Cl[3 * r + n];
          }
        } else {
          k = exact::cc_block_at(c, b);
        }
        act = exact::cc_block_active(c, k);
      }
      if (__ballot(act) == 0ull) continue;
      // t is read only by lanes with an active block, which form all of it
      double t[81];
      if (act) exact::cc_block_terms(c, k, [&](int i, double v) { t[i] = v; });
#pragma unroll 1
      for (int sl = 0; sl < K; sl++) {
        if (j == sl && act) {
#pragma unroll
          for (int i = 0; i < 81; i++) CC += t[i];
        }
        if (sl + 1 < K) CC = from_lane_below(CC);
      }
      CC = from_group_last<K>(CC, lead);  // back to lane 0 of the group
    }
    const double out = __shfl(CC, lead);
    if constexpr (lds) __syncthreads();  // the factors are read before the next call's writes
    return out;
  }
  // The nine Ce_inv entries: in round r lane j forms entry r K + j (all lanes at once, each
  // entry one sum in the reference's order), then every lane of the group takes each entry
  // from the lane that formed it.
  __device__ void ce_inv(const exact::Mat& FaCont, const exact::Mat& gm0, const exact::Mat& Ci,
                         exact::Mat& Ch) const {
    constexpr int R = (9 + K - 1) / K;
    const int lead = (int)(threadIdx.x & 63) - j;
    double v[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int k = r * K + j;
      v[r] = k < 9 ? exact::ce_entry(FaCont, gm0, Ci, k / 3, k % 3) : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 9; k++) Ch.m[k / 3][k % 3] = __shfl(v[k / K], lead + k % K);
  }
  // The total-stress entries: S^22 by every lane; in the last iteration the four entries
  // S^00, S^11, S^01, S^22 are divided among the lanes like the Ce_inv entries.
  __device__ void stresses(const exact::Mat& Fi, const exact::Mat& G, const exact::Mat& Se,
                           bool last, exact::TotOut& o) const {
    if (!last) {
      o.S22 = exact::tot_entry(Fi, G, Se, 2, 2);
      return;
    }
    constexpr int R = (4 + K - 1) / K;
    const int lead = (int)(threadIdx.x & 63) - j;
    double v[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int k = r * K + j;  // 0: S^00, 1: S^11, 2: S^01, 3: S^22
      const int ti = k == 1 ? 1 : (k == 3 ? 2 : 0), tj = k == 0 ? 0 : (k == 3 ? 2 : 1);
      v[r] = k < 4 ? exact::tot_entry(Fi, G, Se, ti, tj) : 0.0;
    }
    o.S00 = __shfl(v[0 / K], lead + 0 % K);
    o.S11 = __shfl(v[1 / K], lead + 1 % K);
    o.S01 = __shfl(v[2 / K], lead + 2 % K);
    o.S22 = __shfl(v[3 / K], lead + 3 % K);
  }
};

// Split factor for a hand-over list (split_fits): the largest K that fits.  K = 60, 30, 15
// (short lists) run in bf_exact_solve_wide_kernel (`wide`), K = 8, 6, 4, 2, 1 in
// bf_exact_solve_queue_kernel and the batch list kernel (not `wide`): the two sets compile as
// separate kernels, so neither's register allocation pays for the other's instantiations
// (more K in one kernel slowed its common cases).
__device__ __forceinline__ int split_factor(int N, int N_mid, int P, bool pairs, int n_waves,
                                            bool wide) {
  if (wide) {
    if (split_fits(N, N_mid, P, pairs, 60, n_waves)) return 60;
    if (split_fits(N, N_mid, P, pairs, 30, n_waves)) return 30;
    if (split_fits(N, N_mid, P, pairs, 15, n_waves)) return 15;
  }
  for (int K = 8; K >= 2; K -= 2)
    if (split_fits(N, N_mid, P, pairs, K, n_waves)) return K;
  return 1;
}

// Work items of a hand-over list for G groups per wave: an element list (entries are
// elements, every qp solved; item i = (qp rank i / chunks, chunk of G entries i % chunks)) or
// a pair list (entries qp_pair(e, q); item i = entries i G ..).  at(i, g) gives the entry
// and qp of group g of item i, false past the list.
struct HandItems {
  int N, P, G;
  bool pairs;
  __device__ int n_items() const {
    const int c = (N + G - 1) / G;
    return pairs ? c : c * P;
  }
  __device__ bool at(int i, int g, int& idx, int& q) const {
    if (pairs) {
      idx = i * G + g;
      q = -1;  // from the entry
      return idx < N;
    }
    const int c = (N + G - 1) / G;
    idx = (i % c) * G + g;
    q = qp_of_rank(i / c, P);
    return idx < N;
  }
};

// Solve of qp q of element e by lane j of a K-lane group (K = 1: one lane, serial sum).
template <int K>
__device__ __forceinline__ void exact_solve_k(const DevMesh& m, const KParams& kp,
                                              const double* __restrict__ x,
                                              const double* __restrict__ y,
                                              const double* __restrict__ z, int e, int q,
                                              int j) {
  if constexpr (K == 1) {
    exact_solve(m, kp, x, y, z, e, q);
  } else {
    const int E = kp.E;
    exact::ElemIn in;
    exact_inputs(m, kp, x, y, z, e, in, true);
    const exact::QpOut r = exact::integrate_qp(in, q, [](bool fail, int) { return fail; },
                                               CcSplit<K>{j}, true);
    if (j != 0) return;
    m.c33[(size_t)q * E + e] = r.c33;
    double* o = m.qsv + 3 * ((size_t)q * E + e);
    o[0] = r.S00;
    o[1] = r.S11;
    o[2] = r.S01;
    m.qst[(size_t)q * E + e] = r.status;
  }
}

// 2. finish: element e from the solve's qp results; guard trips go to the redo (returns
// true: redo(e) was called by slice 0).  kFinishLanes consecutive lanes share an element
// (the same e, so the same branches): lane `slice` writes nodes slice, slice +
// kFinishLanes, ...; the lanes combine their finite flags and slice 0 does the element's
// bookkeeping.
constexpr int kFinishLanes = 4;

template <class Redo>
__device__ __forceinline__ bool exact_finish(const DevMesh& m, const KParams& kp,
                                             const double* __restrict__ x,
                                             const double* __restrict__ y,
                                             const double* __restrict__ z, int e,
                                             unsigned char extra_flags, const Redo& redo,
                                             int slice) {
  const int E = kp.E;
  int status = 0;
  double Sv[kMaxQp][3];
  int st[kMaxQp];
#pragma unroll
  for (int j = 0; j < kMaxQp; j++) {  // branch-free: qps past n_qp read the last one
    const int jj = j < kp.n_qp ? j : kp.n_qp - 1;
    const double* o = m.qsv + 3 * ((size_t)jj * E + e);
    Sv[j][0] = o[0];
    Sv[j][1] = o[1];
    Sv[j][2] = o[2];
    st[j] = m.qst[(size_t)jj * E + e];
  }
#pragma unroll
  for (int j = 0; j < kMaxQp; j++) {
    if (j < kp.n_qp) {
      status |= st[j];
    } else {
      Sv[j][0] = Sv[j][1] = Sv[j][2] = 0.0;
    }
  }
  if (status & kStatCC3333) {  // the cross-qp guard rule matters: redo with coupled qps
    if (slice == 0) redo(e);
    return true;
  }
  exact::ElemIn in;
  const ExactRows r = exact_inputs(m, kp, x, y, z, e, in, false);
  int bad = exact_forces(m, kp, e, r, in.cur, Sv, slice, kFinishLanes) ? 0 : 1;
#pragma unroll
  for (int k = 1; k < kFinishLanes; k <<= 1) bad |= __shfl_xor(bad, k, kFinishLanes);
  if (slice == 0) exact_status(m, e, status, bad == 0, extra_flags);
  return false;
}

// 3. redo: one lane group per element, lane q runs qp q; the guard of iteration it skips
//    qp q if any qp <= q failed (wave ballot), then the group exchanges the qp stresses and
//    each lane writes every n_qp-th node.
struct QpGroup {
  int e;      // element of this lane's group
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
  exact::ElemIn in;
  const ExactRows rows = exact_inputs(m, kp, x, y, z, e, in, true);
  const unsigned long long upto = (2ull << q) - 1ull;
  const exact::QpOut r = exact::integrate_qp(in, q, [&](bool fail, int) {
    return ((__ballot(fail) >> grp.base) & upto) != 0ull;
  });
  m.c33[(size_t)q * E + e] = r.c33;
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
  const bool finite = exact_forces(m, kp, e, rows, in.cur, Sv, q, n_qp);
  int bad = finite ? 0 : 1;
#pragma unroll
  for (int j = 0; j < kMaxQp; j++)
    if (j < n_qp) bad |= __shfl(bad, grp.base + j);
  if (q != 0) return;
  exact_status(m, e, status, bad == 0, extra_flags);
}

// ---- single context ----
// Solve work items are issued qp by qp in qp_of_rank order (the mid-surface point last).

// Work item i of the solve = (qp rank i / chunks, chunk of 64 list entries i % chunks).
// list == nullptr: element = entry index.
__global__ void __launch_bounds__(kExactBlock) bf_exact_solve_kernel(DevMesh m, KParams kp,
                                                                 const double* __restrict__ x,
                                                                 const double* __restrict__ y,
                                                                 const double* __restrict__ z,
                                                                 const int* __restrict__ list,
                                                                 const int* __restrict__ count,
                                                                 int n_static) {
  const int N = count ? *count : n_static;
  const int P = kp.n_qp;
  const int chunks = (N + 63) >> 6;
  const int items = chunks * P;
  const int lane = threadIdx.x & 63;
  const int waves = gridDim.x * (kExactBlock / 64);
  for (int i = blockIdx.x * (kExactBlock / 64) + (threadIdx.x >> 6); i < items; i += waves) {
    const int q = qp_of_rank(i / chunks, P), idx = (i % chunks) * 64 + lane;
    if (idx < N) exact_solve(m, kp, x, y, z, list ? list[idx] : idx, q);
  }
}

// Work items of a persistent grid: wave w takes item w first, then items waves + k in the
// order the waves ask a device counter (zero on entry) for them, so the cheap items of the
// tail go to the waves that are free first.  Every wave leaves once the counter has passed
// the item count.
__device__ __forceinline__ int wave_id() {
  return blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
}
__device__ __forceinline__ int next_item(int* counter) {
  int i = 0;
  if ((threadIdx.x & 63) == 0) i = atomicAdd(counter, 1);
  return __shfl(i, 0) + gridDim.x * (blockDim.x >> 6);
}

// The hand-over solve at split factor K over the first N entries of a SegList (HandItems):
// wave w takes item w first, then items from the counter.
template <int K>
__device__ void solve_list_k(const DevMesh& m, const KParams& kp, const double* __restrict__ x,
                             const double* __restrict__ y, const double* __restrict__ z,
                             SegList list, SegCount sc, int N, bool pairs, int* counter) {
  constexpr int NG = split_groups(K);
  const HandItems hi{N, kp.n_qp, NG, pairs};
  const int items = hi.n_items();
  const int lane = threadIdx.x & 63;
  const int g = lane / K, j = lane % K;
  for (int i = wave_id(); i < items; i = next_item(counter)) {
    int idx, q;
    if (g < NG && hi.at(i, g, idx, q)) {
      const int v = seg_at(list, sc, idx);
      const int e = pairs ? pair_elem(v) : v;
      exact_solve_k<K>(m, kp, x, y, z, e, pairs ? pair_qp(v) : q, j);
    }
  }
}

// Mid-surface entries at the end of a hand-over list: the pair classes are the qp ranks, the
// mid-surface point last (has_mid_qp); element lists have none.
__device__ __forceinline__ int list_mid(const SegCount& sc, int P, bool pairs) {
  int v = 0;  // sc.c[P - 1] by selects: a run-time index would put sc in scratch memory
#pragma unroll
  for (int j = 0; j < kRegions; j++) v = (j == P - 1) ? sc.c[j] : v;
  return (pairs && has_mid_qp(P)) ? v : 0;
}

// Solve over the hand-over list: one-wave blocks, persistent grid of n_waves waves, split
// factor from the list length.  pairs: the pair classes of m.qlist, else the element regions
// of m.fallback.  narrow (not null): the wide kernel ran first and left here 0 if it took the
// list.  The first lane also stores the list's length in the status words the host reads
// (pairs: status[8], status[9]; elements: status[4] is stored by the finish kernel).
__global__ void __launch_bounds__(64) bf_exact_solve_queue_kernel(
    DevMesh m, KParams kp, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, SegList list, const int* __restrict__ narrow, int pairs,
    int* __restrict__ counter, int n_waves) {
  const SegCount sc = seg_count(list);
  const int N_mid = list_mid(sc, kp.n_qp, pairs != 0);
  if (pairs && blockIdx.x == 0 && threadIdx.x == 0) {
    m.status[8] = sc.total;
    m.status[9] = N_mid;
  }
  const int N = narrow ? *narrow : sc.total;
  switch (split_factor(N, N > 0 ? N_mid : 0, kp.n_qp, pairs != 0, n_waves, false)) {
    case 8: solve_list_k<8>(m, kp, x, y, z, list, sc, N, pairs != 0, counter); break;
    case 6: solve_list_k<6>(m, kp, x, y, z, list, sc, N, pairs != 0, counter); break;
    case 4: solve_list_k<4>(m, kp, x, y, z, list, sc, N, pairs != 0, counter); break;
    case 2: solve_list_k<2>(m, kp, x, y, z, list, sc, N, pairs != 0, counter); break;
    default: solve_list_k<1>(m, kp, x, y, z, list, sc, N, pairs != 0, counter); break;
  }
}

// The short lists: 15, 30 or 60 lanes per (element, qp), so a round forms 15 to 60 of the
// sum's 81 blocks at once.  Runs before bf_exact_solve_queue_kernel and hands it its count:
// 0 when this kernel takes the list, the list length when the list is too long for a wide
// split (then this kernel does nothing else).
__global__ void __launch_bounds__(64) bf_exact_solve_wide_kernel(
    DevMesh m, KParams kp, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, SegList list, int pairs, int* __restrict__ counter,
    int n_waves, int* __restrict__ narrow_count) {
  const SegCount sc = seg_count(list);
  const int N = sc.total;
  const int N_mid = list_mid(sc, kp.n_qp, pairs != 0);
  const int K = N > 0 ? split_factor(N, N_mid, kp.n_qp, pairs != 0, n_waves, true) : 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) *narrow_count = K >= kWideSplitMin ? 0 : N;
  switch (K) {
    case 60: solve_list_k<60>(m, kp, x, y, z, list, sc, N, pairs != 0, counter); break;
    case 30: solve_list_k<30>(m, kp, x, y, z, list, sc, N, pairs != 0, counter); break;
    case 15: solve_list_k<15>(m, kp, x, y, z, list, sc, N, pairs != 0, counter); break;
    default: break;  // the queue kernel's
  }
}

struct RedoSink {
  int* list;
  int* count;
  __device__ void operator()(int e) const { list[atomicAdd(count, 1)] = e; }
};

// list.v == nullptr: elements 0 .. n_static - 1; else the list's entries, whose length the
// first lane stores in status[4] (the count the host reads).  kInlineRedo (n_qp <=
// kFinishLanes): an element whose guard tripped is redone right here by its own lanes
// (lane q runs qp q), instead of by bf_exact_redo_kernel after every finish -- it depends on
// no other element, and one launch fewer is ~5 us of the hand-over chain.  The redo's
// registers cut this kernel to one wave per SIMD, which a hand-over list does not fill.
struct NoRedo {
  __device__ void operator()(int) const {}
};
template <bool kInlineRedo>
__global__ void __launch_bounds__(256) bf_exact_finish_kernel(DevMesh m, KParams kp,
                                                              const double* __restrict__ x,
                                                              const double* __restrict__ y,
                                                              const double* __restrict__ z,
                                                              SegList list, int n_static,
                                                              unsigned char extra_flags) {
  SegCount sc{};
  if (list.v) {
    sc = seg_count(list);
    if (blockIdx.x == 0 && threadIdx.x == 0) m.status[4] = sc.total;
  }
  const long long N = (list.v ? sc.total : n_static) * (long long)kFinishLanes;
  // the stride is a multiple of kFinishLanes: an element's lanes stay together
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < N;
       t += (long long)gridDim.x * blockDim.x) {
    const int idx = (int)(t / kFinishLanes), slice = (int)(t % kFinishLanes);
    const int e = list.v ? seg_at(list, sc, idx) : idx;
    if constexpr (kInlineRedo) {
      if (exact_finish(m, kp, x, y, z, e, extra_flags, NoRedo{}, slice)) {
        const QpGroup g{e, slice, (int)(threadIdx.x & 63u) - slice};
        exact_element(m, kp, x, y, z, g, extra_flags);
      }
    } else {
      exact_finish(m, kp, x, y, z, e, extra_flags, RedoSink{m.redo, &m.status[5]}, slice);
    }
  }
}

// Lane groups over the redo list (count read on the device; normally empty).
__global__ void __launch_bounds__(kExactBlock) bf_exact_redo_kernel(DevMesh m, KParams kp,
                                                                const double* __restrict__ x,
                                                                const double* __restrict__ y,
                                                                const double* __restrict__ z,
                                                                unsigned char extra_flags) {
  const int count = m.status[5];
  const int P = kp.n_qp, per_wave = 64 / P;
  QpGroup g;
  int slot;
  lane_group(P, threadIdx.x & 63, slot, g.q, g.base);
  if (slot >= per_wave) return;
  const int waves = gridDim.x * (kExactBlock / 64);
  for (int i = (blockIdx.x * (kExactBlock / 64) + (threadIdx.x >> 6)) * per_wave + slot; i < count;
       i += waves * per_wave) {
    g.e = m.redo[i];
    exact_element(m, kp, x, y, z, g, extra_flags);
  }
}

// ---- batches ----
// Solve work item i = (qp rank i / n_xtiles, 64-element tile i % n_xtiles).
__global__ void __launch_bounds__(kExactBlock) bf_batch_exact_solve_kernel(BatchDev b) {
  const int P = b.P;
  const int items = b.n_xtiles * P;
  const int lane = threadIdx.x & 63;
  const int waves = gridDim.x * (kExactBlock / 64);
  for (int i = blockIdx.x * (kExactBlock / 64) + (threadIdx.x >> 6); i < items; i += waves) {
    const int q = qp_of_rank(i / b.n_xtiles, P);
    const Tile tl = b.xtiles[i % b.n_xtiles];
    const CaseDev& cd = b.cases[tl.c];
    if (lane < tl.n && q < cd.kp.n_qp) exact_solve(cd.m, cd.kp, cd.x, cd.y, cd.z, tl.b + lane, q);
  }
}

// Solve over the compacted hand-over list as the single-context form (work counter
// b.rtotal[1]): pairs: b.plist[0 .. ptotal[0]) ((case, qp_pair) entries), else
// b.hlist[0 .. *htotal) ((case, element) entries, every qp below the case's n_qp solved).
template <int K>
__device__ void batch_solve_list_k(const BatchDev& b, int N, bool pairs) {
  constexpr int NG = split_groups(K);
  const HandItems hi{N, b.P, NG, pairs};
  const int items = hi.n_items();
  const int2* list = pairs ? b.plist : b.hlist;
  const int lane = threadIdx.x & 63;
  const int g = lane / K, j = lane % K;
  for (int i = wave_id(); i < items; i = next_item(b.rtotal + 1)) {
    int idx, q;
    if (g >= NG || !hi.at(i, g, idx, q)) continue;
    const int2 v = list[idx];
    const CaseDev& cd = b.cases[v.x];
    const int e = pairs ? pair_elem(v.y) : v.y;
    if (pairs) q = pair_qp(v.y);
    if (q < cd.kp.n_qp) exact_solve_k<K>(cd.m, cd.kp, cd.x, cd.y, cd.z, e, q, j);
  }
}

__global__ void __launch_bounds__(64) bf_batch_exact_solve_list_kernel(BatchDev b, int pairs,
                                                                       int n_waves) {
  const int N = pairs ? b.ptotal[0] : *b.htotal;
  const int N_mid = pairs ? b.ptotal[1] : 0;
  switch (split_factor(N, N_mid, b.P, pairs != 0, n_waves, false)) {
    case 8: batch_solve_list_k<8>(b, N, pairs != 0); break;
    case 6: batch_solve_list_k<6>(b, N, pairs != 0); break;
    case 4: batch_solve_list_k<4>(b, N, pairs != 0); break;
    case 2: batch_solve_list_k<2>(b, N, pairs != 0); break;
    default: batch_solve_list_k<1>(b, N, pairs != 0); break;
  }
}

struct BatchRedoSink {
  int2* list;
  int* count;
  int c;
  __device__ void operator()(int e) const { list[atomicAdd(count, 1)] = make_int2(c, e); }
};

// Finish over the element tiles (list = 0) or the hand-over list (list = 1).
template <bool kInlineRedo>
__global__ void __launch_bounds__(256) bf_batch_exact_finish_kernel(BatchDev b, int from_list,
                                                                    unsigned char extra_flags) {
  // the stride is a multiple of kFinishLanes: an element's lanes stay together
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long t0 = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (from_list) {
    const long long N = *b.htotal * (long long)kFinishLanes;
    for (long long t = t0; t < N; t += stride) {
      const int idx = (int)(t / kFinishLanes), slice = (int)(t % kFinishLanes);
      const int2 v = b.hlist[idx];
      const CaseDev& cd = b.cases[v.x];
      if constexpr (kInlineRedo) {  // as bf_exact_finish_kernel
        if (exact_finish(cd.m, cd.kp, cd.x, cd.y, cd.z, v.y, extra_flags, NoRedo{}, slice)) {
          const QpGroup g{v.y, slice, (int)(threadIdx.x & 63u) - slice};
          exact_element(cd.m, cd.kp, cd.x, cd.y, cd.z, g, extra_flags);
        }
      } else {
        exact_finish(cd.m, cd.kp, cd.x, cd.y, cd.z, v.y, extra_flags,
                     BatchRedoSink{b.rlist, b.rtotal, v.x}, slice);
      }
    }
    return;
  }
  const long long total = (long long)b.n_xtiles * 64 * kFinishLanes;
  for (long long i = t0; i < total; i += stride) {
    const long long entry = i / kFinishLanes;
    const int slice = (int)(i % kFinishLanes);
    const int t = (int)(entry >> 6), k = (int)(entry & 63);
    const Tile tl = b.xtiles[t];
    if (k >= tl.n) continue;
    const CaseDev& cd = b.cases[tl.c];
    exact_finish(cd.m, cd.kp, cd.x, cd.y, cd.z, tl.b + k, extra_flags,
                 BatchRedoSink{b.rlist, b.rtotal, tl.c}, slice);
  }
}

__global__ void __launch_bounds__(kExactBlock) bf_batch_exact_redo_kernel(BatchDev b,
                                                                      unsigned char extra_flags) {
  const int count = *b.rtotal;
  const int P = b.P, per_wave = 64 / P;
  QpGroup g;
  int slot;
  lane_group(P, threadIdx.x & 63, slot, g.q, g.base);
  if (slot >= per_wave) return;
  const int waves = gridDim.x * (kExactBlock / 64);
  for (int i = (blockIdx.x * (kExactBlock / 64) + (threadIdx.x >> 6)) * per_wave + slot; i < count;
       i += waves * per_wave) {
    const int2 v = b.rlist[i];
    const CaseDev& cd = b.cases[v.x];
    g.e = v.y;
    exact_element(cd.m, cd.kp, cd.x, cd.y, cd.z, g, extra_flags);
  }
}

hipError_t launch_batch_exact(const BatchDev& b, int n_cu, hipStream_t s) {
  if (b.n_xtiles <= 0) return hipSuccess;
  const int items = b.n_xtiles * b.P;
  hipLaunchKernelGGL(bf_batch_exact_solve_kernel, dim3((items + 1) / 2), dim3(kExactBlock), 0, s, b);
  const long long total = (long long)b.n_xtiles * 64 * kFinishLanes;
  const int fblocks = (int)std::min<long long>((total + 255) / 256, 8LL * n_cu);
  hipLaunchKernelGGL(bf_batch_exact_finish_kernel<false>, dim3(fblocks), dim3(256), 0, s, b, 0,
                     (unsigned char)0);
  hipLaunchKernelGGL(bf_batch_exact_redo_kernel, dim3(2 * n_cu), dim3(kExactBlock), 0, s, b,
                     (unsigned char)0);
  return hipGetLastError();
}

hipError_t launch_batch_exact_list(const BatchDev& b, int n_cu, bool pairs, hipStream_t s) {
  if (b.n_etiles <= 0) return hipSuccess;
  // persistent: one wave per SIMD at the solve's occupancy of one
  hipLaunchKernelGGL(bf_batch_exact_solve_list_kernel, dim3(4 * n_cu), dim3(64), 0, s, b,
                     pairs ? 1 : 0, 4 * n_cu);
  // 64-lane blocks: a short hand-over list still spreads over the CUs
  if (b.P <= kFinishLanes) {
    hipLaunchKernelGGL(bf_batch_exact_finish_kernel<true>, dim3(8 * n_cu), dim3(64), 0, s, b, 1,
                       (unsigned char)kStatFallback);
  } else {
    hipLaunchKernelGGL(bf_batch_exact_finish_kernel<false>, dim3(8 * n_cu), dim3(64), 0, s, b,
                       1, (unsigned char)kStatFallback);
    hipLaunchKernelGGL(bf_batch_exact_redo_kernel, dim3(2 * n_cu), dim3(kExactBlock), 0, s, b,
                       (unsigned char)kStatFallback);
  }
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

// ---- node pass of the single-pass staged path ----
// Final force of node n: a partial node sums its tiles' partials in tile order (after the
// caller's value when accumulating); a complete node keeps what its tile wrote; a node no
// element touches keeps the caller's value or zero.  A node of a handed-over element then
// adds, in ascending element order, the forces of its handed-over elements from the
// element-force buffer (written by the reference-order finish or the FP64 second tier), and
// its touched mark is cleared for the next evaluation.  Ghost rows are zeroed when not
// accumulating.  Returns whether f holds the force (always for real nodes when need_f).
// Loads in flight per node and chunk: partials kNodeChunk; entries of a marked node
// EntChunk (the single-mesh kernels use 24, the entries of a node interior to a regular
// patch mesh, so one chunk; the batch kernel, with many more waves in flight, 8).
constexpr int kNodeChunk = 8;
template <int EntChunk>
__device__ __forceinline__ bool tile_node_final(const TileDev& td, const int* __restrict__ ptr,
                                                const unsigned* __restrict__ ent,
                                                const double* __restrict__ Fe, int E, int n_v,
                                                int accumulate, double* __restrict__ fint,
                                                int n, bool need_f, double (&f)[3]) {
  double* fo = fint + 3 * (size_t)n;
  if (n >= n_v) {
    if (!accumulate) fo[0] = fo[1] = fo[2] = 0.0;
    return false;
  }
  const int p0 = td.pptr[n], p1 = td.pptr[n + 1];
  const bool tch = td.touched[n] != 0;
  const bool comp = td.complete[n] != 0;
  if (p0 == p1 && comp && !tch && !need_f) return false;  // its tile wrote it
  bool write = false;
  if (p0 < p1) {
    f[0] = accumulate ? fo[0] : 0.0;
    f[1] = accumulate ? fo[1] : 0.0;
    f[2] = accumulate ? fo[2] : 0.0;
    // partials in chunks whose loads are all in flight before the first add
    for (int j0 = p0; j0 < p1; j0 += kNodeChunk) {
      double px[kNodeChunk], py[kNodeChunk], pz[kNodeChunk];
#pragma unroll
      for (int k = 0; k < kNodeChunk; k++) {
        px[k] = py[k] = pz[k] = 0.0;
        if (j0 + k < p1) {
          const double* pr = td.partial + 3 * (size_t)(j0 + k);
          px[k] = pr[0];
          py[k] = pr[1];
          pz[k] = pr[2];
        }
      }
#pragma unroll
      for (int k = 0; k < kNodeChunk; k++)
        if (j0 + k < p1) {
          f[0] += px[k];
          f[1] += py[k];
          f[2] += pz[k];
        }
    }
    write = true;
  } else if (comp || accumulate) {
    f[0] = fo[0];
    f[1] = fo[1];
    f[2] = fo[2];
  } else {
    f[0] = f[1] = f[2] = 0.0;
    write = true;
  }
  if (tch) {
    td.touched[n] = 0;
    // entries in chunks: the chunk's entries, then their deferred flags, then the rows of
    // the deferred ones are each loaded together; the rows are added in entry order
    const int j1 = ptr[n + 1];
    for (int j0 = ptr[n]; j0 < j1; j0 += EntChunk) {
      unsigned v[EntChunk];
#pragma unroll
      for (int k = 0; k < EntChunk; k++) v[k] = j0 + k < j1 ? ent[j0 + k] : 0u;
      bool d[EntChunk];
#pragma unroll
      for (int k = 0; k < EntChunk; k++) d[k] = j0 + k < j1 && td.deferred[v[k] >> 5] != 0;
      double fx[EntChunk], fy[EntChunk], fz[EntChunk];
#pragma unroll
      for (int k = 0; k < EntChunk; k++) {
        fx[k] = fy[k] = fz[k] = 0.0;
        if (d[k]) {
          const double* fe = Fe + 3 * ((size_t)(v[k] & 31u) * E + (v[k] >> 5));
          fx[k] = fe[0];
          fy[k] = fe[1];
          fz[k] = fe[2];
        }
      }
#pragma unroll
      for (int k = 0; k < EntChunk; k++)
        if (d[k]) {
          f[0] += fx[k];
          f[1] += fy[k];
          f[2] += fz[k];
        }
    }
    write = true;
  }
  if (write) {
    fo[0] = f[0];
    fo[1] = f[1];
    fo[2] = f[2];
  }
  return true;
}

// Explicit damped update of node n from its force f (as step_node).
__device__ __forceinline__ void step_update(int n, const double (&f)[3], double* x, double* y,
                                            double* z, double* v, const double* winv,
                                            const unsigned char* fixed, double dt, double damp) {
  if (fixed[n]) return;
  const double w = winv[n];
  double* vn = v + 3 * (size_t)n;
  const double vx = damp * vn[0] - w * f[0];
  const double vy = damp * vn[1] - w * f[1];
  const double vz = damp * vn[2] - w * f[2];
  vn[0] = vx;
  vn[1] = vy;
  vn[2] = vz;
  x[n] = x[n] + dt * vx;
  y[n] = y[n] + dt * vy;
  z[n] = z[n] + dt * vz;
}

// st (not null): the step's status words, handed over to vis by block 0 (status_handoff).
__global__ void __launch_bounds__(256) bf_tile_fixup_kernel(TileDev td, const int* __restrict__ ptr,
                                                            const unsigned* __restrict__ ent,
                                                            const double* __restrict__ Fe, int E,
                                                            int n_v, int n_nodes, int accumulate,
                                                            double* __restrict__ fint, int* st,
                                                            int* vis) {
  if (st && blockIdx.x == 0) status_handoff(st, vis, (int)threadIdx.x);
  const int n = logical_block() * blockDim.x + threadIdx.x;
  if (n >= n_nodes) return;
  double f[3];
  tile_node_final<24>(td, ptr, ent, Fe, E, n_v, accumulate, fint, n, false, f);
}

__global__ void __launch_bounds__(256) bf_tile_fixup_step_kernel(
    TileDev td, const int* __restrict__ ptr, const unsigned* __restrict__ ent,
    const double* __restrict__ Fe, int E, int n_v, int n_nodes, double* __restrict__ fint,
    double* x, double* y, double* z, double* v, const double* winv, const unsigned char* fixed,
    double dt, double damp, int* st, int* vis) {
  if (st && blockIdx.x == 0) status_handoff(st, vis, (int)threadIdx.x);
  const int n = logical_block() * blockDim.x + threadIdx.x;
  if (n >= n_nodes) return;
  double f[3];
  if (tile_node_final<24>(td, ptr, ent, Fe, E, n_v, 0, fint, n, true, f))
    step_update(n, f, x, y, z, v, winv, fixed, dt, damp);
}

__global__ void __launch_bounds__(256) bf_batch_tile_fixup_kernel(BatchDev b, int step) {
  const Tile tl = b.ntiles[logical_block()];
  const CaseDev& cd = b.cases[tl.c];
  const int i = threadIdx.x;
  if (i >= tl.n) return;
  const int n = tl.b + i;
  double f[3];
  const bool adv = step && cd.v;
  if (tile_node_final<8>(cd.td, cd.ptr, cd.ent, cd.m.Fe, cd.kp.E, cd.n_v, 0, cd.fint, n, adv, f) &&
      adv)
    step_update(n, f, cd.x, cd.y, cd.z, cd.v, cd.winv, cd.fixed, cd.dt, cd.damp);
}

hipError_t launch_tile_fixup(const TileDev& td, const int* ptr, const unsigned* ent,
                             const double* Fe, int E, int n_v, int n_nodes, int accumulate,
                             double* fint, bool step, double* x, double* y, double* z, double* v,
                             const double* winv, const unsigned char* fixed, double dt,
                             double damp, int* st, int* vis, hipStream_t s) {
  if (n_nodes <= 0) return hipSuccess;
  const dim3 grid((n_nodes + 255) / 256), block(256);
  if (step)
    hipLaunchKernelGGL(bf_tile_fixup_step_kernel, grid, block, 0, s, td, ptr, ent, Fe, E, n_v,
                       n_nodes, fint, x, y, z, v, winv, fixed, dt, damp, st, vis);
  else
    hipLaunchKernelGGL(bf_tile_fixup_kernel, grid, block, 0, s, td, ptr, ent, Fe, E, n_v, n_nodes,
                       accumulate, fint, st, vis);
  return hipGetLastError();
}

hipError_t launch_batch_tile_fixup(const BatchDev& b, bool step, hipStream_t s) {
  if (b.n_ntiles <= 0) return hipSuccess;
  hipLaunchKernelGGL(bf_batch_tile_fixup_kernel, dim3(b.n_ntiles), dim3(256), 0, s, b,
                     step ? 1 : 0);
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
                                         const double* y, const double* z, int n_cu, bool wide,
                                         bool pairs, const SegList& solve_list,
                                         const SegList& elem_list, hipEvent_t ev_solve0,
                                         hipEvent_t ev_solve1, hipStream_t s) {
  if (kp.E <= 0) return hipSuccess;
  // persistent grids of one wave per SIMD (the solve's occupancy; the list length is known
  // on the device only).  With `wide` the wide solve kernel runs first and leaves the
  // narrow one its count in status[7] (0 if it took the list); without it the narrow kernel
  // reads the list length itself.
  if (ev_solve0) {
    const hipError_t e = hipEventRecord(ev_solve0, s);
    if (e != hipSuccess) return e;
  }
  if (wide)
    hipLaunchKernelGGL(bf_exact_solve_wide_kernel, dim3(4 * n_cu), dim3(64), 0, s, m, kp, x, y,
                       z, solve_list, pairs ? 1 : 0, &m.status[6], 4 * n_cu, &m.status[7]);
  hipLaunchKernelGGL(bf_exact_solve_queue_kernel, dim3(4 * n_cu), dim3(64), 0, s, m, kp, x, y,
                     z, solve_list, wide ? (const int*)&m.status[7] : (const int*)nullptr,
                     pairs ? 1 : 0, &m.status[6], 4 * n_cu);
  if (ev_solve1) {
    const hipError_t e = hipEventRecord(ev_solve1, s);
    if (e != hipSuccess) return e;
  }
  // 64-lane blocks: a short hand-over list still spreads over the CUs
  if (kp.n_qp <= kFinishLanes) {
    hipLaunchKernelGGL(bf_exact_finish_kernel<true>, dim3(8 * n_cu), dim3(64), 0, s, m, kp, x, y,
                       z, elem_list, 0, (unsigned char)kStatFallback);
  } else {
    hipLaunchKernelGGL(bf_exact_finish_kernel<false>, dim3(8 * n_cu), dim3(64), 0, s, m, kp, x,
                       y, z, elem_list, 0, (unsigned char)kStatFallback);
    hipLaunchKernelGGL(bf_exact_redo_kernel, dim3(2 * n_cu), dim3(kExactBlock), 0, s, m, kp, x,
                       y, z, (unsigned char)kStatFallback);
  }
  return hipGetLastError();
}

hipError_t launch_element_exact(const DevMesh& m, const KParams& kp, const double* x,
                                const double* y, const double* z, int n_cu, hipStream_t s) {
  if (kp.E <= 0) return hipSuccess;
  const int items = ((kp.E + 63) / 64) * kp.n_qp;
  hipLaunchKernelGGL(bf_exact_solve_kernel, dim3((items + 1) / 2), dim3(kExactBlock), 0, s, m,
                     kp, x, y, z, (const int*)nullptr, (const int*)nullptr, kp.E);
  hipLaunchKernelGGL(bf_exact_finish_kernel<false>,
                     dim3((unsigned)(((long long)kp.E * kFinishLanes + 255) / 256)), dim3(256), 0,
                     s, m, kp, x, y, z, SegList{}, kp.E, (unsigned char)0);
  hipLaunchKernelGGL(bf_exact_redo_kernel, dim3(2 * n_cu), dim3(kExactBlock), 0, s, m, kp, x, y,
                     z, (unsigned char)0);
  return hipGetLastError();
}

}  // namespace bf
