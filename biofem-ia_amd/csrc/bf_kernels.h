// bf_kernels.h -- device-side structures shared by the kernel translation units.
#pragma once

#include <climits>
#include <hip/hip_runtime.h>

#include "bf_elem.h"

namespace bf {

struct RampParams {
  double gamma, gamma_T, tau_up, tau_down, apd, pulse_t0, pulse_sigma, t, T_default;
  int type;
};

struct KParams {
  int E, NS, n_qp, iters;
  int act_mode;    // 0: I - gamma f f (cart_fib_act), 1: manufactured diag, 2: isotropic diag
  int has_delay;   // per-element activation delay present
  int fallback;    // staged kernels hand ill-conditioned elements to the plain kernel
  double theta[kMaxQp], w[kMaxQp];
  double h0, mu, K;
  double gamma_t;  // GammaOfTimeDelayed(p, 0) evaluated on the host
  double man_gamma;
  RampParams ramp;
};

struct DevMesh {
  const int* patch;    // [NS][E]
  const int* meta;     // kind | val << 8 | nen << 16
  const int* corner;   // [6][E] cap: nv1..3 then their patch slots (NULL if no cap)
  const double* ref;   // [15][E] a1 a2 a3 a31 a32 of geom0
  const double* dA0;
  const double* fib;   // [3][E]
  const double* gscale;
  const double* delay;
  double* Fe;          // [NS][E][3] element contributions: slot rows of (x, y, z)
  double* c33;         // [n_qp][E]
  int* status;         // [kStatusWords] first error elem, code, first warn elem, warn count,
                       // fallback count, ... (below)
  unsigned char* flags;  // [E] per-element status bits of the last evaluation
  int* fallback;       // [E] elements handed from a staged kernel to the plain kernel
  int* qlist;          // [kMaxQp E] their (element, qp) pairs the reference order solves:
                       // qp_pair(e, q)
  double* qsv;         // [kMaxQp][E][3] reference-order path: S^00, S^11, S^01 per qp
  int* qst;            // [kMaxQp][E] reference-order path: status bits per qp
  int* redo;           // [E] elements whose CC3333 guard tripped: redone with coupled qps
};
// Element tiling of the single-pass staged kernel (host side: bf_tiles.h) as the kernels
// see it, plus its per-evaluation arrays.
struct TileDev {
  int n_tiles, NS, T, lmax;     // tiles, patch slots, elements (threads) per tile, node cap
  int t1_tiles;                 // single context: tiles per region of the tile kernel's list
  const int* elem;              // [n_tiles * T] element of each lane, -1 if none
  const unsigned short* lslot;  // [n_tiles][NS][T] local node of each slot, 0xFFFF unused
  const int* node_ptr;          // [n_tiles + 1] range of each tile's local nodes
  const int* node;              // global node of each local node
  const int* out;               // -1 complete (the tile writes fint), -2 none, >= 0 partial
  const int* cptr;              // [local nodes + 1] range of each local node in cent
  const unsigned short* cent;   // contributions: lane | slot << 8, lane order then slot
  const int* pptr;              // [n_v + 1] partial slots of each real node (consecutive,
                                // tile order)
  const unsigned char* complete;  // [n_v] 1: one tile holds every contribution
  // per evaluation
  double* partial;              // [n_partials][3] tile partial sums
  unsigned char* touched;       // [n_nodes] node of a handed-over element (zero between
                                // evaluations: the node pass clears what the tiles set)
  unsigned char* deferred;      // [E] 1: handed over, its force lands in the element buffer
  int* wlist;                   // [n_tiles * T] hand-over segments, 64 entries per wave
  int* wqlist;                  // [n_tiles * T * kMaxQp] (element, qp) pair segments,
                                // 64 entries per wave and qp (pair_segment)
  unsigned long long* wcount;   // [n_tiles * T / 64] packed segment counts (HandSink)
  int* any;                     // batches, FP32-mixed: set to 1 by a tile wave that hands
                                // anything over (the second tier and the compaction skip
                                // their scans of the count words while it is 0)
  // Single context, FP32-mixed: the tile kernel lists its handed-over elements, the second
  // tier's input, in a SegList of kRegions regions (tiles [j t1_tiles, (j + 1) t1_tiles)
  // append to region j of list_e = wlist, t1_tiles * T entries each, counter
  // list_count[kCntStride j]); FP64: one list (below).  `single` is 0 in batches, whose
  // FP32-mixed tiles keep per-wave segments (wlist, wqlist, wcount) and whose FP64 tiles
  // append to one (case, element) list.
  int single;
  // The FP64 path lists its handed-over elements directly in the hand-over list, at offsets
  // from one device-scope atomic per wave on *list_count (the list's order does not matter:
  // every listed element is solved and finished on its own): list_e (single context,
  // elements) or list_ce (batch, (list_case, element)).
  int* list_count;
  int* list_e;
  int2* list_ce;
  int list_case;
};

// status[4]: elements handed over (m.fallback)
// status[5]: number of entries in redo
// status[6]: work counter of the hand-over solve
// status[7]: the narrow hand-over solve's pair count, when the wide solve runs first
// status[8]: (element, qp) pairs handed over (m.qlist, ordered by qp_of_rank)
// status[9]: how many of them, at the end of the list, are mid-surface pairs (has_mid_qp)
constexpr int kStatusWords = 16;
// Counters of the single context's segmented hand-over lists (SegList), after the status
// words in the same allocation, 256 bytes apart: the tile kernel's regions (the FP64
// hand-over list, or the FP32-mixed second tier's input), the second tier's element regions,
// and its pair classes (qp ranks).  Zeroed by bf_status_reset.
constexpr int kRegions = 8;
constexpr int kCntStride = 64;
constexpr int kT1Base = 64;
constexpr int kT2Base = kT1Base + kRegions * kCntStride;
constexpr int kClassBase = kT2Base + kRegions * kCntStride;
constexpr int kStatusAlloc = kClassBase + kMaxQp * kCntStride;

// The status words and SegList counters as a step starts: first error / warning elements
// INT_MAX, everything else 0.  The staged paths' node pass hands them over (status_handoff):
// block 0 copies the step's words into a second area, which the host reads, and resets them
// for the next step, so that step needs no reset launch (~4 us with its gap); each thread
// copies, then resets, words of its own.
__device__ __forceinline__ void status_handoff(int* st, int* vis, int t) {
  if (t < kStatusWords) {
    vis[t] = st[t];
    st[t] = (t == 0 || t == 2) ? INT_MAX : 0;
  } else if (t < kStatusWords + 2 * kRegions + kMaxQp) {
    const int o = kT1Base + kCntStride * (t - kStatusWords);
    vis[o] = st[o];
    st[o] = 0;
  }
}

// A list kept in n <= kRegions regions: region j holds cnt[kCntStride j] entries at
// v + j cap, and the list is the regions read in order.  Its writers take offsets in a
// region from one atomic per wave on the region's counter: a single counter for a whole grid
// queues those atomics (1 536 waves on one address: +8 us).  v == nullptr: no list.
struct SegList {
  const int* v;
  const int* cnt;
  int cap, n;
};
struct SegCount {
  int c[kRegions];
  int total;
};
__device__ __forceinline__ SegCount seg_count(const SegList& L) {
  SegCount s;
  s.total = 0;
#pragma unroll
  for (int j = 0; j < kRegions; j++) {
    s.c[j] = j < L.n ? L.cnt[kCntStride * j] : 0;
    s.total += s.c[j];
  }
  return s;
}
// Entry i < s.total.
__device__ __forceinline__ int seg_at(const SegList& L, const SegCount& s, int i) {
  int j = 0;
#pragma unroll
  for (int r = 0; r + 1 < kRegions; r++)
    if (j == r && i >= s.c[r]) {
      i -= s.c[r];
      j = r + 1;
    }
  return L.v[(size_t)j * L.cap + i];
}

// An (element, qp) pair of the hand-over solve.
__host__ __device__ inline int qp_pair(int e, int q) { return (e << 3) | q; }
__host__ __device__ inline int pair_elem(int v) { return v >> 3; }
__host__ __device__ inline int pair_qp(int v) { return v & 7; }

// GammaOfTimeDelayed (src/active_strain.c:1156-1179) for a per-element delay.
__device__ __host__ inline double ramp_gamma(const RampParams& p, double delay) {
  if (p.type == 0) return p.gamma;
  const double t = p.t - delay;
  if (p.type == 3) {
    const double tau_up = (p.tau_up > 0.0) ? p.tau_up : 0.02;
    const double tau_down = (p.tau_down > 0.0) ? p.tau_down : 0.08;
    const double apd = (p.apd > 0.0) ? p.apd : 0.30;
    const double t_up = 4.0 * tau_up;
    const double t_down = t_up + apd;
    const double phi_up = 1.0 / (1.0 + exp(-(t - t_up) / tau_up));
    const double phi_down = 1.0 / (1.0 + exp((t - t_down) / tau_down));
    return p.gamma * phi_up * phi_down;
  }
  if (p.type == 4) {
    const double s = (t - p.pulse_t0) / p.pulse_sigma;
    return p.gamma * exp(-0.5 * s * s);
  }
  const double T = (p.gamma_T > 0.0) ? p.gamma_T : p.T_default;
  if (t <= 0.0) return 0.0;
  const double s = (p.type == 1) ? sin(M_PI * t / T) : sin(M_PI * 0.5 * t / T);
  return p.gamma * s * s;
}

// XCD-aware block order.  The dispatcher hands consecutive workgroups to the 8 XCDs in
// turn, and each XCD has its own L2; neighbouring elements share patch nodes and
// neighbouring nodes share element entries, so the logical block index is remapped to give
// each XCD one contiguous range of blocks (a bijection on [0, n)).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block(int b, int n) {
  const int q = n / kXcds, r = n % kXcds;
  const int x = b % kXcds, k = b / kXcds;
  return (x < r) ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}
__device__ __forceinline__ int logical_block() {
  return xcd_block((int)blockIdx.x, (int)gridDim.x);
}

// Lane -> (group slot in the wave, q, first lane of the group) for groups of P adjacent
// lanes (one element per group, one through-thickness point per lane; floor(64 / P)
// groups per wave, the remaining lanes idle).
__device__ __forceinline__ void lane_group(int P, int lane, int& slot, int& q, int& base) {
  slot = lane / P;
  q = lane - slot * P;
  base = slot * P;
}

// Where a staged kernel lists what it hands to the reference-order kernels: every wave owns
// a segment of 64 elements and one segment of 64 (element, qp) pairs per qp (pair_segment),
// and writes their counts packed into one word (hand_count), with no atomics (device-scope
// atomics on one counter serialise across the chip and cost more than the kernel's
// arithmetic).  A compaction pass concatenates the segments, the pairs qp by qp.  push() is
// called by every active lane of a wave that holds at least one element; a lane with `take`
// appends its element e and the pairs of the qps set in qmask (qps below P, wave-uniform).
__host__ __device__ inline size_t pair_segment(int wave, int q) {
  return ((size_t)wave * kMaxQp + q) * 64;
}
// Packed counts of a wave's segments: bits 0-6 elements, bits 7 + 7 q .. 13 + 7 q the pairs
// of qp q (each at most 64).
__host__ __device__ inline unsigned long long hand_count(int n_elem) {
  return (unsigned long long)n_elem;
}
__host__ __device__ inline unsigned long long hand_count_pairs(int q, int n) {
  return (unsigned long long)n << (7 + 7 * q);
}
__host__ __device__ inline int count_elems(unsigned long long w) { return (int)(w & 127u); }
__host__ __device__ inline int count_pairs(unsigned long long w, int q) {
  return (int)((w >> (7 + 7 * q)) & 127u);
}

struct HandSink {
  int* elist;                // [n_waves * 64]
  int* plist;                // [n_waves * kMaxQp * 64]
  unsigned long long* count; // [n_waves]
  int wave;
  int* any;                  // set to 1 when this wave lists anything (may be null)
  __device__ void push(bool take, unsigned qmask, int e, int P) const {
    const unsigned long long act = __ballot(true);
    const int lane = (int)(threadIdx.x & 63u);
    const int leader = __ffsll((long long)act) - 1;
    const unsigned long long below = (1ull << lane) - 1ull;
    const unsigned long long be = __ballot(take);
    if (take) elist[(size_t)wave * 64 + __popcll(be & below)] = e;
    unsigned long long word = hand_count(__popcll(be));
    for (int q = 0; q < P; q++) {
      const bool tq = take && ((qmask >> q) & 1u) != 0u;
      const unsigned long long bal = __ballot(tq);
      if (tq) plist[pair_segment(wave, q) + __popcll(bal & below)] = qp_pair(e, q);
      word |= hand_count_pairs(q, __popcll(bal));
    }
    if (lane == leader) {
      count[wave] = word;
      if (word != 0ull && any) *any = 1;  // every writer stores the same value
    }
  }
};

// Appends e of every lane with `take` to the FP64 path's hand-over list (TileDev list_*):
// one atomic per wave that lists anything, in two calls by every active lane of the wave,
// so that other work can run while the atomic is in flight: hand_list_reserve issues it
// (bal = __ballot(take)) and returns the offset in the wave's first taking lane;
// hand_list_write spreads the offset and writes the entries.
// The counter's address goes through an empty asm statement so that the compiler sees a
// per-lane address: its rewrite of a uniform-address atomic into a wave reduction reads the
// returned value right away, which would wait for the atomic here.
__device__ inline int hand_list_reserve(int* count, unsigned long long bal) {
  int base = 0;
  const int lane = (int)(threadIdx.x & 63u);
  if (bal != 0ull && lane == __ffsll((long long)bal) - 1) {
    int zero = 0;
    asm volatile("" : "+v"(zero));
    base = atomicAdd(count + zero, __popcll(bal));
  }
  return base;
}
__device__ inline void hand_list_write(const TileDev& td, int* list_e, unsigned long long bal,
                                       int base, bool take, int e) {
  if (bal == 0ull) return;
  const int lane = (int)(threadIdx.x & 63u);
  base = __shfl(base, __ffsll((long long)bal) - 1);
  if (!take) return;
  const int at = base + __popcll(bal & ((1ull << lane) - 1ull));
  if (td.list_ce) td.list_ce[at] = make_int2(td.list_case, e);
  else list_e[at] = e;
}

// The form of a wave that works through one segment in passes of floor(64 / P) elements
// (the FP32-mixed path's second tier) and appends to a segment; `pushed` (wave-uniform)
// counts what earlier passes appended.  Every lane active in a pass was active in all
// earlier passes of the segment, so its copy of `pushed` is current.
struct TileSink {
  int* list;    // tile segment
  int* pushed;
  __device__ void push(bool take, int e) const {
    const unsigned long long mask = __ballot(take);
    const int lane = (int)(threadIdx.x & 63u);
    if (take) list[*pushed + __popcll(mask & ((1ull << lane) - 1ull))] = e;
    *pushed += __popcll(mask);
  }
};

// Hand-over solve order of the qps: qp by qp, the mid-surface point (theta = 0, odd n_qp)
// last: on shells with a flat reference its blocks of exactly zero terms are skipped and its
// solve costs a fraction of the others', so the expensive qps start first and the cheap
// ones fill the tail.
__host__ __device__ inline int qp_of_rank(int r, int P) {
  const int mid = P >> 1;
  if (!(P & 1)) return r;
  return r < mid ? r : (r == P - 1 ? mid : r + 1);
}
__host__ __device__ inline bool has_mid_qp(int P) { return P > 1 && (P & 1); }

// ---- batches: many meshes (cases) evaluated by the same launches ----
// One case as the batch kernels see it: its mesh, options and resident state.
struct CaseDev {
  DevMesh m;
  KParams kp;
  double *x, *y, *z;      // resident coordinates (the ghost phase writes ghost rows)
  double* fint;           // resident interleaved forces [3 * n_nodes]
  const int* ptr;         // node -> range of its (element, slot) entries in ent
  const unsigned* ent;    // element << 5 | slot, ascending per node
  const int* rules;       // ghost rules [3 * n_ghost]
  int n_v, n_nodes, n_ghost, pad;
  TileDev td;             // tiling of the single-pass staged kernels
  // time-stepping state (StepState); v == nullptr: force evaluation only
  double* v;              // [3 * n_nodes] velocities
  const double* winv;     // [n_nodes] dt / lumped mass
  const unsigned char* fixed;  // [n_nodes] 1 = clamped (never moves)
  double dt, damp;
};
// A contiguous range of one case's elements, nodes or ghosts.
struct Tile {
  int c, b, n, pad;
};
struct BatchDev {
  const CaseDev* cases;
  int n_cases;
  int P;                  // lanes per element group (largest n_qp of the batch)
  const Tile* etiles;     // element tiles of tile_cap elements
  int n_etiles;
  int tile_cap;
  int persistent;         // 1: waves stride over the tiles; 0: one tile per wave
  const Tile* ntiles;     // node tiles (gather)
  int n_ntiles;
  const Tile* gtiles;     // ghost tiles
  int n_gtiles;
  const Tile* xtiles;     // element tiles of 64 for the reference-order path
  int n_xtiles;
  const Tile* ttiles;     // element tiles of the single-pass staged kernel: (case, tile)
  int n_ttiles;
  int* tlist;             // [n_etiles * tile_cap] hand-over segments
  unsigned long long* tcount;  // [n_etiles] packed segment counts (HandSink)
  int2* hlist;            // compacted hand-over list: (case, element)
  int* htotal;
  int* tqlist;            // [n_etiles * kMaxQp * 64] (element, qp) pair segments
  int2* plist;            // compacted pair list: (case, qp_pair(element, qp)), by qp rank
  int* ptotal;            // [3]: pairs, mid-surface pairs at the end of the list, and the
                          // flag TileDev::any of the step
  int2* rlist;            // elements whose CC3333 guard tripped: (case, element)
  int* rtotal;            // [2]: rlist count, work counter of the hand-over solve
};

__device__ inline V3<double> ld3(const double* __restrict__ x, const double* __restrict__ y,
                                 const double* __restrict__ z, int n) {
  return mk(x[n], y[n], z[n]);
}


// Node-ordered deterministic assembly: fint[3n+c] = (accumulate ? fint : 0) + the node's
// (element, slot) contributions in ascending element order, from the list
// ent[ptr[n] .. ptr[n+1]) built at upload: the summation order of VecSetValues(ADD_VALUES)
// on one rank (active_strain.c:2238-2260).  The element-force buffer holds one row per
// patch slot with the (x, y, z) of every element side by side: a wave's stores for one
// slot cover consecutive elements, and each entry the gather reads is one 24-byte piece.
// Ghost rows (n >= n_v) are zeroed when not accumulating, as after the caller's
// VecSet(Fint, 0).
__device__ __forceinline__ void gather_node(const int* __restrict__ ptr,
                                            const unsigned* __restrict__ ent,
                                            const double* __restrict__ Fe, int E, int n_v,
                                            int accumulate, double* __restrict__ fint, int n) {
  if (n >= n_v) {
    if (!accumulate) {
      fint[3 * (size_t)n + 0] = 0.0;
      fint[3 * (size_t)n + 1] = 0.0;
      fint[3 * (size_t)n + 2] = 0.0;
    }
    return;
  }
  double ax = 0.0, ay = 0.0, az = 0.0;
  if (accumulate) {
    ax = fint[3 * (size_t)n + 0];
    ay = fint[3 * (size_t)n + 1];
    az = fint[3 * (size_t)n + 2];
  }
  // entries are read kGatherChunk at a time (their loads overlap), then added in order; 16
  // holds every entry of a regular interior node (12) and of valence-7 nodes in one round
  // trip (measured: the 64-membrane batch gather 222 -> 124 us against 8)
  constexpr int kGatherChunk = 16;
  const int b = ptr[n], en = ptr[n + 1];
  for (int j0 = b; j0 < en; j0 += kGatherChunk) {
    double vx[kGatherChunk], vy[kGatherChunk], vz[kGatherChunk];
#pragma unroll
    for (int k = 0; k < kGatherChunk; k++) {
      if (j0 + k < en) {
        const unsigned v = ent[j0 + k];
        const double* f = Fe + 3 * ((size_t)(v & 31u) * E + (v >> 5));
        vx[k] = f[0];
        vy[k] = f[1];
        vz[k] = f[2];
      }
    }
#pragma unroll
    for (int k = 0; k < kGatherChunk; k++) {
      if (j0 + k < en) {
        ax += vx[k];
        ay += vy[k];
        az += vz[k];
      }
    }
  }
  fint[3 * (size_t)n + 0] = ax;
  fint[3 * (size_t)n + 1] = ay;
  fint[3 * (size_t)n + 2] = az;
}

// Slot a of element e in the element-force buffer.
__device__ __forceinline__ void put_force(double* __restrict__ Fe, int e, int E, int a,
                                          const V3<double>& f) {
  double* d = Fe + 3 * ((size_t)a * E + e);
  d[0] = f.x;
  d[1] = f.y;
  d[2] = f.z;
}

// Loop box-spline derivatives at the element centre, verbatim 4-decimal values of
// src/active_strain.c:69-88.
static __constant__ double c_Na[2][12] = {
    {-0.0247, -0.0309, 0.0000, -0.4815, -0.1852, 0.0247, 0.4815, 0.0000, -0.0062, 0.0309,
     0.1852, 0.0062},
    {-0.0309, -0.0247, -0.1852, -0.4815, 0.0000, -0.0062, 0.0000, 0.4815, 0.0247, 0.0062,
     0.1852, 0.0309}};
static __constant__ double c_Nab[3][12] = {
    {0.1111, 0.2222, -0.2222, -0.2222, 0.4444, 0.1111, -0.2222, -0.8889, 0.0000, 0.2222,
     0.4444, 0.0000},
    {0.2222, 0.1111, 0.4444, -0.2222, -0.2222, 0.0000, -0.8889, -0.2222, 0.1111, 0.0000,
     0.4444, 0.2222},
    {0.1667, 0.1667, -0.1111, 0.2222, -0.1111, -0.0556, -0.4444, -0.4444, -0.0556, 0.0556,
     0.5556, 0.0556}};
// IN* rows per valence (filled from bf_loop_tables at context creation): [v][row][slot].
// Each kernel translation unit holds its own copy (no relocatable device code).
static __constant__ double c_IN[13][5][kMaxNen];


// Launchers of the kernels compiled without floating-point contraction (bf_exact.hip).
hipError_t upload_in_tables_exact(const double* tab, size_t bytes);
hipError_t launch_geom0(const DevMesh& m, int E, const double* x, const double* y,
                        const double* z, double* ref, hipStream_t s);
// The reference-order path: a qp-major solve kernel (every wave one through-thickness point
// of 64 elements), a per-element force kernel, and a redo of the elements whose CC3333
// guard tripped with the qps coupled.  status[5] must be zero on entry.
hipError_t launch_element_exact(const DevMesh& m, const KParams& kp, const double* x,
                                const double* y, const double* z, int n_cu, hipStream_t s);
// The hand-over solve splits each (element, qp) of its list over K lanes of a wave, K from
// the list length (largest first): 60, 30, 15 (the wide solve kernel), 8, 6, 4, 2, 1.  The
// list holds elements, every qp of each solved (qp by qp in qp_of_rank order), or (element,
// qp) pairs in that order (`pairs`).  A split fits when its work items fill at most one
// round of the n_waves persistent waves, a wave holding split_groups(K) groups of K lanes
// and an item one group's worth of consecutive entries of one qp; the mid-surface items
// (has_mid_qp; for a pair list its last N_mid entries) count at 0.4 of one (measured on
// flat reference shells, where their zero blocks are skipped; the waves the expensive items
// leave free absorb them).
constexpr int kWideSplitMin = 15;
__host__ __device__ constexpr int split_groups(int K) { return 64 / K; }
__host__ __device__ inline bool split_fits(int N, int N_mid, int P, bool pairs, int K,
                                           int n_waves) {
  const long long G = split_groups(K);
  long long ce, cm;
  if (pairs) {
    ce = (N - N_mid + G - 1) / G;
    cm = (N_mid + G - 1) / G;
  } else {
    const long long c = (N + G - 1) / G;
    const int mid = has_mid_qp(P) ? 1 : 0;
    ce = (P - mid) * c;
    cm = mid * c;
  }
  return 10 * ce + 4 * cm <= 10LL * n_waves;
}

// Re-evaluates, in the reference's operation order, what a staged kernel handed over and
// forms the forces of the elements of elem_list (in m.fallback): with `pairs` the (element,
// qp) pairs of solve_list (the classes of m.qlist), else every qp of the elements of
// solve_list (= elem_list) (persistent grids of 4 n_cu waves, counts read on the device and
// stored in status[4], and with `pairs` status[8], status[9]; the solve takes its work items
// from status[6], which must be zero on entry).  `wide` also
// launches the solve kernel for short lists (split of kWideSplitMin lanes or more); without
// it the narrow kernel takes every list -- the sums, and so the results, are the same.
// ev_solve0 / ev_solve1 (may be null) are recorded on s right before and after the solve
// kernels.
hipError_t launch_element_exact_fallback(const DevMesh& m, const KParams& kp, const double* x,
                                         const double* y, const double* z, int n_cu, bool wide,
                                         bool pairs, const SegList& solve_list,
                                         const SegList& elem_list, hipEvent_t ev_solve0,
                                         hipEvent_t ev_solve1, hipStream_t s);
// Batch forms: the reference-order path over the element tiles (plain FP64 batches) and
// over the compacted hand-over list b.hlist[0 .. *b.htotal).  b.rtotal[0..1] must be zero
// on entry (the prologue clears them).
hipError_t launch_batch_exact(const BatchDev& b, int n_cu, hipStream_t s);
// Node phase of a time step (one context): gather + explicit damped update, FP64 without
// contraction.  Batch form: one block per node tile of every case.
hipError_t launch_gather_step(const int* ptr, const unsigned* ent, const double* Fe, int E,
                              int n_v, int n_nodes, double* fint, double* x, double* y,
                              double* z, double* v, const double* winv,
                              const unsigned char* fixed, double dt, double damp, hipStream_t s);
hipError_t launch_batch_gather_step(const BatchDev& b, hipStream_t s);
hipError_t launch_batch_exact_list(const BatchDev& b, int n_cu, bool pairs, hipStream_t s);
// Node pass of the single-pass staged path: each node's force from its tile partial sums
// (tile order) or the value its tile wrote, then the forces of handed-over elements
// (element-force buffer, ascending element order) for nodes that have any; step != 0 also
// applies the explicit damped update (as launch_gather_step); st (may be null): the
// step's status words, handed over to vis (status_handoff).  Batch form over node tiles.
hipError_t launch_tile_fixup(const TileDev& td, const int* ptr, const unsigned* ent,
                             const double* Fe, int E, int n_v, int n_nodes, int accumulate,
                             double* fint, bool step, double* x, double* y, double* z, double* v,
                             const double* winv, const unsigned char* fixed, double dt,
                             double damp, int* st, int* vis, hipStream_t s);
hipError_t launch_batch_tile_fixup(const BatchDev& b, bool step, hipStream_t s);

}  // namespace bf
