// bf_topology.cpp -- host-side mesh topology for the force path, O(E log E).
//
// These build the inputs the force path reads: the vertex classification (IrrVer,
// src/bending.c:1188-1267), the 16-slot Loop patch table (Patch, src/bending.c:1270-1667),
// the closed-ring ghost layer (GlobalGhostInit, src/bending.c:1670-1760) and the reference
// areas (AreaNormal, src/io.c:1646-1669).  The reference searches every triangle for every
// query (O(E^2)); here each query visits only the triangles incident to its edge, in
// ascending triangle index, which reproduces the reference's "last match in index order"
// result exactly (a triangle can only match a query if it holds both query nodes).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/biofem_hip.h"

namespace {

constexpr int kSent = 1000000;
constexpr double kPi = 3.14159265358979323846264338327950288;

struct EdgeIndex {
  std::vector<uint64_t> key;  // sorted (min,max) keys, one per (edge, triangle)
  std::vector<int> tri;       // triangle of each key, ascending within a key
  std::vector<int> node_ptr, node_tri;  // node -> incident triangles (ascending)

  static uint64_t mk(int a, int b) {
    uint32_t lo = (uint32_t)std::min(a, b), hi = (uint32_t)std::max(a, b);
    return ((uint64_t)lo << 32) | hi;
  }
  void build(int n_tri, int n_nodes, const int* nv1, const int* nv2, const int* nv3) {
    std::vector<std::pair<uint64_t, int>> kv;
    kv.reserve((size_t)3 * n_tri);
    node_ptr.assign((size_t)n_nodes + 1, 0);
    for (int t = 0; t < n_tri; t++) {
      const int v[3] = {nv1[t], nv2[t], nv3[t]};
      for (int k = 0; k < 3; k++) {
        int a = v[k], b = v[(k + 1) % 3];
        if (a != b) kv.emplace_back(mk(a, b), t);
        if (v[k] >= 0 && v[k] < n_nodes) node_ptr[v[k] + 1]++;
      }
    }
    std::sort(kv.begin(), kv.end());
    kv.erase(std::unique(kv.begin(), kv.end()), kv.end());
    key.resize(kv.size());
    tri.resize(kv.size());
    for (size_t i = 0; i < kv.size(); i++) {
      key[i] = kv[i].first;
      tri[i] = kv[i].second;
    }
    for (int n = 0; n < n_nodes; n++) node_ptr[n + 1] += node_ptr[n];
    node_tri.assign(node_ptr[n_nodes], -1);
    std::vector<int> fill(node_ptr.begin(), node_ptr.end() - 1);
    for (int t = 0; t < n_tri; t++) {
      const int v[3] = {nv1[t], nv2[t], nv3[t]};
      for (int k = 0; k < 3; k++)
        if (v[k] >= 0 && v[k] < n_nodes) {
          // a triangle listing the same node twice is counted once per listing, as the
          // reference's membership test would match it once
          bool dup = false;
          for (int j = 0; j < k; j++) dup |= (v[j] == v[k]);
          if (!dup) node_tri[fill[v[k]]++] = t;
        }
    }
    for (int n = 0; n < n_nodes; n++) {
      // keep only the filled prefix (duplicates skipped above), ascending
      int* b = node_tri.data() + node_ptr[n];
      int* e = node_tri.data() + fill[n];
      std::sort(b, e);
      for (int* p = e; p < node_tri.data() + node_ptr[n + 1]; p++) *p = -1;
    }
  }
  // visit triangles holding both a and b in ascending index
  template <class F>
  void visit(int a, int b, int n_nodes, F f) const {
    if (a == kSent || b == kSent) return;
    if (a == b) {
      if (a < 0 || a >= n_nodes) return;
      for (int i = node_ptr[a]; i < node_ptr[a + 1] && node_tri[i] >= 0; i++) f(node_tri[i]);
      return;
    }
    const uint64_t k = mk(a, b);
    auto it = std::lower_bound(key.begin(), key.end(), k);
    for (size_t i = (size_t)(it - key.begin()); i < key.size() && key[i] == k; i++) f(tri[i]);
  }
};

// The if / else-if chain of every Patch query (e.g. src/bending.c:1306-1314): take the
// first vertex of triangle t that is none of the exclusions.
inline bool pick(const int* nv1, const int* nv2, const int* nv3, int t, int ex0, int ex1,
                 int ex2, int* out) {
  const int n1 = nv1[t], n2 = nv2[t], n3 = nv3[t];
  if (n1 != ex0 && n1 != ex1 && n1 != ex2) { *out = n1; return true; }
  if (n2 != ex0 && n2 != ex1 && n2 != ex2) { *out = n2; return true; }
  if (n3 != ex0 && n3 != ex1 && n3 != ex2) { *out = n3; return true; }
  return false;
}

}  // namespace

extern "C" {

int bf_topo_area0(int n_tri, const int* nv1, const int* nv2, const int* nv3, const double* x,
                  const double* y, const double* z, double* dA) {
  if (n_tri < 0 || !nv1 || !nv2 || !nv3 || !x || !y || !z || !dA) return BF_ERR_ARG;
  for (int e = 0; e < n_tri; e++) {  // src/io.c:1651-1666
    const int n1 = nv1[e], n2 = nv2[e], n3 = nv3[e];
    const double d21x = x[n2] - x[n1], d21y = y[n2] - y[n1], d21z = z[n2] - z[n1];
    const double d31x = x[n3] - x[n1], d31y = y[n3] - y[n1], d31z = z[n3] - z[n1];
    const double cx = d21y * d31z - d31y * d21z;
    const double cy = -d21x * d31z + d31x * d21z;
    const double cz = d21x * d31y - d31x * d21y;
    dA[e] = 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
  }
  return BF_OK;
}

int bf_topo_irrver(int n_v, int n_elmt, int n_elmt_base, const int* nv1, const int* nv2,
                   const int* nv3, int n_bnodes, const int* bnodes, int* ire, int* irv,
                   int* val) {
  if (n_v < 0 || n_elmt < 0 || !nv1 || !nv2 || !nv3 || !ire || !irv || !val) return BF_ERR_ARG;
  for (int e = 0; e < n_elmt; e++) { ire[e] = 0; irv[e] = 0; val[e] = 6; }
  const int lim = n_elmt_base > 0 ? n_elmt_base : n_elmt;
  std::vector<int> ptr((size_t)n_v + 1, 0), lst;
  for (int e = 0; e < lim; e++) {
    const int v[3] = {nv1[e], nv2[e], nv3[e]};
    for (int k = 0; k < 3; k++) {
      bool dup = false;
      for (int j = 0; j < k; j++) dup |= v[j] == v[k];
      if (!dup && v[k] >= 0 && v[k] < n_v) ptr[v[k] + 1]++;
    }
  }
  for (int n = 0; n < n_v; n++) ptr[n + 1] += ptr[n];
  lst.resize(ptr[n_v]);
  std::vector<int> fill(ptr.begin(), ptr.end() - 1);
  for (int e = 0; e < lim; e++) {  // ascending element order, as the reference's elmt list
    const int v[3] = {nv1[e], nv2[e], nv3[e]};
    for (int k = 0; k < 3; k++) {
      bool dup = false;
      for (int j = 0; j < k; j++) dup |= v[j] == v[k];
      if (!dup && v[k] >= 0 && v[k] < n_v) lst[fill[v[k]]++] = e;
    }
  }
  std::vector<char> isb((size_t)n_v, 0), irr((size_t)n_v, 0);
  for (int i = 0; i < n_bnodes; i++)
    if (bnodes[i] >= 0 && bnodes[i] < n_v) isb[bnodes[i]] = 1;
  for (int nc = 0; nc < n_v; nc++) {  // src/bending.c:1212-1252
    const int count = ptr[nc + 1] - ptr[nc];
    const int bcount = isb[nc] ? count + 3 : 0;
    const int vv = bcount == 0 ? count : bcount;
    if (vv != 6) {
      irr[nc] = 1;
      for (int i = ptr[nc]; i < ptr[nc + 1]; i++) { ire[lst[i]] = 1; val[lst[i]] = vv; }
    }
  }
  for (int e = 0; e < n_elmt; e++) {  // :1254-1262: the highest-numbered irregular corner
    int best = -1;
    const int v[3] = {nv1[e], nv2[e], nv3[e]};
    for (int k = 0; k < 3; k++)
      if (v[k] >= 0 && v[k] < n_v && irr[v[k]] && v[k] >= best) { best = v[k]; irv[e] = k + 1; }
  }
  return BF_OK;
}

int bf_topo_patch(int n_v, int n_nodes, int n_elmt, int n_elmt_base, int n_tri, const int* nv1,
                  const int* nv2, const int* nv3, const int* ire, const int* irv,
                  const int* val, int stride, int* patch) {
  if (n_v < 0 || n_elmt < 0 || n_tri < n_elmt || !nv1 || !nv2 || !nv3 || !ire || !irv ||
      !val || !patch || stride < 12 || stride > 64 || n_nodes >= kSent)
    return BF_ERR_ARG;
  EdgeIndex ix;
  ix.build(n_tri, n_nodes, nv1, nv2, nv3);
  for (int ec = 0; ec < n_elmt; ec++) {
    int* P = patch + (size_t)stride * ec;
    const int v = val[ec];
    if (n_elmt_base > 0 && ec >= n_elmt_base) {  // :1281-1294
      for (int i = 0; i < stride; i++) P[i] = kSent;
      P[3] = nv1[ec]; P[6] = nv2[ec]; P[7] = nv3[ec];
      continue;
    }
    if (v < 0 || v + 6 > 64) return BF_ERR_ARG;
    int p[64];
    for (int i = 0; i < v + 6; i++) p[i] = kSent;
    // one query: over triangles holding (a, b), excluding triangle `skip`, last match wins
    auto q = [&](int a, int b, int e0, int e1, int e2, int skip, int slot) {
      ix.visit(a, b, n_nodes, [&](int t) {
        if (t == skip) return;
        pick(nv1, nv2, nv3, t, e0, e1, e2, &p[slot]);
      });
    };
    if (ire[ec] == 0) {  // :1300-1430
      p[3] = nv1[ec]; p[6] = nv2[ec]; p[7] = nv3[ec];
      q(p[3], p[6], p[3], p[6], p[6], ec, 2);
      q(p[3], p[7], p[3], p[7], p[7], ec, 4);
      q(p[6], p[7], p[6], p[7], p[7], ec, 10);
      q(p[3], p[2], p[3], p[2], p[6], -1, 0);
      q(p[3], p[4], p[3], p[4], p[7], -1, 1);
      q(p[6], p[2], p[6], p[2], p[3], -1, 5);
      q(p[6], p[10], p[6], p[10], p[7], -1, 9);
      q(p[7], p[10], p[7], p[10], p[6], -1, 11);
      q(p[7], p[4], p[7], p[4], p[3], -1, 8);
    } else if (ire[ec] == 1) {  // :1432-1660
      if (irv[ec] == 1) { p[0] = nv1[ec]; p[1] = nv2[ec]; p[v] = nv3[ec]; }
      else if (irv[ec] == 2) { p[0] = nv2[ec]; p[1] = nv3[ec]; p[v] = nv1[ec]; }
      else if (irv[ec] == 3) { p[0] = nv3[ec]; p[1] = nv1[ec]; p[v] = nv2[ec]; }
      // Within one scan the reference evaluates its queries per triangle, so a query can
      // see a slot written earlier in the same scan.  Queries of one scan only depend on
      // each other when slot indices alias (v <= 5), so those scans are replayed per
      // triangle in index order, exactly as written.
      auto scan = [&](auto body) {
        for (int t = 0; t < n_tri; t++) body(t);
      };
      auto q1 = [&](int t, int a, int b, int e0, int e1, int e2, bool ok, int slot) {
        int m = 0;
        if (a == nv1[t] || a == nv2[t] || a == nv3[t]) m++;
        if (b == nv1[t] || b == nv2[t] || b == nv3[t]) m++;
        if (m == 2 && ok) pick(nv1, nv2, nv3, t, e0, e1, e2, &p[slot]);
      };
      if (v <= 5) {
        scan([&](int t) {
          const bool ok = t != ec;
          q1(t, p[0], p[1], p[0], p[1], p[1], ok, 2);
          q1(t, p[0], p[v], p[0], p[v], p[v], ok, v - 1);
          q1(t, p[1], p[v], p[1], p[v], p[v], ok, v + 1);
        });
        scan([&](int t) {
          q1(t, p[0], p[2], p[0], p[2], p[1], true, 3);
          q1(t, p[0], p[v - 1], p[0], p[v - 1], p[v], true, v - 2);
          q1(t, p[1], p[2], p[1], p[2], p[0], true, v + 3);
          q1(t, p[1], p[v + 1], p[1], p[v + 1], p[v], true, v + 2);
          q1(t, p[v], p[v + 1], p[v], p[v + 1], p[1], true, v + 4);
          q1(t, p[v], p[v - 1], p[v], p[v - 1], p[0], true, v + 5);
        });
      } else {
        q(p[0], p[1], p[0], p[1], p[1], ec, 2);
        q(p[0], p[v], p[0], p[v], p[v], ec, v - 1);
        q(p[1], p[v], p[1], p[v], p[v], ec, v + 1);
        q(p[0], p[2], p[0], p[2], p[1], -1, 3);
        q(p[0], p[v - 1], p[0], p[v - 1], p[v], -1, v - 2);
        q(p[1], p[2], p[1], p[2], p[0], -1, v + 3);
        q(p[1], p[v + 1], p[1], p[v + 1], p[v], -1, v + 2);
        q(p[v], p[v + 1], p[v], p[v + 1], p[1], -1, v + 4);
        q(p[v], p[v - 1], p[v], p[v - 1], p[0], -1, v + 5);
        // extra valence walk (:1622-1660): two queries per scan that may alias
        for (int k = 0; k < v - 6; k++) {
          if (v - 3 - k == k + 4 || v - 2 - k == k + 4 || 3 + k == v - 3 - k) {
            scan([&](int t) {
              q1(t, p[0], p[3 + k], p[0], p[3 + k], p[2 + k], true, k + 4);
              q1(t, p[0], p[v - 2 - k], p[0], p[v - 2 - k], p[v - 1 - k], true, v - 3 - k);
            });
          } else {
            q(p[0], p[3 + k], p[0], p[3 + k], p[2 + k], -1, k + 4);
            q(p[0], p[v - 2 - k], p[0], p[v - 2 - k], p[v - 1 - k], -1, v - 3 - k);
          }
        }
      }
    }
    // sentinel fix-up (:1662-1680)
    int fix = -1, gfix = -1;
    for (int i = 0; i < v + 6 && i < stride; i++) {
      if (p[i] == kSent) continue;
      if (fix < 0) fix = p[i];
      if (p[i] >= n_v) { gfix = p[i]; break; }
    }
    if (gfix >= 0) fix = gfix;
    if (fix >= 0)
      for (int i = 0; i < v + 6 && i < stride; i++)
        if (p[i] == kSent) p[i] = fix;
    for (int i = 0; i < stride; i++) P[i] = kSent;
    for (int i = 0; i < v + 6 && i < stride; i++) P[i] = p[i];
  }
  return BF_OK;
}

int bf_topo_ghost_ring(int n_v, int n_elmt, int n_elmt_base, int n_edge, const int* n_bnodes,
                       const int* bnodes, int* nv1, int* nv2, int* nv3, int* rules) {
  // GlobalGhostInit (src/bending.c:1674-1756): ghost g of boundary edge (b=bnodes[i],
  // a=bnodes[i+1 mod ring]) sits at x[b] + x[a] - x[c], c the third corner of the real
  // triangle on that edge; rules[3g..] = (b, a, c).  Two ghost triangles per ghost are
  // appended after the n_elmt real ones.
  if (!n_bnodes || !bnodes || !nv1 || !nv2 || !nv3 || !rules) return BF_ERR_ARG;
  int total = 0;
  for (int r = 0; r < n_edge; r++) total += n_bnodes[r];
  EdgeIndex ix;
  ix.build(n_elmt, n_v + total, nv1, nv2, nv3);
  int count = 0, start = 0;
  for (int r = 0; r < n_edge; r++) {
    const int nl = n_bnodes[r];
    const int lim = (r == 0 && n_elmt_base > 0) ? n_elmt_base : n_elmt;
    const int first = count;
    for (int li = 0; li < nl; li++) {
      const int b = bnodes[start + li], a = bnodes[start + (li + 1) % nl];
      int c = -1;
      ix.visit(a, b, n_v + total, [&](int t) {
        if (t >= lim) return;
        if (nv1[t] != a && nv1[t] != b) c = nv1[t];
        else if (nv2[t] != a && nv2[t] != b) c = nv2[t];
        else c = nv3[t];
      });
      if (c < 0) return BF_ERR_ARG;  // boundary edge without a triangle: not a closed ring
      rules[3 * count + 0] = b;
      rules[3 * count + 1] = a;
      rules[3 * count + 2] = c;
      nv2[n_elmt + 2 * count] = b;
      nv1[n_elmt + 2 * count] = a;
      nv3[n_elmt + 2 * count] = n_v + count;
      const int prev = (li == 0) ? (n_v + first + nl - 1) : (n_v + count - 1);
      nv1[n_elmt + 2 * count + 1] = n_v + count;
      nv2[n_elmt + 2 * count + 1] = b;
      nv3[n_elmt + 2 * count + 1] = prev;
      count++;
    }
    start += nl;
  }
  return BF_OK;
}

int bf_loop_tables(int v, double* out) {
  // IN* rows of an irregular Loop patch of valence v (active_strain.c:594-666); out is
  // [5][v+6]: INa0, INa1, INab0, INab1, INab2.
  if (v < 3 || v + 12 > 64) return BF_ERR_ARG;
  static const double Na[2][12] = {
      {-0.0247, -0.0309, 0.0000, -0.4815, -0.1852, 0.0247, 0.4815, 0.0000, -0.0062, 0.0309,
       0.1852, 0.0062},
      {-0.0309, -0.0247, -0.1852, -0.4815, 0.0000, -0.0062, 0.0000, 0.4815, 0.0247, 0.0062,
       0.1852, 0.0309}};
  static const double Nab[3][12] = {
      {0.1111, 0.2222, -0.2222, -0.2222, 0.4444, 0.1111, -0.2222, -0.8889, 0.0000, 0.2222,
       0.4444, 0.0000},
      {0.2222, 0.1111, 0.4444, -0.2222, -0.2222, 0.0000, -0.8889, -0.2222, 0.1111, 0.0000,
       0.4444, 0.2222},
      {0.1667, 0.1667, -0.1111, 0.2222, -0.1111, -0.0556, -0.4444, -0.4444, -0.0556, 0.0556,
       0.5556, 0.0556}};
  const int nen = v + 6, nr = v + 12;
  const double w = (1.0 / (double)v) * (0.625 - std::pow(0.375 + 0.25 * std::cos(2 * kPi / (double)v), 2.0));
  // B1: Loop subdivision mask from the (v+6)-node patch to the (v+12)-node refined patch,
  // B2: picks the 12 refined nodes forming the regular sub-patch (:604-639)
  std::vector<double> B1((size_t)nr * nen, 0.0);
  auto b1 = [&](int i, int j) -> double& { return B1[(size_t)i * nen + j]; };
  int pick12[12] = {v + 9, v + 6, v + 4, v + 1, v + 2, v + 5, v, 1, v + 3, v - 1, 0, 2};
  b1(0, 0) = 1 - v * w;
  for (int j = 0; j < v; j++) b1(0, 1 + j) = w;
  b1(1, 0) = 0.375; b1(1, 1) = 0.375; b1(1, 2) = 0.125; b1(1, v) = 0.125;
  b1(2, 0) = 0.375; b1(2, 1) = 0.125; b1(2, 2) = 0.375; b1(2, 3) = 0.125;
  if (v > 5) { b1(v - 4, 0) = 0.375; b1(v - 4, v - 5) = 0.125; b1(v - 4, v - 4) = 0.375; b1(v - 4, v - 3) = 0.125; }
  if (v > 4) { b1(v - 3, 0) = 0.375; b1(v - 3, v - 4) = 0.125; b1(v - 3, v - 3) = 0.375; b1(v - 3, v - 2) = 0.125; }
  b1(v - 2, 0) = 0.375; b1(v - 2, v - 3) = 0.125; b1(v - 2, v - 2) = 0.375; b1(v - 2, v - 1) = 0.125;
  b1(v - 1, 0) = 0.375; b1(v - 1, v - 2) = 0.125; b1(v - 1, v - 1) = 0.375; b1(v - 1, v) = 0.125;
  b1(v, 0) = 0.375; b1(v, 1) = 0.125; b1(v, v - 1) = 0.125; b1(v, v) = 0.375;
  b1(v + 1, 0) = 0.125; b1(v + 1, 1) = 0.375; b1(v + 1, v) = 0.375; b1(v + 1, v + 1) = 0.125;
  b1(v + 2, 0) = 0.0625; b1(v + 2, 1) = 0.625; b1(v + 2, 2) = 0.0625; b1(v + 2, v) = 0.0625;
  b1(v + 2, v + 1) = 0.0625; b1(v + 2, v + 2) = 0.0625; b1(v + 2, v + 3) = 0.0625;
  b1(v + 3, 0) = 0.125; b1(v + 3, 1) = 0.375; b1(v + 3, 2) = 0.375; b1(v + 3, v + 3) = 0.125;
  b1(v + 4, 0) = 0.0625; b1(v + 4, 1) = 0.0625; b1(v + 4, v - 1) = 0.0625; b1(v + 4, v) = 0.625;
  b1(v + 4, v + 1) = 0.0625; b1(v + 4, v + 4) = 0.0625; b1(v + 4, v + 5) = 0.0625;
  b1(v + 5, 0) = 0.125; b1(v + 5, v - 1) = 0.375; b1(v + 5, v) = 0.375; b1(v + 5, v + 5) = 0.125;
  b1(v + 6, 1) = 0.375; b1(v + 6, v) = 0.125; b1(v + 6, v + 1) = 0.375; b1(v + 6, v + 2) = 0.125;
  b1(v + 7, 1) = 0.375; b1(v + 7, v + 1) = 0.125; b1(v + 7, v + 2) = 0.375; b1(v + 7, v + 3) = 0.125;
  b1(v + 8, 1) = 0.375; b1(v + 8, 2) = 0.125; b1(v + 8, v + 2) = 0.125; b1(v + 8, v + 3) = 0.375;
  b1(v + 9, 1) = 0.125; b1(v + 9, v) = 0.375; b1(v + 9, v + 1) = 0.375; b1(v + 9, v + 4) = 0.125;
  b1(v + 10, v) = 0.375; b1(v + 10, v + 1) = 0.125; b1(v + 10, v + 4) = 0.375; b1(v + 10, v + 5) = 0.125;
  b1(v + 11, v - 1) = 0.125; b1(v + 11, v) = 0.375; b1(v + 11, v + 4) = 0.125; b1(v + 11, v + 5) = 0.375;
  // B3 = B2 B1: row i of B3 is row pick12[i] of B1, accumulated over all v+12 columns of B2
  // (B2 has a single 1 per row; the sum over m visits the zeros too, as at :642-648)
  for (int j = 0; j < nen; j++) {
    double s1 = 0., s2 = 0., s3 = 0., s4 = 0., s5 = 0.;
    for (int i = 0; i < 12; i++) {
      double s = 0.0;
      for (int m = 0; m < nr; m++) s += (m == pick12[i] ? 1.0 : 0.0) * b1(m, j);
      s1 += s * Na[0][i];
      s2 += s * Na[1][i];
      s3 += s * Nab[0][i];
      s4 += s * Nab[1][i];
      s5 += s * Nab[2][i];
    }
    out[0 * nen + j] = -2.0 * s1;
    out[1 * nen + j] = -2.0 * s2;
    out[2 * nen + j] = 4.0 * s3;
    out[3 * nen + j] = 4.0 * s4;
    out[4 * nen + j] = 4.0 * s5;
  }
  return BF_OK;
}

}  // extern "C"
