// bf_tiles.cpp -- element tiling of the single-pass staged force kernel (see bf_tiles.h).
#include "bf_tiles.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

namespace bf {

namespace {

// 10-bit Morton code of a point of the unit cube.
uint32_t spread10(uint32_t v) {
  v &= 0x3FFu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

}  // namespace

int build_tiles(int E, int n_v, int n_nodes, int NS, const int* patch, const int* nen,
                const double* x0, const double* y0, const double* z0, int TE, TileHost& T) {
  T = TileHost();
  T.NS = NS;
  T.T = TE;
  const int kTileElems = TE, kTileNodes = tile_node_cap(TE);
  if (TE != 64 && TE != 128) return -1;
  if (E <= 0) {
    T.node_ptr.assign(1, 0);
    T.cptr.assign(1, 0);
    T.pptr.assign((size_t)std::max(n_v, 0) + 1, 0);
    T.complete.assign((size_t)std::max(n_v, 0), 0);
    return 0;
  }
  // ---- Morton order of the patch centroids ----
  std::vector<double> c((size_t)3 * E);
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int e = 0; e < E; e++) {
    double s[3] = {0.0, 0.0, 0.0};
    int cnt = 0;
    for (int a = 0; a < nen[e]; a++) {
      const int n = patch[(size_t)a * E + e];
      if (n < 0) continue;
      s[0] += x0[n];
      s[1] += y0[n];
      s[2] += z0[n];
      cnt++;
    }
    for (int k = 0; k < 3; k++) {
      c[3 * (size_t)e + k] = s[k] / std::max(cnt, 1);
      lo[k] = std::min(lo[k], c[3 * (size_t)e + k]);
      hi[k] = std::max(hi[k], c[3 * (size_t)e + k]);
    }
  }
  const double span = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-300});
  std::vector<uint32_t> key((size_t)E);
  for (int e = 0; e < E; e++) {
    uint32_t q[3];
    for (int k = 0; k < 3; k++) {
      const double u = (c[3 * (size_t)e + k] - lo[k]) / span;
      q[k] = (uint32_t)std::min(1023.0, std::max(0.0, std::floor(u * 1024.0)));
    }
    key[e] = spread10(q[0]) | (spread10(q[1]) << 1) | (spread10(q[2]) << 2);
  }
  std::vector<int> order((size_t)E);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });

  // ---- tiles: consecutive elements of the order, at most kTileElems and kTileNodes ----
  std::vector<int> loc((size_t)n_nodes, -1);       // local index of a node in the open tile
  std::vector<int> tile_nodes;                     // nodes of the open tile
  std::vector<int> tile_elems;
  std::vector<int> node_tiles((size_t)n_nodes, 0);  // tiles that hold a node
  T.node_ptr.push_back(0);
  T.cptr.push_back(0);
  std::vector<int> cnt;
  auto close_tile = [&]() {
    const int t = T.n_tiles++;
    T.elem.resize((size_t)T.n_tiles * kTileElems, -1);
    T.lslot.resize((size_t)T.n_tiles * NS * kTileElems, kTileIdle);
    for (int l = 0; l < (int)tile_elems.size(); l++) T.elem[(size_t)t * kTileElems + l] = tile_elems[l];
    // slot codes; contributions of each local node in lane order, then slot order
    const int L = (int)tile_nodes.size();
    cnt.assign((size_t)L + 1, 0);
    for (int l = 0; l < (int)tile_elems.size(); l++) {
      const int e = tile_elems[l];
      for (int a = 0; a < nen[e]; a++) {
        const int n = patch[(size_t)a * E + e];
        if (n < 0) continue;
        const int ln = loc[n];
        T.lslot[((size_t)t * NS + a) * kTileElems + l] = (uint16_t)ln;
        cnt[ln + 1]++;
      }
    }
    for (int k = 0; k < L; k++) cnt[k + 1] += cnt[k];
    const size_t c0 = T.cent.size();
    T.cent.resize(c0 + (size_t)cnt[L]);
    for (int k = 0; k < L; k++) T.cptr.push_back((int)c0 + cnt[k + 1]);
    for (int l = 0; l < (int)tile_elems.size(); l++) {
      const int e = tile_elems[l];
      for (int a = 0; a < nen[e]; a++) {
        const int n = patch[(size_t)a * E + e];
        if (n < 0) continue;
        T.cent[c0 + cnt[loc[n]]++] = (uint16_t)(l | (a << 8));
      }
    }
    for (int n : tile_nodes) {
      T.node.push_back(n);
      node_tiles[n]++;
      loc[n] = -1;
    }
    T.node_ptr.push_back((int)T.node.size());
    tile_nodes.clear();
    tile_elems.clear();
  };
  for (int i = 0; i < E; i++) {
    const int e = order[i];
    int fresh = 0;
    for (int a = 0; a < nen[e]; a++) {
      const int n = patch[(size_t)a * E + e];
      if (n < 0) continue;
      if (loc[n] < 0) {
        bool seen = false;  // a node repeated within this element's own patch
        for (int b = 0; b < a; b++) seen = seen || patch[(size_t)b * E + e] == n;
        fresh += seen ? 0 : 1;
      }
    }
    if (!tile_elems.empty() &&
        ((int)tile_elems.size() == kTileElems || (int)tile_nodes.size() + fresh > kTileNodes))
      close_tile();
    for (int a = 0; a < nen[e]; a++) {
      const int n = patch[(size_t)a * E + e];
      if (n < 0) continue;
      if (loc[n] < 0) {
        loc[n] = (int)tile_nodes.size();
        tile_nodes.push_back(n);
      }
    }
    tile_elems.push_back(e);
  }
  close_tile();

  // ---- outputs: complete nodes, partial slots (consecutive per node, tile order) ----
  T.out.assign(T.node.size(), -1);
  T.complete.assign((size_t)n_v, 0);
  T.pptr.assign((size_t)n_v + 1, 0);
  for (int n = 0; n < n_v; n++) {
    if (node_tiles[n] == 1) T.complete[n] = 1;
    if (node_tiles[n] > 1) T.pptr[n + 1] = node_tiles[n];
  }
  for (int n = 0; n < n_v; n++) T.pptr[n + 1] += T.pptr[n];
  T.n_partials = T.pptr[n_v];
  std::vector<int> fill(T.pptr.begin(), T.pptr.end() - 1);
  for (int t = 0; t < T.n_tiles; t++)
    for (int j = T.node_ptr[t]; j < T.node_ptr[t + 1]; j++) {
      const int n = T.node[j];
      if (n >= n_v)
        T.out[j] = -2;
      else if (node_tiles[n] > 1)
        T.out[j] = fill[n]++;
    }
  return 0;
}

}  // namespace bf
