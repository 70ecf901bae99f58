// bf_tiles.h -- host-side element tiling of the single-pass staged force kernel.
//
// The staged paths assemble nodal forces without an element-force buffer: a workgroup
// takes one tile of up to T (64, 128 or 256) spatially close elements (Morton order of their
// reference patch centroids), stages the coordinates of every node the tile's patches touch
// in LDS, evaluates one element per lane, and adds the element's nodal forces into LDS node
// accumulators slot by slot.  Within one slot the lanes are split into phases such that no
// two lanes of a phase address the same node, with a barrier between phases, so every node
// sum is formed in the fixed (slot, phase, lane) order: deterministic, without atomics.
// A node whose contributions all lie in one tile is written to the force vector by that
// tile; the others ("partial" nodes) get one partial sum per tile, which a node pass adds in
// tile order.
#pragma once

#include <cstdint>
#include <vector>

namespace bf {

constexpr int kTileElemsDefault = 128;  // elements (lanes, threads) per tile: two waves
constexpr int kTilePhases = 16;         // most phases of one slot (4 bits)
// Most local nodes of a tile of T elements (LDS: 48 B each; below 4096).
constexpr int tile_node_cap(int T) { return 2 * T + T / 2 + 64; }
constexpr uint16_t kTileIdle = 0xFFFFu;  // lane/slot without a contribution

// Local slot entry: bits 0..11 local node, bits 12..15 phase; kTileIdle when unused.
struct TileHost {
  int n_tiles = 0;
  int NS = 0;
  int T = 0;                      // elements per tile
  std::vector<int> elem;          // [n_tiles * T] element of each lane, -1 if none
  std::vector<uint16_t> lslot;    // [n_tiles][NS][T] local node | phase << 12
  std::vector<uint8_t> nphase;    // [n_tiles][NS] phases of each slot (0 if unused)
  std::vector<int> node_ptr;      // [n_tiles + 1] range of each tile in node / out
  std::vector<int> node;          // global node of each local node
  std::vector<int> out;           // per local node: -1 complete (written by the tile),
                                  // -2 no output (ghost node), >= 0 partial slot
  int n_partials = 0;
  std::vector<int> pptr;          // [n_v + 1] partial slots of each real node, tile order
  std::vector<int> plist;         // [n_partials]
  std::vector<uint8_t> complete;  // [n_v] 1: one tile holds every contribution
};

// Builds the tiling of E elements into tiles of at most T elements (T a multiple of 64)
// and tile_node_cap(T) nodes.  patch: [NS][E] node of each slot (slot-major; a slot below
// nen holds a valid node index, or -1 when it adds no force: the slots of an LV apex-cap
// element other than its three corners), nen: [E] slots in use, x0/y0/z0: reference
// coordinates [n_nodes].  Returns 0, or -1 if a slot would need more than kTilePhases phases.
int build_tiles(int E, int n_v, int n_nodes, int NS, const int* patch, const int* nen,
                const double* x0, const double* y0, const double* z0, int T, TileHost& out);

}  // namespace bf
