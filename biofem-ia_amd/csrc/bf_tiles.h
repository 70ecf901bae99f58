// bf_tiles.h -- host-side element tiling of the single-pass staged force kernel.
//
// The staged paths assemble nodal forces without an element-force buffer in HBM: a
// workgroup takes one tile of up to T (64 or 128) spatially close elements (Morton order of
// their reference patch centroids), stages the coordinates of every node the tile's patches
// touch in LDS, evaluates the tile's elements, leaves each element's nodal forces per patch
// slot in LDS, and then forms every local node's sum from its (lane, slot) contributions in
// the fixed order of a per-tile list (lane order, then slot order): deterministic, without
// atomics, one barrier.  A node whose contributions all lie in one tile is written to the
// force vector by that tile; the others ("partial" nodes) get one partial sum per tile,
// stored consecutively per node in tile order, which a node pass adds in that order.
#pragma once

#include <cstdint>
#include <vector>

namespace bf {

// Elements per tile unless a batch's tuned launch asks for 128: 64 measured faster on single
// meshes (256 x 192 membrane 54 -> 45 us, 184k-triangle ellipsoid 93 -> 67 us per element
// phase: twice the workgroups, so more than one round of the chip to overlap phases).
constexpr int kTileElemsDefault = 64;
// Most local nodes of a tile of T elements.
constexpr int tile_node_cap(int T) { return 2 * T + T / 2 + 64; }
constexpr uint16_t kTileIdle = 0xFFFFu;  // lane/slot without a contribution

struct TileHost {
  int n_tiles = 0;
  int NS = 0;
  int T = 0;                      // elements per tile
  std::vector<int> elem;          // [n_tiles * T] element of each lane, -1 if none
  std::vector<uint16_t> lslot;    // [n_tiles][NS][T] local node of each slot, kTileIdle if unused
  std::vector<int> node_ptr;      // [n_tiles + 1] range of each tile in node / out / cptr
  std::vector<int> node;          // global node of each local node
  std::vector<int> out;           // per local node: -1 complete (written by the tile),
                                  // -2 no output (ghost node), >= 0 its partial slot
  std::vector<int> cptr;          // [node.size() + 1] contributions of each local node in cent
  std::vector<uint16_t> cent;     // contribution: lane | slot << 8 (lane order, then slot)
  int n_partials = 0;
  std::vector<int> pptr;          // [n_v + 1] partial slots of each real node (consecutive,
                                  // tile order)
  std::vector<uint8_t> complete;  // [n_v] 1: one tile holds every contribution
};

// Builds the tiling of E elements into tiles of at most T elements (64 or 128) and
// tile_node_cap(T) nodes.  patch: [NS][E] node of each slot (slot-major; a slot below nen
// holds a valid node index, or -1 when it adds no force: the slots of an LV apex-cap element
// other than its three corners), nen: [E] slots in use, x0/y0/z0: reference coordinates
// [n_nodes].  Returns 0, or -1 for an argument the tiling cannot hold.
int build_tiles(int E, int n_v, int n_nodes, int NS, const int* patch, const int* nen,
                const double* x0, const double* y0, const double* z0, int T, TileHost& out);

}  // namespace bf
