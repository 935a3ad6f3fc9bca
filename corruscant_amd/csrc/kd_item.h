// Work items of the pair kernels: one (query tile, run of tree tiles) pair
// with the radius window whose counts it still has to decide.
#pragma once

#include <hip/hip_runtime.h>

namespace kd {

// Items of the enumeration (16 bytes).  `run` packs three small numbers:
// bits 0..6 the tree tiles of the run (1..127; the enumeration merges at most
// 64), bits 7..10 and 11..14 the quads of four points that the trim pass cut
// off the first and the last tile (0..15 each).  It is private so that every
// read and write goes through the accessors below; a kernel that indexed
// tile arrays with the packed value would read far past their ends.
struct Item {
    int q;    // query tile
    int t;    // first tree tile; for a subtree item the hierarchy node
    int job;  // bit 31: mirrored (symmetric off-diagonal) item; bit 30: subtree item
    unsigned char elo, ehi;

    static constexpr int kMaxTiles = 0x7F;
    static constexpr int kMaxCut = 0xF;

    __host__ __device__ __forceinline__ int tiles() const { return run & 0x7F; }
    __host__ __device__ __forceinline__ int cut_lo() const { return (run >> 7) & 0xF; }
    __host__ __device__ __forceinline__ int cut_hi() const { return (run >> 11) & 0xF; }
    // a run of n tiles (1..127) without cuts
    __host__ __device__ __forceinline__ void set_tiles(int n) { run = (unsigned short)(n & 0x7F); }
    // the trim pass's cuts (0..15 each); the tile count is kept
    __host__ __device__ __forceinline__ void set_cuts(int lo, int hi)
    {
        run = (unsigned short)((run & 0x7F) | ((lo & 0xF) << 7) | ((hi & 0xF) << 11));
    }

  private:
    unsigned short run;
};
static_assert(sizeof(Item) == 16, "Item is 16 bytes");

static constexpr int kSubtreeItem = 0x40000000;
static constexpr int kJobMask = 0x3fffffff;

// Bits of the pair kernels' error word.
static constexpr unsigned kErrFieldIndex = 1u;  // a field id outside the count matrix
static constexpr unsigned kErrStack = 2u;       // enumeration stack overflow
static constexpr unsigned kErrItem = 4u;        // an item outside its job's tiles or quads
static constexpr unsigned kErrPlan = 8u;        // a replayed plan enumerated other items than when it was made

}  // namespace kd
