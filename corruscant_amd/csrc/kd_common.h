// Shared definitions of the gfx950 pair-counting library (host + device).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/corruscant_amd.h"

namespace kd {

struct Error : std::runtime_error {
    explicit Error(const std::string &m) : std::runtime_error(m) {}
};

#define KD_HIP(expr)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            throw ::kd::Error(std::string("HIP error ") + hipGetErrorString(e_) + " at " + \
                              __FILE__ + ":" + std::to_string(__LINE__) + " (" #expr ")");  \
    } while (0)

#define KD_REQUIRE(cond, msg)                 \
    do {                                      \
        if (!(cond)) throw ::kd::Error(msg);  \
    } while (0)

// Device memory of the library: stream-ordered allocations on the library's
// stream from a pool that keeps freed blocks for reuse (no device-wide
// synchronisation per free).  Defined with the stream in kd_api.hip.
void *dev_alloc(size_t bytes);
void dev_free(void *p);

// While a call runs with an arena (the entry point that allocates nothing),
// dev_alloc carves 256-byte aligned blocks from it instead, throws when it
// is used up, and dev_free of its blocks does nothing.
struct Arena {
    unsigned char *base = nullptr;
    size_t size = 0;
    size_t used = 0;
    const char *who = "";  // the entry point, for the error message
};
bool arena_active();

// Owning device buffer.
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        if (count) p = (T *)dev_alloc(sizeof(T) * count);
    }
    void release() {
        if (p) dev_free(p);
        p = nullptr;
        n = 0;
    }
    size_t bytes() const { return sizeof(T) * n; }
};

// 32-byte tree-side record: one deduplicated point of the tree, in tile order.
struct TPoint {
    double x, y, z;
    uint32_t meta;  // weight (bits 0..7) | field (bits 8..15)
    int32_t heap;   // heap index of the node in the reference layout
};
static_assert(sizeof(TPoint) == 32, "TPoint is 32 bytes");


struct Box {
    double lo[3], hi[3];
};
// A box in FP32, rounded outwards: it contains the FP64 box it was made from.
struct BoxF {
    float lo[3], hi[3];
};

// A point set cut into spatially compact tiles (Morton order, <= tile_cap points).
struct Tiles {
    int n = 0;
    int ntiles = 0;
    int cap = 0;
    double cmax = 0.0;   // max |coordinate| of the set
    double hmax = 0.0;   // tree side: every point lies within hmax of its tile box's midpoint
    DevBuf<int> start;   // ntiles + 1
    DevBuf<Box> box;     // ntiles
    // hierarchy over tiles: implicit binary heap of 2*P boxes, leaves at P + t
    int P = 0;
    DevBuf<Box> hbox;
};

struct QuerySide {  // raw points, weight 1 each
    Tiles t;
    DevBuf<int> order;       // raw index of each tile-order position
    DevBuf<double> x, y, z;  // n, tile order
    DevBuf<int> f;           // raw field ids
    DevBuf<int> fmin, fmax;  // per tile field range
};

struct TreeSide {  // deduplicated weighted points; tile t covers the same
                   // raw range as query tile t of the same catalogue
    Tiles t;
    DevBuf<TPoint> pts;      // n, tile order
    DevBuf<double> centre;   // ntiles * 3: tile box midpoints
    // the FP32 form's operand groups: per tile ceil(points / 16) groups of 64
    // floats (kd_layout.hip k_fill_groups), starting at gstart[tile]
    DevBuf<float> grp;
    DevBuf<int> gstart;      // ntiles + 1
    DevBuf<unsigned> meta;   // ntiles: bit 0 one tree field, bit 1 all weights 1, bits 8..15 the field,
                             // bits 16..19 the size class (FP32 form)
    DevBuf<int> raw;         // n: raw index (input order) of each point's representative
    // the coordinates again in quads of 4 points, {x0..x3, y0..y3, z0..z3}
    // (96 bytes, two wide scalar loads).  The points of a tile are grouped
    // by tree field; each run of one field within a tile starts a new quad
    // and pads its last one with +inf points, so every quad has one field.
    DevBuf<double> quad;
    DevBuf<int> quad_start;      // ntiles + 1: first quad of each tile
    DevBuf<unsigned> quad_meta;  // per quad: field (bits 0..7) | points in the quad (bits 8..10)
    DevBuf<BoxF> quad_box;       // per quad: bounding box of its points (the padding left out)
    // the same quads' FP32 operands for the weighted kernels' FP32 form: per
    // point {bx, by, bz, |b|^2} (16 bytes, b relative to the tile centre;
    // padding {0, 0, 0, +inf}), so lane l of a wave loads point l % 4 of a
    // quad with one 16-byte load
    DevBuf<float> quad32;
    // per node of the tile hierarchy (2P entries, node 1 the root, leaf P + t
    // tile t): the tree field of every point of the subtree (kHMixed if
    // several, kHEmpty if none) and the subtree's total weight
    DevBuf<unsigned> hfield;
    DevBuf<long long> hweight;
    // per hierarchy node spanning at least 128 tile slots (nodes 1 .. P/64-1):
    // the subtree's total weight of each tree field, nfw entries per node
    // (nfw == 0: not kept, because a field id did not fit)
    DevBuf<long long> hfw;
    int nfw = 0;
};

constexpr unsigned kHMixed = 0xFFFFFFFFu;
constexpr unsigned kHEmpty = 0xFFFFFFFEu;

}  // namespace kd

// Opaque handle behind kdtree_t::dev.
struct kd_device_tree {
    int n_raw = 0;
    int n_u = 0;
    int memsize = 0;
    int num_fields = 1;
    bool query_only = false;
    double cmax = 0.0;  // max |coordinate| over the raw points
    bool symmetric_ok = false;  // no duplicates and raw fields in [0, 255]: DD-type
                                // jobs may count each unordered tile pair once
    kd::DevBuf<unsigned char> nodes;  // memsize * 40 bytes, reference node layout
    kd::DevBuf<double> split;         // memsize: value[dim] of each node
    kd::DevBuf<uint32_t> nbits;       // memsize: has_l | has_r << 1 | dim << 2 (0 for empty slots)
    kd::DevBuf<int> perm;             // memsize: heap -> unique index (-1 empty)
    kd::DevBuf<int> inorder;          // memsize: heap -> median position
    kd::QuerySide q;
    kd::TreeSide tr;
    node_t *host_nodes = nullptr;
};
