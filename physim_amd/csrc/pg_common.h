// Shared host/device definitions of the physim_gpu engine.
#pragma once

#include <stdint.h>

#include "physim_gpu.h"

namespace pg {

constexpr int kBlock = 256;        // threads per block (4 wave64)
constexpr int kSortTile = 2048;    // keys per radix-sort tile
constexpr int kSortThreads = 512;  // threads per radix-sort tile
constexpr int kSortItems = kSortTile / kSortThreads;  // keys per thread
constexpr int kScanItems = 4;      // counts per thread in the node-offset scan
constexpr int kScanTile = kBlock * kScanItems;
constexpr int kMaxPasses = 8;      // 8-bit digits over a 64-bit key
constexpr int kKeyThreads = 512;   // threads of a key-building block (one sort tile)
constexpr int kBucketKeys = 1664;  // keys per bucket of the sample sort (on average; steady-state
                                   // buckets stay within 2048, the slot sort's reach)
constexpr int kBucketCap = 4096;   // largest bucket sorted in LDS (larger ones: in HBM)
constexpr int kMinBucketKeys = 256;  // smallest mean bucket a tuning may ask for
constexpr int kBucketThreads = 512;   // threads of a bucket-sorting block (up to 8 keys each)
constexpr int kMaxBuckets = 1024;  // buckets whose splitters every key tile holds in LDS
constexpr int kMaxWideBuckets = 8192;  // more buckets: splitters in HBM behind a coarse table
constexpr int kWideGroup = 8;      // splitters per coarse-table entry (kMaxWideBuckets / kWideGroup
                                   // entries fit a tile's LDS)
constexpr int kSampleOver = 4;     // splitter samples per bucket (sampling from unsorted keys)
constexpr int kWideSampleOver = 2; // the same with more than kMaxBuckets buckets (128 KB of LDS)
constexpr int kSplitThreads = 1024;  // threads of the splitter block
constexpr int kQuantiles = 256;    // per bucket: quantiles of the previous sorted keys
constexpr int kScanThreads = 1024;  // threads of a per-digit scan over the sort tiles
constexpr int kExtentChunk = 4096; // bodies per extent partial
constexpr uint32_t kExtentBlocks = 256;  // blocks of the extent pass (each takes chunks in turn)
constexpr int kNodeTile = 1024;    // sorted bodies per node-building block
constexpr int kNodeThreads = 512;  // threads per node-building block (two bodies each)
constexpr int kNodeLds = 1792;     // kNodeTile bodies + up to kNodeLds - kNodeTile internal nodes: built one
                                   // node per thread (more internal nodes: body by body)
constexpr int kTopSlots = 65;      // tree levels 0..64: one tile-spanning node each
constexpr int kMaxTieRun = 256;    // longest run of equal keys ordered by their deep keys
constexpr int kMaxLcp = 64;        // deepest common prefix, in levels, of two bodies: the
                                   // reference recurses to depth 64 (octree.rs:60-62)
constexpr int kDeepWords = 3;      // words of a tied body's deep key
static_assert(kMaxLcp + 1 == kTopSlots, "every level of an internal node has its slot");

// Deep key of a body whose first key word ties with a neighbour's: the
// digits of the levels past the first word, from the same descent.
struct DeepKey {
    uint64_t w[kDeepWords];
};
constexpr int kComSpanGroups = 5;   // level groups (of 4 waves) per tile in k_com_span
static_assert(kScanTile == kNodeTile, "node offsets and node building share their tiles");
static_assert(kNodeTile == 2 * kNodeThreads, "the partial sums take two bodies per thread");
static_assert(kTopSlots <= kBlock, "per-level tables are filled by one thread per level");
static_assert(kSortTile % kKeyThreads == 0 && kKeyThreads >= 256, "whole key tiles, a digit per thread");
static_assert(kSortTile % kSortThreads == 0 && kSortThreads >= 256, "whole sort tiles, a digit per thread");

// Device-side error bits (sticky per segment), mapped to pg_status on the host.
enum : uint32_t {
    kErrDepth = 1u << 0,
    kErrCoincident = 1u << 1,
    kErrNonFinite = 1u << 2,
    kErrCapacity = 1u << 3,
};

struct SegState;

// A body's tag, the fourth word of its previous-position record: its upload
// index (low 32 bits) and its fixed flag (bit 32), a bit pattern that no
// arithmetic touches.
__host__ __device__ inline double make_tag(uint32_t id, bool fixed) {
    return __builtin_bit_cast(double, (uint64_t)id | ((uint64_t)(fixed ? 1u : 0u) << 32));
}
__host__ __device__ inline uint32_t tag_id(double t) {
    return (uint32_t)__builtin_bit_cast(uint64_t, t);
}
__host__ __device__ inline bool tag_fixed(double t) {
    return (__builtin_bit_cast(uint64_t, t) >> 32) != 0;
}

// One simulation's view inside a batched launch: the current state (src) and
// the buffer the step writes (dst), both SoA in the order of the previous
// build's Morton sort.
struct SegDesc {
    // per body: (x, y, z, m) and (x_prev, y_prev, z_prev, tag); velocities
    const double4* spos; const double4* sprv;
    const double* svx; const double* svy; const double* svz;
    double4* dpos; double4* dprv;
    double* dvx; double* dvy; double* dvz;
    uint64_t off;   // first global (batch) index of this segment
    uint64_t n;
    double theta, e, dt;
    uint32_t seg;
    uint32_t count; // per-body interaction counts into counts[] and step statistics
    int64_t* counts;
    struct SegState* st;
    uint32_t chunk0;  // first extent chunk of this segment
    uint32_t nchunks;
    uint32_t tile0;   // first node tile of this segment (node tiles start at its first body)
    uint32_t nsplit;  // direct sum: blocks summing one target tile, each over a range of sources
    uint64_t part_off;  // direct sum with nsplit > 1: first partial sum (x3 doubles) of the segment
    double near;      // near field in widths of the body's own leaf cell (0: off)
};

// Per-segment device scalars.
struct SegState {
    uint64_t extent_bits;   // max |coord| as the bit pattern of a non-negative double
    uint64_t interactions;
    uint64_t steps_done;
    uint32_t history;       // 1 when the Stormer history is valid for this step
    uint32_t err;
    uint32_t max_level;
    uint32_t vperm;         // velocities must travel with the state (not yet derivable)
};

// Internal tree node record, one 64-byte line (what the walk reads per
// visited node). Every field is a whole number of 32-bit words, so the walk
// fetches a record through the scalar unit (which has no 16-bit loads).
struct NodeHot {
    double cx, cy, cz;      // geometric centre (reference rounding)
    double mx, my, mz;      // centre of mass
    double mass;
    uint32_t skip;          // index of the next internal node after this subtree
    uint32_t lvl_oct;       // level (bits 0-15), child index within the parent (bits 16-23)
    __host__ __device__ int level() const { return (int)(lvl_oct & 0xFFFFu); }
    __host__ __device__ uint32_t octant() const { return (lvl_oct >> 16) & 0xFFu; }
};

// Batch-wide counters reset at the start of every step.
struct StepCounters {
    uint32_t gen;           // bumped per step: versions tile-spanning node slots
    uint32_t top_levels;    // 1 + deepest level of a tile-spanning node (0: none)
    uint64_t sort_tag;      // batch whose sorted keys the sample sort left (0: none)
    uint64_t ext_tag;       // batch whose current positions the extent partials hold (0: none)
    uint32_t split_prev;    // this step's splitters may come from those keys
    uint32_t quantiles;     // they did: finer quantiles per bucket were taken too
};

}  // namespace pg
