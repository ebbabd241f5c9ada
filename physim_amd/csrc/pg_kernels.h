// Device kernels of the physim_gpu engine (gfx950, wave64). Included once,
// by engine.hip. Compiled with -ffp-contract=off: every FP64 expression that
// must round as the reference does is written in the reference's operation
// order, and FP32 force code asks for fused multiply-adds explicitly.
#pragma once

#include <hip/hip_runtime.h>

#include "pg_common.h"

namespace pg {

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ double bits_f64(uint64_t b) { return __longlong_as_double((long long)b); }
__device__ __forceinline__ uint64_t f64_bits(double d) { return (uint64_t)__double_as_longlong(d); }
__device__ __forceinline__ float bits_f32(uint32_t b) { return __uint_as_float(b); }
__device__ __forceinline__ uint32_t f32_bits(float f) { return __float_as_uint(f); }

// The reference's `extent / 2.0` (octree.rs:97-100). Halving by a multiply
// rounds exactly as the IEEE division does (scaling by a power of two, also
// into the subnormal range), without relying on the compiler to see that.
__device__ __forceinline__ double half(double x) { return x * 0.5; }

// A uniform double copied into a vector register. The walk copies an item's
// source (a node's centre of mass and mass, read through the scalar unit)
// before it loads the next node record: the record then lands in the
// registers it is carried in, and the pair term runs while the load is in
// flight, instead of the compiler copying the new record over and waiting
// for it first. (A move, not a load: nothing is in flight in it.)
__device__ __forceinline__ double vcopy(double v) {
    double r;
    asm("v_mov_b64 %0, %1" : "=v"(r) : "s"(v));
    return r;
}

__device__ __forceinline__ uint32_t uniform_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

// Element idx of an array, its byte offset formed in 32 bits (idx * sizeof(T)
// below 2^32): a uniform load then takes the offset in a scalar register
// instead of a 64-bit address add per load.
template <class T>
__device__ __forceinline__ const T& at32(const T* __restrict__ base, uint32_t idx) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) +
                                       (uint32_t)(idx * (uint32_t)sizeof(T)));
}

// Global-memory access through a pointer held in a descriptor. The compiler
// cannot tell where such a pointer points and emits flat instructions, which
// also count against the scalar-load counter: every wait for a scalar load
// (a descriptor field, a node record) then waits for them too. Through the
// global address space they are plain global loads and stores.
// (As 32-bit words: HIP's vector types do not copy from address-space
// qualified references.)
template <class T>
__device__ __forceinline__ T gload(const T* p) {
    static_assert(sizeof(T) % 4 == 0, "whole words");
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, *(const __attribute__((address_space(1))) uint32_t*)p);
    } else {
        typedef uint32_t words __attribute__((ext_vector_type(sizeof(T) / 4)));
        return __builtin_bit_cast(T, *(const __attribute__((address_space(1))) words*)p);
    }
}
template <class T>
__device__ __forceinline__ void gstore(T* p, const T& v) {
    static_assert(sizeof(T) % 4 == 0 && sizeof(T) >= 8, "whole words");
    typedef uint32_t words __attribute__((ext_vector_type(sizeof(T) / 4)));
    *(__attribute__((address_space(1))) words*)p = __builtin_bit_cast(words, v);
}

// floor(a * b / c) for a * b < 2^53 and c > 0 by one FP64 division (a 64-bit
// integer division is a software sequence of about 150 instructions). Both
// operands are exact in FP64; a quotient that is not an integer lies at
// least 1/c from the integers around it, and the division's rounding moves
// it by at most q 2^-53 = a b / c 2^-53 < 1/c, so its floor is exact. The
// sample positions below stay in range: indices below 2^32 (the batch's
// body count) times at most kMaxWideBuckets * kQuantiles = 2^21.
__device__ __forceinline__ uint64_t muldiv_floor(uint64_t a, uint64_t b, uint64_t c) {
    return (uint64_t)floor((double)(a * b) / (double)c);
}
static_assert((uint64_t)kMaxWideBuckets * kQuantiles <= (1ull << 21) &&
                  (uint64_t)kMaxWideBuckets * kWideSampleOver <= (1ull << 21) &&
                  (uint64_t)kMaxBuckets * kSampleOver <= (1ull << 21),
              "sample positions: index times count below 2^53");

// Segment of global index g: largest s with seg_off[s] <= g (empty segments
// share an offset with their successor and are skipped).
__device__ __forceinline__ uint32_t seg_of(const uint64_t* __restrict__ seg_off, uint32_t nseg,
                                           uint64_t g) {
    uint32_t lo = 0, hi = nseg;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (seg_off[mid] <= g) lo = mid; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint32_t seg_of_key(uint64_t key, int segbits) {
    return segbits ? (uint32_t)(key >> (64 - segbits)) : 0u;
}

// Node tile of global index g < total: node tiles of kNodeTile bodies start
// at every segment's first body.
__device__ __forceinline__ uint32_t tile_of(const SegDesc* __restrict__ desc,
                                            const uint64_t* __restrict__ seg_off, uint32_t nseg,
                                            uint64_t g) {
    const SegDesc& d = desc[seg_of(seg_off, nseg, g)];
    return d.tile0 + (uint32_t)((g - d.off) / kNodeTile);
}

// Key prefix holding the segment bits and the top `level` tree digits.
template <int DIM>
__device__ __forceinline__ uint64_t key_prefix(uint64_t key, int lmax, int level) {
    int sh = DIM * (lmax - level);
    return sh >= 64 ? 0ull : key >> sh;
}

// Common prefix length, in tree levels, of two keys of one segment.
template <int DIM>
__device__ __forceinline__ int key_lcp(uint64_t a, uint64_t b, int lmax) {
    uint64_t x = a ^ b;
    int mbits = DIM * lmax;
    if (mbits < 64) x &= (1ull << mbits) - 1ull;
    if (x == 0) return lmax;
    int lz = __clzll((long long)x) - (64 - mbits);
    return lz / DIM;
}

// Deep key. Bodies whose first words tie (they share every one of the lmax
// levels a key holds) carry the next levels of the same descent, up to the
// deepest level two bodies may share (kMaxLcp): kWordLevels levels per word,
// level lmax + t in word t / W at bits [DIM (W - 1 - t % W), DIM (W - t % W)).
// Defined for tied bodies only.
template <int DIM>
constexpr int kWordLevels = DIM == 3 ? 21 : 32;
static_assert(21 + kDeepWords * 21 > kMaxLcp && 32 + kDeepWords * 32 > kMaxLcp,
              "a first word and the deep key resolve every level a prefix may reach");

template <int DIM>
__device__ __forceinline__ uint32_t deep_digit(const DeepKey& k, int t) {
    constexpr int W = kWordLevels<DIM>;
    return (uint32_t)((k.w[t / W] >> (DIM * (W - 1 - t % W))) & ((1u << DIM) - 1u));
}

// Octant digit taken at tree depth d (from level d to level d + 1).
template <int DIM>
__device__ __forceinline__ uint32_t key_digit(uint64_t key, const DeepKey& dk, int lmax, int d) {
    const uint64_t dmask = (1u << DIM) - 1u;
    return d < lmax ? (uint32_t)((key >> (DIM * (lmax - 1 - d))) & dmask)
                    : deep_digit<DIM>(dk, d - lmax);
}

// Common prefix length, in levels, of two deep keys.
template <int DIM>
__device__ __forceinline__ int deep_lcp(const DeepKey& a, const DeepKey& b) {
    constexpr int W = kWordLevels<DIM>;
    int l = 0;
    for (int i = 0; i < kDeepWords; ++i) {
        const int c = key_lcp<DIM>(a.w[i], b.w[i], W);
        l += c;
        if (c < W) break;
    }
    return l;
}

// Common prefix length in levels of bodies ia and ib (keys a, b), reading
// the deep keys only when the first words are equal.
template <int DIM>
__device__ __forceinline__ int key_lcp2(uint64_t a, uint64_t b, const DeepKey* __restrict__ dk,
                                        uint64_t ia, uint64_t ib, int lmax) {
    const int l = key_lcp<DIM>(a, b, lmax);
    return l < lmax ? l : lmax + deep_lcp<DIM>(dk[ia], dk[ib]);
}

// DPP move across a wave's lanes (CTRL: the DPP control, ROWS: the rows
// written); a lane without a source, or in a row not written, takes old.
template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t dpp_move(uint32_t v, uint32_t old) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, CTRL, ROWS, 0xF, false);
}
constexpr int kDppWaveShr1 = 0x138;  // lane l takes lane l - 1 (lane 0: old)

// Inclusive scan over a wave's lanes in lane order with an associative
// operation of identity id: steps of 1, 2, 4 and 8 lanes within rows of 16
// (row_shr), then rows 1 and 3 take the last lane of the row before
// (row_bcast:15), rows 2 and 3 lane 31 (row_bcast:31).
template <class Op>
__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v, uint32_t id, Op op) {
    v = op(dpp_move<0x111, 0xF>(v, id), v);
    v = op(dpp_move<0x112, 0xF>(v, id), v);
    v = op(dpp_move<0x114, 0xF>(v, id), v);
    v = op(dpp_move<0x118, 0xF>(v, id), v);
    v = op(dpp_move<0x142, 0xA>(v, id), v);
    v = op(dpp_move<0x143, 0xC>(v, id), v);
    return v;
}
struct OpAdd {
    __device__ __forceinline__ uint32_t operator()(uint32_t a, uint32_t b) const { return a + b; }
};
struct OpMin {
    __device__ __forceinline__ uint32_t operator()(uint32_t a, uint32_t b) const { return min(a, b); }
};

// Exclusive scan of one value per thread over a block of NT threads
// (s_wave holds NT / 64 values).
template <int NT = kBlock>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_wave,
                                                         uint32_t* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t inc = wave_incl_scan_u32(v, 0u, OpAdd{});
    if (lane == 63) s_wave[wave] = inc;
    __syncthreads();
    uint32_t wbase = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) {
        uint32_t c = s_wave[w];
        if (w < wave) wbase += c;
        tot += c;
    }
    __syncthreads();
    *total = tot;
    return wbase + inc - v;
}

// ------------------------------------------------------------ step prologue
// Opens a segment's step: its extent (bit pattern of max |coord|; 0 for the
// direct sum, which has no tree) and step state.
__device__ __forceinline__ void open_segment(SegState* st, uint64_t extent_bits, int integrate) {
    st->extent_bits = extent_bits;
    st->interactions = 0;
    st->max_level = 0;
    st->history = st->steps_done > 0 ? 1u : 0u;
    // v is (x - x_prev) / dt after any step but the Taylor first one
    // (verlet.rs:68-70), so it only has to be carried before and right
    // after that step.
    st->vperm = st->steps_done <= 1 ? 1u : 0u;
    if (integrate) st->steps_done += 1;
}

// The direct step's prologue: batch counters and every segment's step.
__global__ void k_prep(const SegDesc* __restrict__ desc, uint32_t nseg, StepCounters* sc,
                       int integrate) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    if (t == 0) {
        sc->gen += 1;
        sc->top_levels = 0;
    }
    for (uint32_t s = t; s < nseg; s += stride) open_segment(desc[s].st, 0, integrate);
}

// Block-wide max of one u64 per thread over NT threads.
template <int NT = kBlock>
__device__ __forceinline__ uint64_t block_max_u64(uint64_t v, uint64_t* s_wave) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint64_t t = __shfl_xor(v, o);
        v = t > v ? t : v;
    }
    if (lane == 0) s_wave[wave] = v;
    __syncthreads();
    uint64_t m = s_wave[0];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w) m = s_wave[w] > m ? s_wave[w] : m;
    __syncthreads();
    return m;
}

// transformers.rs:126-131: max |coord| over a segment (exact). Each block
// reduces one chunk of kExtentChunk bodies of one segment into partial[];
// NaN/Inf coordinates are flagged. Nothing to do when the batch's previous
// integrating walk left the partials of the current positions (ext_tag).
__global__ __launch_bounds__(kBlock) void k_extent_partial(const SegDesc* __restrict__ desc,
                                                           const uint32_t* __restrict__ chunk_seg,
                                                           const uint32_t* __restrict__ chunk_idx,
                                                           uint64_t* __restrict__ partial,
                                                           const StepCounters* __restrict__ sc,
                                                           uint64_t tag, uint32_t nchunks) {
    __shared__ uint64_t s_wave[kBlock / 64];
    if (tag != 0 && sc->ext_tag == tag) return;
    // (a small grid: the skip above is the common case)
    for (uint32_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const SegDesc& d = desc[chunk_seg[c]];
        const uint64_t j0 = (uint64_t)chunk_idx[c] * kExtentChunk;
        const uint64_t j1 = min(j0 + kExtentChunk, d.n);
        uint64_t m = 0;
        bool bad = false;
        for (uint64_t j = j0 + threadIdx.x; j < j1; j += kBlock) {
            const double4 p = d.spos[j];
            const double x = p.x, y = p.y, z = p.z;
            bad |= !isfinite(x) || !isfinite(y) || !isfinite(z);
            const uint64_t b = f64_bits(fmax(fmax(fabs(x), fabs(y)), fabs(z)));
            m = b > m ? b : m;
        }
        if (__ballot(bad) && (threadIdx.x & 63) == 0) atomicOr(&d.st->err, kErrNonFinite);
        m = block_max_u64(m, s_wave);
        if (threadIdx.x == 0) partial[c] = m;
    }
}

// Adds one digit per lane (256 = none) to the block's LDS digit counters:
// lanes of a wave that hold the same digit are found with match-any ballots
// and their leader adds the group's size. Every lane of the wave must call.
__device__ __forceinline__ void wave_count_digit(uint32_t* s_h, uint32_t d) {
    const bool valid = d < 256u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bb = __ballot(bit);
        peers &= bit ? bb : ~bb;
    }
    if (valid && (int)(threadIdx.x & 63) == __ffsll((long long)peers) - 1)
        atomicAdd(&s_h[d], (uint32_t)__popcll(peers));
}

// Bit b of q at bit DIM * b (q < 2^(64 / DIM)).
template <int DIM>
__device__ __forceinline__ uint64_t spread_bits(uint64_t q) {
    if (DIM == 3) {
        q &= 0x1FFFFFull;
        q = (q | (q << 32)) & 0x1F00000000FFFFull;
        q = (q | (q << 16)) & 0x1F0000FF0000FFull;
        q = (q | (q << 8)) & 0x100F00F00F00F00Full;
        q = (q | (q << 4)) & 0x10C30C30C30C30C3ull;
        q = (q | (q << 2)) & 0x1249249249249249ull;
    } else {
        q &= 0xFFFFFFFFull;
        q = (q | (q << 16)) & 0x0000FFFF0000FFFFull;
        q = (q | (q << 8)) & 0x00FF00FF00FF00FFull;
        q = (q | (q << 4)) & 0x0F0F0F0F0F0F0F0Full;
        q = (q | (q << 2)) & 0x3333333333333333ull;
        q = (q | (q << 1)) & 0x5555555555555555ull;
    }
    return q;
}

// One axis of the descent below without it: with exact arithmetic the axis'
// digit at level l (p > c) is bit (lmax - 1 - l) of floor(w), w = (p / E + 1)
// 2^(lmax - 1), every cell boundary of levels < lmax being an integer in w.
// The reference's centres are rounded, by at most lmax ulp(E) in all (below
// 2^(lmax - 48) in w), and w here is within 2^(lmax - 51) of its exact
// value; so a coordinate farther than 2^(lmax - 41) (in w) from every
// integer has the descent's digits. Returns false (with q = 0) for any other
// coordinate or a non-finite one: those take the descent.
__device__ __forceinline__ bool axis_digits(double p, double inv, double scale, double margin,
                                            uint64_t& q) {
    const double w = (p * inv + 1.0) * scale;
    const double f = w - floor(w);
    const bool ok = f > margin && f < 1.0 - margin;  // (false for NaN)
    q = ok ? (uint64_t)w : 0ull;
    return ok;
}

// Bucket of a key: the number of splitters (sorted, nb - 1 of them) <= key.
__device__ __forceinline__ uint32_t bucket_of(const uint64_t* s_spl, uint32_t nb, uint64_t key) {
    uint32_t lo = 0, hi = nb - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s_spl[mid] <= key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// The same count for K keys at once: a fixed number of halvings (no branch,
// no divergence), the K searches' loads in flight together.
template <int K>
__device__ __forceinline__ void buckets_of(const uint64_t* s_spl, uint32_t nb, const uint64_t* key,
                                           uint32_t* b) {
    static_assert(kMaxBuckets <= 1024, "ten halvings cover the splitters");
    const uint32_t ns = nb - 1;
#pragma unroll
    for (int k = 0; k < K; ++k) b[k] = 0;
#pragma unroll
    for (uint32_t step = kMaxBuckets / 2; step; step >>= 1) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t q = b[k] + step;
            if (q <= ns && s_spl[q - 1] <= key[k]) b[k] = q;
        }
    }
}

// The same count with more than kMaxBuckets buckets: a search of the coarse
// table s_c (every kWideGroup-th splitter, (nb - 1) / kWideGroup of them) for
// the group c of the key's bucket, then of the group's splitters in spl (L2):
// the bucket lies in [kWideGroup c, kWideGroup c + kWideGroup - 1]. Fixed
// halvings, no divergence.
template <int K>
__device__ __forceinline__ void buckets_of_wide(const uint64_t* s_c, const uint64_t* __restrict__ spl,
                                                uint32_t nb, const uint64_t* key, uint32_t* b) {
    constexpr uint32_t kCoarseMax = kMaxWideBuckets / kWideGroup;
    static_assert((kCoarseMax & (kCoarseMax - 1)) == 0 && (kWideGroup & (kWideGroup - 1)) == 0,
                  "halvings cover the tables");
    const uint32_t ns = nb - 1, nc = ns / kWideGroup;
#pragma unroll
    for (int k = 0; k < K; ++k) b[k] = 0;
#pragma unroll
    for (uint32_t step = kCoarseMax / 2; step; step >>= 1) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t q = b[k] + step;
            if (q <= nc && s_c[q - 1] <= key[k]) b[k] = q;
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) b[k] *= kWideGroup;
#pragma unroll
    for (uint32_t step = kWideGroup / 2; step; step >>= 1) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t q = b[k] + step;
            if (q <= ns && spl[q - 1] <= key[k]) b[k] = q;
        }
    }
}

// Morton key by exact FP64 descent (octree.rs:160-215 / quadtree.rs:161-194):
// at each level the octant is (x > cx) | (y > cy) << 1 [| (z > cz) << 2] and
// the child centre is c -/+ extent/2, rounded as the reference rounds it;
// a body clear of every cell boundary takes the same digits from its scaled
// coordinates (axis_digits). `ext_bits`: the root's half-width as the bits of
// a double.
template <int DIM>
__device__ __forceinline__ uint64_t morton_key(double px, double py, double pz,
                                               uint64_t ext_bits, int lmax) {
    double ext = bits_f64(ext_bits);
    const double inv = 1.0 / ext;
    const double scale = ldexp(1.0, lmax - 1), margin = ldexp(1.0, lmax - 41);
    uint64_t qx, qy, qz = 0;
    // (an extent this far from 1 makes 1/E or E's halvings subnormal)
    bool fast = ext >= 0x1p-900 && ext <= 0x1p+1000;
    fast &= axis_digits(px, inv, scale, margin, qx);
    fast &= axis_digits(py, inv, scale, margin, qy);
    if (DIM == 3) fast &= axis_digits(pz, inv, scale, margin, qz);
    uint64_t key = spread_bits<DIM>(qx) | (spread_bits<DIM>(qy) << 1);
    if (DIM == 3) key |= spread_bits<DIM>(qz) << 2;
    if (!fast) {
        double cx = 0.0, cy = 0.0, cz = 0.0;
        key = 0;
        for (int l = 0; l < lmax; ++l) {
            uint32_t o = (uint32_t)(px > cx) | ((uint32_t)(py > cy) << 1);
            if (DIM == 3) o |= (uint32_t)(pz > cz) << 2;
            const double h2 = half(ext);
            cx = (o & 1u) ? cx + h2 : cx - h2;
            cy = (o & 2u) ? cy + h2 : cy - h2;
            if (DIM == 3) cz = (o & 4u) ? cz + h2 : cz - h2;
            ext = h2;
            key = (key << DIM) | o;
        }
    }
    return key;
}

// Each body's Morton key (morton_key) and its index into keys / vals.
// For the radix sort (sort_tag 0) each tile also counts its digits of the
// first pass. With FUSE (the sample sort, the batch's previous sorted keys
// in `prev`) it does the work of k_splitters and k_bucket_count instead: the
// same splitters (prev at the quantile positions), each key's bucket, the
// bucket sizes (one atomic per tile and bucket; the bucket sort leaves them
// zero) and the finer quantiles for the slot sort (FUSE 1). FUSE 2 (more than
// kMaxBuckets buckets) does the work of k_bucket_count_wide: the splitters
// (spl) and their coarse table (spc) come from k_splitters_wide, run first.
// One block per radix-sort tile of kSortTile bodies (the batch's segments are
// concatenated), which also writes the tile's digit counts of the first radix
// pass (digit-major: cnt[digit * tiles + tile]). The extent
// (transformers.rs:126-131) is the maximum of the segment's chunk partials,
// reduced by the whole block when the tile lies in one segment; the thread of
// a segment's first body stores it and opens the segment's step, and block 0
// resets the batch counters.
template <int DIM, int FUSE>
__global__ __launch_bounds__(kKeyThreads) void k_keys(
    const SegDesc* __restrict__ desc, const uint64_t* __restrict__ seg_off, uint32_t nseg,
    uint64_t total, int segbits, int lmax, const uint64_t* __restrict__ partial,
    StepCounters* __restrict__ sc, int integrate, uint64_t sort_tag, uint64_t* __restrict__ keys,
    uint32_t* __restrict__ vals, uint32_t* __restrict__ cnt, uint32_t tiles,
    const uint64_t* __restrict__ prev, uint32_t nb, uint16_t* __restrict__ bkt,
    uint32_t* __restrict__ bcount, uint64_t* __restrict__ quant, const uint64_t* __restrict__ spl,
    const uint64_t* __restrict__ spc) {
    static_assert(kMaxWideBuckets / kWideGroup <= kMaxBuckets, "the coarse table in s_spl");
    __shared__ uint64_t s_wave[kKeyThreads / 64];
    __shared__ uint32_t s_h[256];
    // the splitters (FUSE 1) or the coarse table (FUSE 2)
    __shared__ uint64_t s_spl[kMaxBuckets];
    // bucket counts: of all buckets (FUSE 1), or of a window from s_wb (FUSE 2)
    __shared__ uint32_t s_bh[kMaxBuckets];
    __shared__ uint32_t s_wb;
    constexpr int kItems = kSortTile / kKeyThreads;
    constexpr int kSplitItems = kMaxBuckets / kKeyThreads;
    static_assert(kMaxBuckets % kKeyThreads == 0, "splitter loads cover the table");
    const uint64_t g0 = (uint64_t)blockIdx.x * kSortTile;
    const uint64_t glast = min(g0 + (uint64_t)kSortTile, total) - 1;
    const uint32_t sa = seg_of(seg_off, nseg, g0);
    const uint32_t sb = seg_of(seg_off, nseg, glast);
    const bool one = sa == sb;  // block-uniform
    // A tile in one segment (the usual case) requests its bodies' positions,
    // the splitters and its first extent partials before anything waits on
    // them, so the block pays one memory round trip for all of them instead
    // of one per key (indices clamped into range rather than guarded: a load
    // inside a divergent branch is waited for at the branch's end).
    double4 q[kItems];
    uint64_t mblock = 0;
    if (one) {
        const SegDesc& d = desc[sa];
        const double4* sp = d.spos;
        const uint64_t off = d.off;
#pragma unroll
        for (int k = 0; k < kItems; ++k)
            q[k] = gload(sp + (min(g0 + (uint64_t)k * kKeyThreads + threadIdx.x, glast) - off));
        if (d.nchunks) mblock = partial[d.chunk0 + min((uint32_t)threadIdx.x, d.nchunks - 1)];
    }
    uint64_t sv[kSplitItems];
    if (FUSE == 1) {
#pragma unroll
        for (int i = 0; i < kSplitItems; ++i) {
            const uint32_t b = threadIdx.x + (uint32_t)i * kKeyThreads;
            sv[i] = prev[muldiv_floor(min(b + 1u, nb - 1u), total, nb)];
        }
    } else if (FUSE == 2) {
        const uint32_t nc = (nb - 1) / kWideGroup;  // >= 128: nb > kMaxBuckets
#pragma unroll
        for (int i = 0; i < kSplitItems; ++i)
            sv[i] = spc[min(threadIdx.x + (uint32_t)i * kKeyThreads, nc - 1u)];
    }
    if (threadIdx.x < 256) s_h[threadIdx.x] = 0;
    if (FUSE == 1) {
#pragma unroll
        for (int i = 0; i < kSplitItems; ++i) {
            const uint32_t b = threadIdx.x + (uint32_t)i * kKeyThreads;
            if (b < nb) s_bh[b] = 0;
            if (b + 1 < nb) s_spl[b] = sv[i];
        }
    } else if (FUSE == 2) {
        const uint32_t nc = (nb - 1) / kWideGroup;
#pragma unroll
        for (int i = 0; i < kSplitItems; ++i) {
            const uint32_t b = threadIdx.x + (uint32_t)i * kKeyThreads;
            s_bh[b] = 0;
            if (b < nc) s_spl[b] = sv[i];
        }
    }
    if (FUSE) {
        const uint64_t m = (uint64_t)nb * kQuantiles;
        for (uint64_t t = (uint64_t)blockIdx.x * kKeyThreads + threadIdx.x; t < m;
             t += (uint64_t)gridDim.x * kKeyThreads)
            quant[t] = prev[muldiv_floor(t, total, m)];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc->gen += 1;
        sc->top_levels = 0;
        // the sample sort may take its splitters from the sorted keys of this
        // batch's previous step (sort_tag 0: the radix sort, which leaves none)
        sc->split_prev = sort_tag != 0 && sc->sort_tag == sort_tag ? 1u : 0u;
        sc->sort_tag = sort_tag;
        if (FUSE) sc->quantiles = 1;
        sc->ext_tag = 0;  // the partials are read here; k_scan_gather clears them
    }
    // the tile's extent when it lies in one segment
    if (one) {
        const SegDesc& d = desc[sa];
        for (uint32_t c = threadIdx.x + kKeyThreads; c < d.nchunks; c += kKeyThreads) {
            const uint64_t b = partial[d.chunk0 + c];
            mblock = b > mblock ? b : mblock;
        }
        mblock = block_max_u64<kKeyThreads>(mblock, s_wave);
    } else {
        __syncthreads();
    }
    uint64_t tk[kItems];  // this thread's keys, for the fused bucket count
    if (one) {
        const SegDesc& d = desc[sa];
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            const uint64_t g = g0 + (uint64_t)k * kKeyThreads + threadIdx.x;
            uint32_t dig = 256u;
            tk[k] = 0;
            if (g <= glast) {
                if (g == d.off) open_segment(d.st, mblock, integrate);
                uint64_t key = morton_key<DIM>(q[k].x, q[k].y, q[k].z, mblock, lmax);
                if (segbits) key |= (uint64_t)sa << (64 - segbits);
                keys[g] = key;
                vals[g] = (uint32_t)g;
                dig = (uint32_t)(key & 255u);
                tk[k] = key;
            }
            if (sort_tag == 0) wave_count_digit(s_h, dig);
        }
    } else {
        // a tile across segments: each body finds its segment and extent
        for (int k = 0; k < kItems; ++k) {
            const uint64_t g = g0 + (uint64_t)k * kKeyThreads + threadIdx.x;
            uint32_t dig = 256u;
            tk[k] = 0;
            if (g < total) {
                const uint32_t s = seg_of(seg_off, nseg, g);
                const SegDesc& d = desc[s];
                uint64_t m = 0;
                for (uint32_t c = 0; c < d.nchunks; ++c) {
                    const uint64_t b = partial[d.chunk0 + c];
                    m = b > m ? b : m;
                }
                if (g == d.off) open_segment(d.st, m, integrate);
                const double4 p = d.spos[g - d.off];
                uint64_t key = morton_key<DIM>(p.x, p.y, p.z, m, lmax);
                if (segbits) key |= (uint64_t)s << (64 - segbits);
                keys[g] = key;
                vals[g] = (uint32_t)g;
                dig = (uint32_t)(key & 255u);
                tk[k] = key;
            }
            if (sort_tag == 0) wave_count_digit(s_h, dig);
        }
    }
    __syncthreads();
    if (sort_tag == 0) {
        if (threadIdx.x < 256) cnt[(uint64_t)threadIdx.x * tiles + blockIdx.x] = s_h[threadIdx.x];
    } else if (FUSE) {
        uint32_t bk[kItems];
        // the first bucket counted in LDS: 0 (every bucket), or, with more
        // buckets, a quarter of the window below the tile's first key's (the
        // keys arrive in the previous step's order; bodies that moved far
        // count straight to HBM)
        uint32_t wb = 0;
        if (FUSE == 1) {
            buckets_of<kItems>(s_spl, nb, tk, bk);
        } else {
            buckets_of_wide<kItems>(s_spl, spl, nb, tk, bk);
            if (threadIdx.x == 0) s_wb = bk[0] > kMaxBuckets / 4 ? bk[0] - kMaxBuckets / 4 : 0u;
            __syncthreads();
            wb = s_wb;
        }
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            const uint64_t g = g0 + (uint64_t)k * kKeyThreads + threadIdx.x;
            const bool valid = g < total;
            const uint32_t b = valid ? bk[k] : 0u;
            if (valid) bkt[g] = (uint16_t)b;
            const uint32_t r = b - wb;  // (wraps past the window below it)
            const bool in = valid && r < (uint32_t)kMaxBuckets;
            // a wave's keys mostly share a bucket (nearly sorted input): then
            // one atomic for the wave, else one per key
            const uint64_t vm = __ballot(in);
            const uint32_t rf = (uint32_t)__builtin_amdgcn_readlane(
                (int)r, vm ? __ffsll((long long)vm) - 1 : 0);
            if (__ballot(in && r != rf) == 0) {
                if (vm && (int)(threadIdx.x & 63) == __ffsll((long long)vm) - 1)
                    atomicAdd(&s_bh[rf], (uint32_t)__popcll(vm));
            } else if (in) {
                atomicAdd(&s_bh[r], 1u);
            }
            if (valid && !in) atomicAdd(&bcount[b], 1u);
        }
        __syncthreads();
        const uint32_t nw = min(nb - wb, (uint32_t)kMaxBuckets);
        for (uint32_t r = threadIdx.x; r < nw; r += kKeyThreads)
            if (s_bh[r]) atomicAdd(&bcount[wb + r], s_bh[r]);
    }
}

// Stable LSD radix sort of (key, value) pairs, 8-bit digits, three kernels
// per pass and no inter-tile waiting: (1) each tile counts its digits,
// (2) one block per digit scans that digit's counts over the tiles,
// (3) each tile ranks its keys and scatters them. (Tiles that wait on their
// predecessors, as in a single-pass look-back scheme, pay a chain of
// cross-XCD round trips when every tile of a pass is resident at once.)

// (1) Digit counts of one tile: cnt[digit * tiles + tile].
__global__ __launch_bounds__(kSortThreads) void k_sort_count(const uint64_t* __restrict__ kin,
                                                             uint64_t n, int shift,
                                                             uint32_t* __restrict__ cnt,
                                                             uint32_t tiles) {
    __shared__ uint32_t s_h[256];
    if (threadIdx.x < 256) s_h[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
    uint64_t key[kSortItems];
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const uint64_t idx = base + (uint64_t)k * kSortThreads + threadIdx.x;
        key[k] = idx < n ? kin[idx] : 0ull;
    }
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const uint64_t idx = base + (uint64_t)k * kSortThreads + threadIdx.x;
        wave_count_digit(s_h, idx < n ? (uint32_t)((key[k] >> shift) & 255u) : 256u);
    }
    __syncthreads();
    if (threadIdx.x < 256) cnt[(uint64_t)threadIdx.x * tiles + blockIdx.x] = s_h[threadIdx.x];
}

// (2) One block per digit: exclusive scan of the digit's per-tile counts (in
// place) and the digit's total.
__global__ __launch_bounds__(kScanThreads) void k_sort_scan(uint32_t* __restrict__ cnt,
                                                            uint32_t tiles,
                                                            uint32_t* __restrict__ tot) {
    __shared__ uint32_t s_wave[kScanThreads / 64];
    uint32_t* row = cnt + (uint64_t)blockIdx.x * tiles;
    const uint32_t per = (tiles + kScanThreads - 1) / kScanThreads;
    const uint32_t b = threadIdx.x * per;
    uint32_t s = 0;
    for (uint32_t k = 0; k < per; ++k)
        if (b + k < tiles) s += row[b + k];
    uint32_t total;
    uint32_t run = block_exclusive_scan<kScanThreads>(s, s_wave, &total);
    for (uint32_t k = 0; k < per; ++k) {
        if (b + k < tiles) {
            const uint32_t c = row[b + k];
            row[b + k] = run;
            run += c;
        }
    }
    if (threadIdx.x == 0) tot[blockIdx.x] = total;
}

// (3) Ranks inside the tile come from wave64 match-any ballots and per-wave
// LDS counters; a key's output index is the digit's global base (scan of the
// digit totals) + the tile's offset within the digit (pass 2) + its rank in
// the tile. The tile is reordered in LDS and written out in runs per digit.
__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint64_t* __restrict__ kout,
    uint32_t* __restrict__ vout, uint64_t n, int shift, const uint32_t* __restrict__ cnt,
    uint32_t tiles, const uint32_t* __restrict__ tot, uint32_t* err) {
    constexpr int kWaves = kSortThreads / 64;
    __shared__ uint32_t s_wcnt[kWaves][256];
    __shared__ uint32_t s_lstart[256];
    __shared__ uint32_t s_gbase[256];
    __shared__ uint32_t s_scan[kWaves];
    __shared__ uint64_t s_keys[kSortTile];
    __shared__ uint32_t s_vals[kSortTile];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < kWaves * 256; i += kSortThreads) (&s_wcnt[0][0])[i] = 0;
    __syncthreads();
    const uint32_t tile = blockIdx.x;
    const uint64_t base = (uint64_t)tile * kSortTile;

    uint64_t key[kSortItems];
    uint32_t val[kSortItems], dig[kSortItems], rank[kSortItems];
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const uint64_t idx = base + (uint64_t)wave * 64 * kSortItems + k * 64 + lane;
        const bool valid = idx < n;
        key[k] = valid ? kin[idx] : 0ull;
        val[k] = valid ? vin[idx] : 0u;
        dig[k] = valid ? (uint32_t)((key[k] >> shift) & 255u) : 256u;
    }
    const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const uint32_t d = dig[k];
        const bool valid = d < 256u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            peers &= bit ? bb : ~bb;
        }
        if (!valid) peers = 1ull << lane;
        const int leader = __ffsll((long long)peers) - 1;
        uint32_t old = 0;
        if (valid && lane == leader) {
            old = s_wcnt[wave][d];
            s_wcnt[wave][d] = old + (uint32_t)__popcll(peers);
        }
        old = __shfl(old, leader);
        rank[k] = old + (uint32_t)__popcll(peers & lt);
    }
    __syncthreads();
    // Per digit (the first 256 threads): exclusive offsets over waves, tile count.
    const uint32_t d = tid;
    uint32_t c = 0, t = 0;
    if (tid < 256) {
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t cw = s_wcnt[w][d];
            s_wcnt[w][d] = c;
            c += cw;
        }
        t = tot[d];
    }
    uint32_t dummy;
    const uint32_t lstart = block_exclusive_scan<kSortThreads>(c, s_scan, &dummy);
    const uint32_t gbase = block_exclusive_scan<kSortThreads>(t, s_scan, &dummy);
    if (tid < 256) {
        s_lstart[d] = lstart;
        s_gbase[d] = gbase + cnt[(uint64_t)d * tiles + tile] - lstart;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        if (dig[k] < 256u) {
            const uint32_t p = s_lstart[dig[k]] + s_wcnt[wave][dig[k]] + rank[k];
            s_keys[p] = key[k];
            s_vals[p] = val[k];
        }
    }
    __syncthreads();
    const uint32_t valid_n = (uint32_t)min((uint64_t)kSortTile, n - base);
    for (uint32_t p = tid; p < valid_n; p += kSortThreads) {
        const uint64_t k = s_keys[p];
        const uint32_t o = s_gbase[(k >> shift) & 255u] + p;
        if (o < n) {
            kout[o] = k;
            vout[o] = s_vals[p];
        } else {
            atomicOr(err, kErrCapacity);
        }
    }
}

// The reference's tree keeps splitting bodies that share every level a key
// holds (octree.rs:92-125, down to depth 64). After the sort, each run of
// equal keys (in practice two or three bodies of a dense core) gets the
// levels past the first word of the same descent as deep keys, and is put
// in (deep key, index) order by the run's first thread (insertion sort).
// Bodies that share more than kMaxLcp levels, or a run longer than
// kMaxTieRun, are a depth error (the reference panics past depth 64).
template <int DIM>
__device__ __forceinline__ bool deep_less(const DeepKey& a, uint32_t va, const DeepKey& b,
                                          uint32_t vb) {
    for (int i = 0; i < kDeepWords; ++i)
        if (a.w[i] != b.w[i]) return a.w[i] < b.w[i];
    return va < vb;
}
template <int DIM>
__device__ void order_tie_run(const SegDesc* __restrict__ desc, uint64_t total,
                              const uint64_t* __restrict__ skeys, uint32_t* __restrict__ svals,
                              DeepKey* __restrict__ deep, uint32_t nseg, int segbits, int lmax,
                              uint64_t g) {
    if (g + 1 >= total) return;
    const uint64_t key = skeys[g];
    if (skeys[g + 1] != key || (g > 0 && skeys[g - 1] == key)) return;  // run heads only
    const uint32_t s = seg_of_key(key, segbits);
    if (s >= nseg) return;
    const SegDesc& d = desc[s];
    uint64_t r = g + 1;
    while (r + 1 < total && skeys[r + 1] == key && r + 1 - g < (uint64_t)kMaxTieRun) ++r;
    if (r + 1 < total && skeys[r + 1] == key) {
        atomicOr(&d.st->err, kErrDepth);
        return;
    }
    const double E = bits_f64(d.st->extent_bits);
    constexpr int W = kWordLevels<DIM>;
    // the descent down to the first level no two bodies may share
    const int deepest = min(kMaxLcp + 1, lmax + kDeepWords * W);
    for (uint64_t j = g; j <= r; ++j) {
        const uint64_t src = (uint64_t)svals[j] - d.off;
        if (src >= d.n) {
            atomicOr(&d.st->err, kErrCapacity);
            return;
        }
        const double4 q = d.spos[src];
        const double px = q.x, py = q.y, pz = q.z;
        double ext = E, cx = 0.0, cy = 0.0, cz = 0.0;
        DeepKey dk;
        for (int i = 0; i < kDeepWords; ++i) dk.w[i] = 0;
        for (int l = 0; l < deepest; ++l) {
            uint32_t o = (uint32_t)(px > cx) | ((uint32_t)(py > cy) << 1);
            if (DIM == 3) o |= (uint32_t)(pz > cz) << 2;
            const double h2 = half(ext);
            cx = (o & 1u) ? cx + h2 : cx - h2;
            cy = (o & 2u) ? cy + h2 : cy - h2;
            if (DIM == 3) cz = (o & 4u) ? cz + h2 : cz - h2;
            ext = h2;
            if (l >= lmax) {
                const int t = l - lmax;
                dk.w[t / W] |= (uint64_t)o << (DIM * (W - 1 - t % W));
            }
        }
        deep[j] = dk;
    }
    for (uint64_t j = g + 1; j <= r; ++j) {
        const DeepKey kv = deep[j];
        const uint32_t vv = svals[j];
        uint64_t m = j;
        while (m > g && deep_less<DIM>(kv, vv, deep[m - 1], svals[m - 1])) {
            deep[m] = deep[m - 1];
            svals[m] = svals[m - 1];
            --m;
        }
        deep[m] = kv;
        svals[m] = vv;
    }
}

// Tie runs after the radix sort (the sample sort's bucket blocks order their
// own).
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_ties(const SegDesc* __restrict__ desc, uint64_t total,
                                                 const uint64_t* __restrict__ skeys,
                                                 uint32_t* __restrict__ svals,
                                                 DeepKey* __restrict__ deep, uint32_t nseg,
                                                 int segbits, int lmax) {
    order_tie_run<DIM>(desc, total, skeys, svals, deep, nseg, segbits, lmax,
                       (uint64_t)blockIdx.x * kBlock + threadIdx.x);
}

// ------------------------------------------------------------ sample sort
// The keys arrive in the previous step's Morton order, so they are nearly
// sorted, and keys taken every kBucketKeys positions split them into buckets
// of about kBucketKeys keys each. (1) k_splitters: one block takes the
// quantiles of the batch's previous sorted keys (exact for the previous
// step, so buckets stay near kBucketKeys as bodies move), or, on a batch's
// first step, sorts kSampleOver evenly spaced keys per bucket and keeps
// every kSampleOver-th; (2) k_bucket_count: each key's bucket (the
// number of splitters <= key) and the bucket sizes; (3) k_bucket_scatter:
// each tile reserves a range in every bucket it holds and writes its keys
// there, in runs; (4) k_bucket_sort: each bucket is sorted by key (a
// larger one in HBM, by one block). Runs of equal keys then take (second key
// word, index) order in k_ties, the order a stable sort by key gives.

// Ascending bitonic sort of s[0, p2) in LDS (p2 a power of two) by a block of
// NT threads; the caller has synchronised the block after filling s.
template <int NT>
__device__ __forceinline__ void lds_bitonic_sort(uint64_t* s, uint32_t p2) {
    for (uint32_t k = 2; k <= p2; k <<= 1) {
        for (uint32_t h = k >> 1; h > 0; h >>= 1) {
            for (uint32_t i = threadIdx.x; i < p2 / 2; i += NT) {
                const uint32_t a = ((i & ~(h - 1)) << 1) | (i & (h - 1)), b = a + h;
                const uint64_t x = s[a], y = s[b];
                if ((x > y) == ((a & k) == 0)) {
                    s[a] = y;
                    s[b] = x;
                }
            }
            __syncthreads();
        }
    }
}

__global__ __launch_bounds__(kSplitThreads) void k_splitters(
    const uint64_t* __restrict__ prev, const uint64_t* __restrict__ keys, uint64_t n, uint32_t nb,
    uint64_t* __restrict__ spl, uint32_t* __restrict__ bcount, uint32_t* __restrict__ bfill,
    StepCounters* __restrict__ sc) {
    __shared__ uint64_t s[kMaxBuckets * kSampleOver];
    __shared__ uint32_t s_unsorted;
    for (uint32_t b = threadIdx.x; b < nb; b += kSplitThreads) {
        bcount[b] = 0;
        bfill[b] = 0;
    }
    const bool prev_ok = sc->split_prev != 0;
    if (threadIdx.x == 0) sc->quantiles = 0;
    if (nb <= 1) return;
    if (prev_ok) {
        // quantiles of the previous step's sorted keys, if they are in order
        if (threadIdx.x == 0) s_unsorted = 0;
        for (uint32_t b = threadIdx.x; b + 1 < nb; b += kSplitThreads)
            s[b] = prev[muldiv_floor(b + 1, n, nb)];
        __syncthreads();
        bool bad = false;
        for (uint32_t b = threadIdx.x; b + 2 < nb; b += kSplitThreads) bad |= s[b] > s[b + 1];
        if (bad) s_unsorted = 1;
        __syncthreads();
        if (!s_unsorted) {
            for (uint32_t b = threadIdx.x; b + 1 < nb; b += kSplitThreads) spl[b] = s[b];
            if (threadIdx.x == 0) sc->quantiles = 1;
            return;
        }
        __syncthreads();
    }
    // kSampleOver evenly spaced keys per bucket, padded to a power of two,
    // bitonic-sorted; every kSampleOver-th is a splitter
    const uint32_t m = nb * kSampleOver;
    uint32_t p2 = 1;
    while (p2 < m) p2 <<= 1;
    for (uint32_t j = threadIdx.x; j < p2; j += kSplitThreads)
        s[j] = j < m ? keys[muldiv_floor(j, n, m)] : ~0ull;
    __syncthreads();
    lds_bitonic_sort<kSplitThreads>(s, p2);
    for (uint32_t b = threadIdx.x; b + 1 < nb; b += kSplitThreads) spl[b] = s[(b + 1) * kSampleOver];
}

// With the previous step's quantiles as splitters, it also copies kQuantiles
// finer ones per bucket (bucket b's i-th: prev[(b kQuantiles + i) n / (nb
// kQuantiles)], the first being the bucket's lower splitter) for the slot sort.
__global__ __launch_bounds__(kSortThreads) void k_bucket_count(
    const uint64_t* __restrict__ keys, uint64_t n, const uint64_t* __restrict__ spl, uint32_t nb,
    uint16_t* __restrict__ bkt, uint32_t* __restrict__ bcount, const uint64_t* __restrict__ prev,
    uint64_t* __restrict__ quant, const StepCounters* __restrict__ sc) {
    __shared__ uint64_t s_spl[kMaxBuckets];
    __shared__ uint32_t s_h[kMaxBuckets];
    // the independent loads first: the tile's keys, the splitters, the flag
    const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
    uint64_t key[kSortItems];
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const uint64_t idx = base + (uint64_t)k * kSortThreads + threadIdx.x;
        key[k] = idx < n ? keys[idx] : 0ull;
    }
    const bool copy_quant = sc->quantiles != 0;
    for (uint32_t b = threadIdx.x; b < nb; b += kSortThreads) {
        s_h[b] = 0;
        if (b + 1 < nb) s_spl[b] = spl[b];
    }
    if (copy_quant) {
        const uint64_t m = (uint64_t)nb * kQuantiles;
        for (uint64_t t = (uint64_t)blockIdx.x * kSortThreads + threadIdx.x; t < m;
             t += (uint64_t)gridDim.x * kSortThreads)
            quant[t] = prev[muldiv_floor(t, n, m)];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const uint64_t idx = base + (uint64_t)k * kSortThreads + threadIdx.x;
        if (idx < n) {
            const uint32_t b = bucket_of(s_spl, nb, key[k]);
            bkt[idx] = (uint16_t)b;
            atomicAdd(&s_h[b], 1u);
        }
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < nb; b += kSortThreads)
        if (s_h[b]) atomicAdd(&bcount[b], s_h[b]);
}

__global__ __launch_bounds__(kSortThreads) void k_bucket_scatter(
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    const uint16_t* __restrict__ bkt, uint64_t n, const uint32_t* __restrict__ bcount,
    uint32_t* __restrict__ bfill, uint32_t* __restrict__ bbase, uint32_t nb,
    uint64_t* __restrict__ kout, uint32_t* __restrict__ vout) {
    constexpr int kWaves = kSortThreads / 64;
    static_assert(kMaxBuckets == 2 * kSortThreads, "two buckets per thread in the scans");
    // per wave and bucket counts, then starts (at most a tile: 16 bits), so
    // three blocks fit in a CU
    static_assert(kSortTile < 65536, "16-bit tile offsets");
    __shared__ __align__(16) uint16_t s_wcnt[kWaves][kMaxBuckets];
    __shared__ uint16_t s_lstart[kMaxBuckets];
    __shared__ uint32_t s_dst[kMaxBuckets];
    __shared__ uint32_t s_scan[kWaves];
    __shared__ uint64_t s_keys[kSortTile];
    __shared__ uint32_t s_vals[kSortTile];
    __shared__ uint16_t s_b[kSortTile];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // the bucket sizes of this thread's two buckets, loaded early
    const uint32_t b0 = 2u * (uint32_t)tid, b1 = b0 + 1u;
    const uint32_t n0 = b0 < nb ? bcount[b0] : 0u, n1 = b1 < nb ? bcount[b1] : 0u;
    for (int i = tid; i < kWaves * kMaxBuckets / 2; i += kSortThreads)
        reinterpret_cast<uint32_t*>(&s_wcnt[0][0])[i] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
    uint64_t key[kSortItems];
    uint32_t val[kSortItems], bk[kSortItems], rank[kSortItems];
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const uint64_t idx = base + (uint64_t)wave * 64 * kSortItems + k * 64 + lane;
        const bool valid = idx < n;
        key[k] = valid ? kin[idx] : 0ull;
        val[k] = valid ? vin[idx] : 0u;
        bk[k] = valid ? (uint32_t)bkt[idx] : kMaxBuckets;
    }
    // rank within the wave among equal buckets (match-any over 10 bits; the
    // keys arrive nearly sorted, so a wave's keys mostly share one bucket)
    const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const uint32_t b = bk[k];
        const bool valid = b < (uint32_t)kMaxBuckets;
        uint64_t peers = __ballot(valid);
        const uint32_t bf = (uint32_t)__builtin_amdgcn_readlane(
            (int)b, peers ? __ffsll((long long)peers) - 1 : 0);
        if (__ballot(valid && b != bf)) {
#pragma unroll
            for (int bit = 0; bit < 10; ++bit) {
                const bool on = (b >> bit) & 1u;
                const uint64_t bb = __ballot(on);
                peers &= on ? bb : ~bb;
            }
        }
        if (!valid) peers = 1ull << lane;
        const int leader = __ffsll((long long)peers) - 1;
        uint32_t old = 0;
        if (valid && lane == leader) {
            old = s_wcnt[wave][b];
            s_wcnt[wave][b] = (uint16_t)(old + (uint32_t)__popcll(peers));
        }
        old = __shfl(old, leader);
        rank[k] = old + (uint32_t)__popcll(peers & lt);
    }
    __syncthreads();
    // buckets 2t and 2t + 1: counts over the waves, tile-local starts,
    // global bucket bases, and the reserved output ranges
    uint32_t c0 = 0, c1 = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        const uint32_t x0 = s_wcnt[w][b0], x1 = s_wcnt[w][b1];
        s_wcnt[w][b0] = (uint16_t)c0;
        s_wcnt[w][b1] = (uint16_t)c1;
        c0 += x0;
        c1 += x1;
    }
    // the tile's ranges in its buckets, reserved before the scans so that the
    // atomics' round trip overlaps them (each result is used only under its
    // own condition: no wait where the branches join)
    uint32_t f0, f1;
    if (c0) f0 = atomicAdd(&bfill[b0], c0);
    if (c1) f1 = atomicAdd(&bfill[b1], c1);
    uint32_t dummy;
    const uint32_t l0 = block_exclusive_scan<kSortThreads>(c0 + c1, s_scan, &dummy);
    const uint32_t g0 = block_exclusive_scan<kSortThreads>(n0 + n1, s_scan, &dummy);
    s_lstart[b0] = (uint16_t)l0;
    s_lstart[b1] = (uint16_t)(l0 + c0);
    s_dst[b0] = 0u;
    s_dst[b1] = 0u;
    if (c0) s_dst[b0] = g0 + f0;
    if (c1) s_dst[b1] = g0 + n0 + f1;
    if (blockIdx.x == 0) {
        if (b0 < nb) bbase[b0] = g0;
        if (b1 < nb) bbase[b1] = g0 + n0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const uint32_t b = bk[k];
        if (b < (uint32_t)kMaxBuckets) {
            const uint32_t p = s_lstart[b] + s_wcnt[wave][b] + rank[k];
            s_keys[p] = key[k];
            s_vals[p] = val[k];
            s_b[p] = (uint16_t)b;
        }
    }
    __syncthreads();
    const uint32_t valid_n = (uint32_t)min((uint64_t)kSortTile, n - base);
    for (uint32_t p = tid; p < valid_n; p += kSortThreads) {
        const uint32_t b = s_b[p];
        const uint32_t o = s_dst[b] + (p - s_lstart[b]);
        kout[o] = s_keys[p];
        vout[o] = s_vals[p];
    }
}

// Stable ranks of a chunk of up to kBucketChunk keys by an 8-bit digit (256
// = no key), one block of kBucketThreads threads, item k of thread (wave,
// lane) at chunk index wave * 256 + k * 64 + lane: pos[k] is the item's index
// in the chunk ordered by digit; s_dstart[d] and s_dcnt[d] are the chunk's
// start and count of digit d.
constexpr int kBucketChunk = 4 * kBucketThreads;
__device__ __forceinline__ void chunk_rank(const uint32_t (&dig)[4], uint32_t (&pos)[4],
                                           uint32_t (*s_wcnt)[256], uint32_t* s_dstart,
                                           uint32_t* s_dcnt, uint32_t* s_scan) {
    constexpr int kWaves = kBucketThreads / 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < kWaves * 256; i += kBucketThreads) (&s_wcnt[0][0])[i] = 0;
    __syncthreads();
    const uint64_t lt = (1ull << lane) - 1ull;
    uint32_t rank[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t d = dig[k];
        const bool valid = d < 256u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            peers &= bit ? bb : ~bb;
        }
        if (!valid) peers = 1ull << lane;
        const int leader = __ffsll((long long)peers) - 1;
        uint32_t old = 0;
        if (valid && lane == leader) {
            old = s_wcnt[wave][d];
            s_wcnt[wave][d] = old + (uint32_t)__popcll(peers);
        }
        old = __shfl(old, leader);
        rank[k] = old + (uint32_t)__popcll(peers & lt);
    }
    __syncthreads();
    uint32_t c = 0;
    if (tid < 256) {
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t x = s_wcnt[w][tid];
            s_wcnt[w][tid] = c;
            c += x;
        }
        s_dcnt[tid] = c;
    }
    uint32_t dummy;
    const uint32_t st = block_exclusive_scan<kBucketThreads>(c, s_scan, &dummy);
    if (tid < 256) s_dstart[tid] = st;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k)
        pos[k] = dig[k] < 256u ? s_dstart[dig[k]] + s_wcnt[wave][dig[k]] + rank[k] : 0u;
}

// ---------------------------------------------- sample sort, many buckets
// With more than kMaxBuckets buckets (up to kMaxWideBuckets: a batch of 8.8M
// bodies takes about 5,300) the splitters no longer fit every key tile's LDS
// (k_keys and k_bucket_scatter hold a slot per bucket). They stay in HBM,
// behind a coarse table of every kWideGroup-th splitter that does fit: a
// key's bucket is a search of the coarse table in LDS, then three halvings
// over the splitters in L2. A tile ranks its keys by bucket with the stable
// chunk rank (the bucket less the tile's lowest, 8 bits a pass) rather than
// through a per-bucket table. The kernels: k_splitters_wide,
// k_bucket_count_wide, k_bucket_base, k_bucket_scatter_wide, then
// k_bucket_sort as for fewer buckets; the order is the same.

// (1) One block: the splitters (spl[0, nb - 1)) and the coarse table
// (spc[i] = spl[kWideGroup i + kWideGroup - 1] for i < (nb - 1) /
// kWideGroup), from the previous step's sorted keys where they are this
// batch's and in order, else from kWideSampleOver sorted samples per bucket.
// Also clears the bucket counters.
__global__ __launch_bounds__(kSplitThreads) void k_splitters_wide(
    const uint64_t* __restrict__ prev, const uint64_t* __restrict__ keys, uint64_t n, uint32_t nb,
    uint64_t* __restrict__ spl, uint64_t* __restrict__ spc, uint32_t* __restrict__ bcount,
    uint32_t* __restrict__ bfill, StepCounters* __restrict__ sc, int fused) {
    static_assert(kMaxWideBuckets * kWideSampleOver * 8 <= 160 * 1024, "samples in one CU's LDS");
    __shared__ uint64_t s[kMaxWideBuckets * kWideSampleOver];
    __shared__ uint32_t s_unsorted;
    for (uint32_t b = threadIdx.x; b < nb; b += kSplitThreads) {
        bcount[b] = 0;
        bfill[b] = 0;
    }
    // (fused: this step's keys are not built yet and k_keys takes the
    // quantiles; the caller knows the previous step's sorted keys are there)
    const bool prev_ok = fused || sc->split_prev != 0;
    if (threadIdx.x == 0) {
        if (!fused) sc->quantiles = 0;
        s_unsorted = 0;
    }
    const uint32_t ncoarse = (nb - 1) / kWideGroup;
    __syncthreads();
    if (prev_ok) {
        for (uint32_t b = threadIdx.x; b + 1 < nb; b += kSplitThreads)
            s[b] = prev[muldiv_floor(b + 1, n, nb)];
        __syncthreads();
        bool bad = false;
        if (!fused)
            for (uint32_t b = threadIdx.x; b + 2 < nb; b += kSplitThreads) bad |= s[b] > s[b + 1];
        if (bad) s_unsorted = 1;
        __syncthreads();
        if (!s_unsorted) {
            for (uint32_t b = threadIdx.x; b + 1 < nb; b += kSplitThreads) spl[b] = s[b];
            for (uint32_t i = threadIdx.x; i < ncoarse; i += kSplitThreads)
                spc[i] = s[i * kWideGroup + kWideGroup - 1];
            if (threadIdx.x == 0 && !fused) sc->quantiles = 1;
            return;
        }
        __syncthreads();
    }
    const uint32_t m = nb * kWideSampleOver;
    uint32_t p2 = 1;
    while (p2 < m) p2 <<= 1;
    for (uint32_t j = threadIdx.x; j < p2; j += kSplitThreads)
        s[j] = j < m ? keys[muldiv_floor(j, n, m)] : ~0ull;
    __syncthreads();
    lds_bitonic_sort<kSplitThreads>(s, p2);
    for (uint32_t b = threadIdx.x; b + 1 < nb; b += kSplitThreads)
        spl[b] = s[(b + 1) * kWideSampleOver];
    for (uint32_t i = threadIdx.x; i < ncoarse; i += kSplitThreads)
        spc[i] = s[(i + 1) * kWideGroup * kWideSampleOver];
}

// (2) Each key's bucket (the number of splitters <= key) and the bucket
// sizes. The keys arrive in the previous step's order, so a tile's buckets
// lie mostly within a window from its first key's: counts in that window go
// through LDS, the rest (bodies that moved far) straight to HBM. Copies the
// finer quantiles for the slot sort as k_bucket_count does.
__global__ __launch_bounds__(kSortThreads) void k_bucket_count_wide(
    const uint64_t* __restrict__ keys, uint64_t n, const uint64_t* __restrict__ spl,
    const uint64_t* __restrict__ spc, uint32_t nb, uint16_t* __restrict__ bkt,
    uint32_t* __restrict__ bcount, const uint64_t* __restrict__ prev,
    uint64_t* __restrict__ quant, const StepCounters* __restrict__ sc) {
    constexpr uint32_t kCoarseMax = kMaxWideBuckets / kWideGroup;
    constexpr uint32_t kWindow = 1024;
    static_assert(kMaxWideBuckets <= 65536, "16-bit bucket numbers");
    __shared__ uint64_t s_c[kCoarseMax];
    __shared__ uint32_t s_h[kWindow];
    __shared__ uint32_t s_wb;
    const int tid = threadIdx.x;
    const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
    uint64_t key[kSortItems];
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const uint64_t idx = base + (uint64_t)k * kSortThreads + tid;
        key[k] = idx < n ? keys[idx] : 0ull;
    }
    const bool copy_quant = sc->quantiles != 0;
    const uint32_t nc = (nb - 1) / kWideGroup;
    for (uint32_t i = tid; i < nc; i += kSortThreads) s_c[i] = spc[i];
    for (uint32_t i = tid; i < kWindow; i += kSortThreads) s_h[i] = 0;
    if (copy_quant) {
        const uint64_t m = (uint64_t)nb * kQuantiles;
        for (uint64_t t = (uint64_t)blockIdx.x * kSortThreads + tid; t < m;
             t += (uint64_t)gridDim.x * kSortThreads)
            quant[t] = prev[muldiv_floor(t, n, m)];
    }
    __syncthreads();
    uint32_t b[kSortItems];
    buckets_of_wide<kSortItems>(s_c, spl, nb, key, b);
    // the window starts a quarter of it below the tile's first key's bucket
    if (tid == 0) s_wb = b[0] > kWindow / 4 ? b[0] - kWindow / 4 : 0u;
    __syncthreads();
    const uint32_t wb = s_wb;
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const uint64_t idx = base + (uint64_t)k * kSortThreads + tid;
        if (idx < n) {
            bkt[idx] = (uint16_t)b[k];
            const uint32_t r = b[k] - wb;  // (wraps past kWindow below the window)
            if (r < kWindow) atomicAdd(&s_h[r], 1u);
            else atomicAdd(&bcount[b[k]], 1u);
        }
    }
    __syncthreads();
    for (uint32_t r = tid; r < kWindow; r += kSortThreads)
        if (s_h[r]) atomicAdd(&bcount[wb + r], s_h[r]);
}

// (3) One block: each bucket's first key (exclusive scan of the sizes).
__global__ __launch_bounds__(kSplitThreads) void k_bucket_base(const uint32_t* __restrict__ bcount,
                                                               uint32_t nb,
                                                               uint32_t* __restrict__ bbase) {
    constexpr int kPer = kMaxWideBuckets / kSplitThreads;
    static_assert(kPer * kSplitThreads == kMaxWideBuckets, "whole buckets per thread");
    __shared__ uint32_t s_scan[kSplitThreads / 64];
    const uint32_t b0 = threadIdx.x * (uint32_t)kPer;
    uint32_t c[kPer], sum = 0;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        c[j] = b0 + j < nb ? bcount[b0 + j] : 0u;
        sum += c[j];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan<kSplitThreads>(sum, s_scan, &total);
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        if (b0 + j < nb) bbase[b0 + j] = run;
        run += c[j];
    }
}

// (4) Each tile ranks its keys by bucket in LDS, reserves one range per run
// of a bucket (bfill, which the bucket sort clears) and writes the runs.
__global__ __launch_bounds__(kSortThreads) void k_bucket_scatter_wide(
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    const uint16_t* __restrict__ bkt, uint64_t n, const uint32_t* __restrict__ bbase,
    uint32_t* __restrict__ bfill, uint64_t* __restrict__ kout, uint32_t* __restrict__ vout) {
    static_assert(kSortThreads == kBucketThreads && kSortTile == kBucketChunk,
                  "a tile is one chunk of the chunk rank");
    static_assert(kSortTile < 65536, "16-bit tile positions");
    constexpr int kWaves = kSortThreads / 64;
    __shared__ uint64_t s_keys[kSortTile];
    __shared__ uint32_t s_vals[kSortTile];
    __shared__ uint16_t s_b[kSortTile];
    __shared__ uint16_t s_rid[kSortTile];        // each position's run
    __shared__ uint16_t s_rhead[kSortTile + 1];  // each run's first position (+ the end)
    __shared__ __align__(16) uint32_t s_rank[kWaves * 256];  // the chunk rank's counts, then
                                                              // each run's destination
    __shared__ uint32_t s_dstart[256], s_dcnt[256], s_scan[kWaves];
    __shared__ uint64_t s_red[kWaves];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
    const uint32_t valid_n = (uint32_t)min((uint64_t)kSortTile, n - base);
    uint64_t key[4];
    uint32_t val[4], bk[4], dig[4], pos[4];
    uint32_t lo = 0xFFFFu, hi = 0u;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t ci = (uint32_t)wave * 256u + (uint32_t)k * 64u + (uint32_t)lane;
        const bool valid = ci < valid_n;
        key[k] = valid ? kin[base + ci] : 0ull;
        val[k] = valid ? vin[base + ci] : 0u;
        bk[k] = valid ? (uint32_t)bkt[base + ci] : 0xFFFFu;
        lo = valid ? min(lo, bk[k]) : lo;
        hi = valid ? max(hi, bk[k]) : hi;
    }
    const uint32_t bmin = 0xFFFFu - (uint32_t)block_max_u64<kSortThreads>(0xFFFFu - lo, s_red);
    const uint32_t bmax = (uint32_t)block_max_u64<kSortThreads>(hi, s_red);
    uint32_t(*s_wcnt)[256] = reinterpret_cast<uint32_t(*)[256]>(s_rank);
    // stable ranks by the low 8 bits of (bucket - bmin), then, where the
    // tile spans 256 buckets or more, by the high 8
#pragma unroll
    for (int k = 0; k < 4; ++k)
        dig[k] = bk[k] != 0xFFFFu ? (bk[k] - bmin) & 255u : 256u;
    chunk_rank(dig, pos, s_wcnt, s_dstart, s_dcnt, s_scan);
    if (bmax - bmin >= 256u) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (dig[k] < 256u) {
                s_keys[pos[k]] = key[k];
                s_vals[pos[k]] = val[k];
                s_b[pos[k]] = (uint16_t)bk[k];
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t ci = (uint32_t)wave * 256u + (uint32_t)k * 64u + (uint32_t)lane;
            const bool valid = ci < valid_n;
            key[k] = valid ? s_keys[ci] : 0ull;
            val[k] = valid ? s_vals[ci] : 0u;
            bk[k] = valid ? (uint32_t)s_b[ci] : 0xFFFFu;
            dig[k] = valid ? (bk[k] - bmin) >> 8 : 256u;
        }
        // (the rank's first barrier orders these reads before the writes below)
        chunk_rank(dig, pos, s_wcnt, s_dstart, s_dcnt, s_scan);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (dig[k] < 256u) {
            s_keys[pos[k]] = key[k];
            s_vals[pos[k]] = val[k];
            s_b[pos[k]] = (uint16_t)bk[k];
        }
    }
    __syncthreads();
    // runs of one bucket: thread t flags the run heads among positions
    // 4t .. 4t + 3, a scan numbers them
    uint32_t head[4], heads = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t p = 4u * (uint32_t)tid + (uint32_t)j;
        head[j] = p < valid_n && (p == 0 || s_b[p] != s_b[p - 1]) ? 1u : 0u;
        heads += head[j];
    }
    uint32_t nruns;
    uint32_t rid = block_exclusive_scan<kSortThreads>(heads, s_scan, &nruns);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t p = 4u * (uint32_t)tid + (uint32_t)j;
        if (p < valid_n) {
            if (head[j]) s_rhead[rid++] = (uint16_t)p;
            s_rid[p] = (uint16_t)(rid - 1u);
        }
    }
    if (tid == 0) s_rhead[nruns] = (uint16_t)valid_n;
    __syncthreads();
    uint32_t* s_rdst = s_rank;
    for (uint32_t r = tid; r < nruns; r += kSortThreads) {
        const uint32_t h = s_rhead[r], b = s_b[h];
        s_rdst[r] = bbase[b] + atomicAdd(&bfill[b], (uint32_t)s_rhead[r + 1] - h);
    }
    __syncthreads();
    for (uint32_t p = tid; p < valid_n; p += kSortThreads) {
        const uint32_t r = s_rid[p];
        const uint32_t o = s_rdst[r] + (p - (uint32_t)s_rhead[r]);
        kout[o] = s_keys[p];
        vout[o] = s_vals[p];
    }
}

// Highest bit in which any of the block's keys differs from `ref` (-1: none).
__device__ __forceinline__ int block_top_diff_bit(uint64_t x, uint64_t* s_or) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x |= __shfl_xor(x, o);
    if (lane == 0) s_or[wave] = x;
    __syncthreads();
    uint64_t a = 0;
#pragma unroll
    for (int w = 0; w < kBucketThreads / 64; ++w) a |= s_or[w];
    __syncthreads();
    return a ? 63 - __clzll((long long)a) : -1;
}

// Bitonic sort of P = E * kBucketThreads (key, value) pairs, E per thread in
// registers (element i = thread * E + e): exchanges between a thread's own
// elements in registers, between threads of a wave by shuffles, and the
// widest ones through LDS (s_k / s_v, element i at (i % E) * kBucketThreads +
// i / E, conflict-free). Ascending in (key, value).
// The value of lane (lane ^ M) of a wave, on the VALU instead of a
// bpermute through the LDS unit: DPP quad_perm for M = 1, 2; row_ror:n
// (lane i of a row of 16 takes lane i - n mod 16) for M = 8, and for M = 4
// by lane (row_ror:4 where bit 2 is set, row_ror:12 where it is clear);
// permlane16/32_swap of the value with itself for M = 16, 32 (the swap
// leaves the partner row / half in one of its two results).
template <int M>
__device__ __forceinline__ uint32_t xor_lanes(uint32_t v) {
    static_assert(M == 1 || M == 2 || M == 4 || M == 8 || M == 16 || M == 32, "lane distance");
    const uint32_t lane = threadIdx.x & 63u;
    if (M == 1) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);
    if (M == 2) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false);
    if (M == 4) {
        const uint32_t r4 = (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x124, 0xF, 0xF, false);
        const uint32_t r12 = (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x12C, 0xF, 0xF, false);
        return (lane & 4u) ? r4 : r12;
    }
    if (M == 8) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x128, 0xF, 0xF, false);
    if (M == 16) {
        const auto p = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        return (lane & 16u) ? p[0] : p[1];
    }
    const auto p = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return (lane & 32u) ? p[0] : p[1];
}

template <int M>
__device__ __forceinline__ uint64_t xor_lanes64(uint64_t v) {
    return (uint64_t)xor_lanes<M>((uint32_t)v) | ((uint64_t)xor_lanes<M>((uint32_t)(v >> 32)) << 32);
}

// FULL: order by (key, value); else by key alone (equal keys in no fixed
// order), which needs no key to equal the padding ~0.
template <bool FULL>
__device__ __forceinline__ bool pair_gt(uint64_t ka, uint32_t va, uint64_t kb, uint32_t vb) {
    return FULL ? (ka > kb || (ka == kb && va > vb)) : ka > kb;
}

// Compare-exchange with a partner held elsewhere: keep the smaller (keep_min)
// or the larger of the two, by selects.
template <bool FULL>
__device__ __forceinline__ void bitonic_keep(uint64_t& k, uint32_t& v, uint64_t ok, uint32_t ov,
                                             bool keep_min) {
    const bool take = keep_min ? pair_gt<FULL>(k, v, ok, ov) : pair_gt<FULL>(ok, ov, k, v);
    k = take ? ok : k;
    v = take ? ov : v;
}

template <bool FULL, int E, int H>
__device__ __forceinline__ void bitonic_regs(uint64_t (&key)[E], uint32_t (&val)[E], uint32_t k) {
    const uint32_t t0 = threadIdx.x * E;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        if (e & H) continue;
        const bool up = ((t0 + (uint32_t)e) & k) == 0;
        const uint64_t ka = key[e], kb = key[e | H];
        const uint32_t va = val[e], vb = val[e | H];
        const bool sw = pair_gt<FULL>(ka, va, kb, vb) == up;
        key[e] = sw ? kb : ka;
        key[e | H] = sw ? ka : kb;
        val[e] = sw ? vb : va;
        val[e | H] = sw ? va : vb;
    }
}

// One bitonic stage against the partner thread t ^ M of the same wave.
template <bool FULL, int E, int M>
__device__ __forceinline__ void bitonic_wave(uint64_t (&key)[E], uint32_t (&val)[E], bool keep_min) {
    uint64_t ok[E];
    uint32_t ov[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        ok[e] = xor_lanes64<M>(key[e]);
        ov[e] = xor_lanes<M>(val[e]);
    }
#pragma unroll
    for (int e = 0; e < E; ++e) bitonic_keep<FULL>(key[e], val[e], ok[e], ov[e], keep_min);
}

template <bool FULL, int E>
__device__ void bitonic_block(uint64_t (&key)[E], uint32_t (&val)[E], uint64_t* s_k,
                              uint32_t* s_v) {
    constexpr uint32_t P = (uint32_t)E * kBucketThreads;
    const uint32_t t = threadIdx.x;
    for (uint32_t k = 2; k <= P; k <<= 1) {
        for (uint32_t h = k >> 1; h > 0; h >>= 1) {
            if (h >= (uint32_t)E) {
                // k > h >= E: the direction is the same for a thread's elements
                const bool up = ((t * E) & k) == 0;
                const uint32_t m = h / E;
                const bool keep_min = ((t & m) == 0) == up;
                if (h >= 64u * E) {
                    // partner in another wave: through LDS
#pragma unroll
                    for (int e = 0; e < E; ++e) {
                        s_k[(uint32_t)e * kBucketThreads + t] = key[e];
                        s_v[(uint32_t)e * kBucketThreads + t] = val[e];
                    }
                    __syncthreads();
                    const uint32_t pt = t ^ m;
                    uint64_t ok[E];
                    uint32_t ov[E];
#pragma unroll
                    for (int e = 0; e < E; ++e) {
                        ok[e] = s_k[(uint32_t)e * kBucketThreads + pt];
                        ov[e] = s_v[(uint32_t)e * kBucketThreads + pt];
                    }
#pragma unroll
                    for (int e = 0; e < E; ++e) bitonic_keep<FULL>(key[e], val[e], ok[e], ov[e], keep_min);
                    __syncthreads();
                } else {
                    // partner in the same wave: lane exchanges on the VALU
                    switch (m) {
                        case 1: bitonic_wave<FULL, E, 1>(key, val, keep_min); break;
                        case 2: bitonic_wave<FULL, E, 2>(key, val, keep_min); break;
                        case 4: bitonic_wave<FULL, E, 4>(key, val, keep_min); break;
                        case 8: bitonic_wave<FULL, E, 8>(key, val, keep_min); break;
                        case 16: bitonic_wave<FULL, E, 16>(key, val, keep_min); break;
                        default: bitonic_wave<FULL, E, 32>(key, val, keep_min); break;
                    }
                }
            } else {
                // within the thread
                if (E > 4 && h == 4) bitonic_regs<FULL, E, (E > 4 ? 4 : 1)>(key, val, k);
                else if (E > 2 && h == 2) bitonic_regs<FULL, E, (E > 2 ? 2 : 1)>(key, val, k);
                else bitonic_regs<FULL, E, 1>(key, val, k);
            }
        }
    }
}

template <int E>
__device__ __forceinline__ void bucket_sort_regs(uint64_t* __restrict__ keys,
                                                 uint32_t* __restrict__ vals, uint64_t base,
                                                 uint32_t cnt, uint64_t* s_k, uint32_t* s_v) {
    uint64_t key[E];
    uint32_t val[E];
    const uint32_t t = threadIdx.x;
    bool top = false;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const uint32_t i = t * E + (uint32_t)e;
        // padding sorts last: (~0, ~0) exceeds every (key, index)
        key[e] = i < cnt ? keys[base + i] : ~0ull;
        val[e] = i < cnt ? vals[base + i] : ~0u;
        top |= i < cnt && key[e] == ~0ull;
    }
    // by key alone unless a key ties with the padding
    if (__syncthreads_or(top)) bitonic_block<true, E>(key, val, s_k, s_v);
    else bitonic_block<false, E>(key, val, s_k, s_v);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const uint32_t i = t * E + (uint32_t)e;
        if (i < cnt) {
            keys[base + i] = key[e];
            vals[base + i] = val[e];
        }
    }
}

// Slot sort of a bucket of at most kSlots keys (buckets hold kBucketKeys on
// average). Keys are spread over kSlots slots by a monotone map: a coarse bin
// (between two of the bucket's quantiles of the previous step's keys, which
// follow the keys' distribution at every scale; without them, the top
// kCoarseBits bits of the key's offset from the bucket's smallest), bins
// getting slots in proportion to their key counts (a piecewise-linear fit of
// the bucket's key distribution), and the key's place within its bin picking
// the slot within the bin's. Each slot so holds one
// contiguous key range, the slots in key order, a key or two each. One LDS
// atomic per key counts its slot and returns its rank there; the scanned
// counts place every key; in a slot of up to kSlotRun keys each key then
// counts the slot's keys below it (independent LDS reads, where a per-thread
// insertion sort made a chain of dependent ones) and moves to that rank, and
// a slot of more than kSlotRun keys is sorted by a wave (one key per lane).
// Equal keys end in no fixed order, as in the bitonic sort. Returns false,
// with the bucket untouched, when a slot holds more than 64 keys (clustered
// keys: the bitonic sort takes the bucket).
constexpr int kSlotBits = 11;
constexpr uint32_t kSlots = 1u << kSlotBits;
constexpr uint32_t kSlotRun = 8;
constexpr int kCoarseBits = 8;
constexpr uint32_t kCoarse = 1u << kCoarseBits;
static_assert(kQuantiles == (int)kCoarse, "a coarse bin between consecutive quantiles");
constexpr uint32_t kWideSlots = 128;  // slots of more than kSlotRun keys sorted by waves

// Sorts one (key, value) per lane across a wave by (key, value), ascending in
// lane order (a register bitonic network).
__device__ __forceinline__ void wave_sort64(uint64_t& key, uint32_t& val) {
    const uint32_t lane = threadIdx.x & 63u;
    uint64_t k1[1] = {key};
    uint32_t v1[1] = {val};
#define PG_WAVE_STAGE(K, H) \
    bitonic_wave<true, 1, H>(k1, v1, ((lane & (H)) == 0) == ((lane & (K)) == 0));
    PG_WAVE_STAGE(2, 1)
    PG_WAVE_STAGE(4, 2) PG_WAVE_STAGE(4, 1)
    PG_WAVE_STAGE(8, 4) PG_WAVE_STAGE(8, 2) PG_WAVE_STAGE(8, 1)
    PG_WAVE_STAGE(16, 8) PG_WAVE_STAGE(16, 4) PG_WAVE_STAGE(16, 2) PG_WAVE_STAGE(16, 1)
    PG_WAVE_STAGE(32, 16) PG_WAVE_STAGE(32, 8) PG_WAVE_STAGE(32, 4) PG_WAVE_STAGE(32, 2)
    PG_WAVE_STAGE(32, 1)
    PG_WAVE_STAGE(64, 32) PG_WAVE_STAGE(64, 16) PG_WAVE_STAGE(64, 8) PG_WAVE_STAGE(64, 4)
    PG_WAVE_STAGE(64, 2) PG_WAVE_STAGE(64, 1)
#undef PG_WAVE_STAGE
    key = k1[0];
    val = v1[0];
}

// The offset `off` of a key within a bin of width `width` (off <= width) as
// a 16-bit fraction: monotone in off.
__device__ __forceinline__ uint32_t frac16(uint64_t off, uint64_t width) {
    const int bl = width ? 64 - __clzll((long long)width) : 0;
    return bl > 16 ? (uint32_t)(off >> (bl - 16)) : (uint32_t)(off << (16 - bl));
}
static_assert(kSlots == 4 * kBucketThreads, "four slots (and up to four keys) per thread");

template <int M>
__device__ __forceinline__ void minmax_lanes(uint64_t& lo, uint64_t& hi) {
    const uint64_t a = xor_lanes64<M>(lo), b = xor_lanes64<M>(hi);
    lo = a < lo ? a : lo;
    hi = b > hi ? b : hi;
}

__device__ __forceinline__ bool bucket_sort_slots(uint64_t* __restrict__ keys,
                                                  uint32_t* __restrict__ vals, uint64_t base,
                                                  uint32_t cnt, uint8_t* s_raw,
                                                  const uint64_t* __restrict__ quant,
                                                  uint32_t& heads) {
    constexpr int kWaves = kBucketThreads / 64;
    uint64_t* s_k = reinterpret_cast<uint64_t*>(s_raw);
    uint32_t* s_v = reinterpret_cast<uint32_t*>(s_raw + kSlots * 8);
    // two 16-bit fields per word: slot counts, then slot starts
    uint32_t* s_pos = reinterpret_cast<uint32_t*>(s_raw + kSlots * 12);
    uint64_t* s_red = reinterpret_cast<uint64_t*>(s_raw + kSlots * 14);
    uint32_t* s_scan = reinterpret_cast<uint32_t*>(s_raw + kSlots * 14 + 16 * kWaves);
    // coarse bins: counts, then first slot | slot count << 16
    uint32_t* s_bin = s_scan + kWaves;
    uint64_t* s_q = reinterpret_cast<uint64_t*>(s_bin + kCoarse);  // the bucket's quantiles
    uint32_t* s_wide = reinterpret_cast<uint32_t*>(s_q + kCoarse);   // count, then wide slots
    const uint32_t t = threadIdx.x;
    const int lane = (int)(t & 63u), wave = (int)(t >> 6);
    // every load goes out before anything waits on one (indices clamped into
    // the bucket, cnt >= 2: a load inside a divergent branch is waited for at
    // the branch's end)
    const uint64_t qv = quant ? quant[min(t, kCoarse - 1u)] : 0ull;
    uint64_t key[4];
    uint32_t val[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const uint64_t i = base + min((uint32_t)e * kBucketThreads + t, cnt - 1u);
        key[e] = keys[i];
        val[e] = vals[i];
    }
    if (t == 0) s_wide[0] = 0;
    s_pos[2 * t] = 0;
    s_pos[2 * t + 1] = 0;
    if (t < kCoarse) {
        s_bin[t] = 0;
        if (quant) s_q[t] = qv;
    }
    uint64_t lo = ~0ull, hi = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        if ((uint32_t)e * kBucketThreads + t < cnt) {
            lo = key[e] < lo ? key[e] : lo;
            hi = key[e] > hi ? key[e] : hi;
        }
    }
    minmax_lanes<1>(lo, hi);
    minmax_lanes<2>(lo, hi);
    minmax_lanes<4>(lo, hi);
    minmax_lanes<8>(lo, hi);
    minmax_lanes<16>(lo, hi);
    minmax_lanes<32>(lo, hi);
    if (lane == 0) {
        s_red[wave] = lo;
        s_red[kWaves + wave] = hi;
    }
    __syncthreads();
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        const uint64_t a = s_red[w], b = s_red[kWaves + w];
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    }
    const uint64_t range = hi - lo;
    const int bits = range ? 64 - __clzll((long long)range) : 0;
    const int sh = bits > kCoarseBits ? bits - kCoarseBits : 0;
    uint32_t slot[4], rank[4], frac[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const uint32_t i = (uint32_t)e * kBucketThreads + t;
        if (quant) {
            // the last quantile <= key (the first if none)
            uint32_t j = 0;
#pragma unroll
            for (uint32_t st = kCoarse / 2; st > 0; st >>= 1)
                if (s_q[j + st] <= key[e]) j += st;
            const uint64_t a = j ? s_q[j] : (s_q[0] < lo ? s_q[0] : lo);
            const uint64_t z = j + 1 < kCoarse ? s_q[j + 1] : (s_q[kCoarse - 1] > hi ? s_q[kCoarse - 1] : hi);
            slot[e] = j;  // the coarse bin, for now
            frac[e] = frac16(key[e] - a, z - a);
        } else {
            slot[e] = (uint32_t)((key[e] - lo) >> sh);
            frac[e] = frac16((key[e] - lo) - ((uint64_t)slot[e] << sh), (1ull << sh) - 1ull);
        }
        if (i < cnt) atomicAdd(&s_bin[slot[e]], 1u);
    }
    __syncthreads();
    {
        // bin b's slots: [kSlots * (keys before b) / cnt, ... of b + 1)
        const uint32_t c = t < kCoarse ? s_bin[t] : 0u;
        uint32_t all;
        const uint32_t before = block_exclusive_scan<kBucketThreads>(c, s_scan, &all);
        if (t < kCoarse) {
            const uint32_t f = (uint32_t)(((uint64_t)before * kSlots) / cnt);
            const uint32_t l = (uint32_t)(((uint64_t)(before + c) * kSlots) / cnt);
            s_bin[t] = f | ((l - f) << 16);
        }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const uint32_t i = (uint32_t)e * kBucketThreads + t;
        rank[e] = 0;
        if (i < cnt) {
            const uint32_t b = s_bin[slot[e]];
            slot[e] = min((b & 0xFFFFu) + ((frac[e] * (b >> 16)) >> 16), kSlots - 1u);
            const uint32_t h = (slot[e] & 1u) * 16u;
            rank[e] = (atomicAdd(&s_pos[slot[e] >> 1], 1u << h) >> h) & 0xFFFFu;
        }
    }
    __syncthreads();
    // this thread's slots 4t .. 4t + 3
    const uint32_t w0 = s_pos[2 * t], w1 = s_pos[2 * t + 1];
    const uint32_t c0 = w0 & 0xFFFFu, c1 = w0 >> 16, c2 = w1 & 0xFFFFu, c3 = w1 >> 16;
    if (__syncthreads_or(max(max(c0, c1), max(c2, c3)) > 64u)) return false;
    const uint32_t cs[4] = {c0, c1, c2, c3};
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (cs[k] > kSlotRun) {
            const uint32_t q = atomicAdd(&s_wide[0], 1u);
            if (q < kWideSlots) s_wide[1 + q] = 4u * t + (uint32_t)k;
        }
    uint32_t all;
    const uint32_t s0 = block_exclusive_scan<kBucketThreads>(c0 + c1 + c2 + c3, s_scan, &all);
    const uint32_t s1 = s0 + c0, s2 = s1 + c1, s3 = s2 + c2;
    // (the scan's barriers ordered the wide-slot count)
    if (s_wide[0] > kWideSlots) return false;
    s_pos[2 * t] = s0 | (s1 << 16);
    s_pos[2 * t + 1] = s2 | (s3 << 16);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const uint32_t i = (uint32_t)e * kBucketThreads + t;
        if (i < cnt) {
            const uint32_t p =
                ((s_pos[slot[e] >> 1] >> ((slot[e] & 1u) * 16u)) & 0xFFFFu) + rank[e];
            s_k[p] = key[e];
            s_v[p] = val[e];
        }
    }
    __syncthreads();
    // a slot of 2 .. kSlotRun keys: each key's place is its rank among the
    // slot's keys, equal keys keeping their arrival order (a stable sort);
    // the ranks are read first (independent loads), then every key moves
    uint32_t dst[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const uint32_t i = (uint32_t)e * kBucketThreads + t;
        dst[e] = ~0u;
        if (i < cnt) {
            const uint32_t sl = slot[e];
            const uint32_t a = (s_pos[sl >> 1] >> ((sl & 1u) * 16u)) & 0xFFFFu;
            const uint32_t z = sl + 1u < kSlots ? (s_pos[(sl + 1u) >> 1] >> (((sl + 1u) & 1u) * 16u)) & 0xFFFFu : cnt;
            const uint32_t c = z - a, p = a + rank[e];
            if (c > 1u && c <= kSlotRun) {
                uint32_t r = 0;
#pragma unroll
                for (uint32_t j = 0; j < kSlotRun; ++j) {
                    if (j < c) {
                        const uint64_t y = s_k[a + j];
                        r += (y < key[e] || (y == key[e] && a + j < p)) ? 1u : 0u;
                    }
                }
                dst[e] = a + r;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        if (dst[e] != ~0u) {
            s_k[dst[e]] = key[e];
            s_v[dst[e]] = val[e];
        }
    }
    const uint32_t nwide = s_wide[0];
    for (uint32_t q = (uint32_t)wave; q < nwide; q += kWaves) {
        const uint32_t sl = s_wide[1 + q];
        const uint32_t a = (s_pos[sl >> 1] >> ((sl & 1u) * 16u)) & 0xFFFFu;
        const uint32_t z = sl + 1u < kSlots ? (s_pos[(sl + 1u) >> 1] >> (((sl + 1u) & 1u) * 16u)) & 0xFFFFu : cnt;
        const uint32_t c = z - a;  // kSlotRun < c <= 64
        uint64_t x = (uint32_t)lane < c ? s_k[a + lane] : ~0ull;
        uint32_t xv = (uint32_t)lane < c ? s_v[a + lane] : ~0u;
        wave_sort64(x, xv);
        if ((uint32_t)lane < c) {
            s_k[a + lane] = x;
            s_v[a + lane] = xv;
        }
    }
    __syncthreads();
    // the sorted bucket goes out; bit e of `heads` marks the thread's key e
    // as the head of a run of equal keys (equal keys share a bucket)
    uint32_t h = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const uint32_t i = (uint32_t)e * kBucketThreads + t;
        if (i < cnt) {
            const uint64_t k = s_k[i];
            keys[base + i] = k;
            vals[base + i] = s_v[i];
            if (i + 1u < cnt && s_k[i + 1u] == k && (i == 0 || s_k[i - 1u] != k)) h |= 1u << e;
        }
    }
    heads = h;
    return true;
}

// (4) One block per bucket: sorts keys[base, base + count) (with vals) by
// key. Up to `cap` keys (at most kBucketCap) by a bitonic sort held in
// registers; more by stable LSD passes in HBM, ping-ponging with tkeys /
// tvals over the same range. Equal keys fall in one bucket: the block then
// orders their runs (order_tie_run).
template <int DIM>
__global__ __launch_bounds__(kBucketThreads) __attribute__((amdgpu_waves_per_eu(6))) void k_bucket_sort(
    uint64_t* __restrict__ keys, uint32_t* __restrict__ vals, uint64_t* __restrict__ tkeys,
    uint32_t* __restrict__ tvals, uint32_t* __restrict__ bcount, uint32_t* __restrict__ bfill,
    const uint32_t* __restrict__ bbase, uint32_t cap, const SegDesc* __restrict__ desc,
    uint64_t total, DeepKey* __restrict__ deep, uint32_t nseg, int segbits, int lmax,
    const uint64_t* __restrict__ quant, const StepCounters* __restrict__ sc) {
    constexpr int kWaves = kBucketThreads / 64;
    static_assert(kBucketCap == 8 * kBucketThreads, "up to eight keys per thread");
    // LDS: the bitonic exchange arrays, or (HBM path) the ranking tables
    constexpr int kRankBytes = (kWaves * 256 + 3 * 256 + kWaves) * 4 + kWaves * 8;
    constexpr int kSortBytes = kBucketCap * 12;
    static_assert(kSlots * 14 + 16 * kWaves + 4 * kWaves + 12 * kCoarse + 4 * (1 + kWideSlots) <=
                      kSortBytes,
                  "slot sort in the sort space");
    __shared__ __align__(16) uint8_t s_raw[kSortBytes > kRankBytes ? kSortBytes : kRankBytes];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t cnt = bcount[blockIdx.x];
    const uint64_t base = bbase[blockIdx.x];
    // the next step's bucket count (fused into k_keys) starts from zero
    __syncthreads();
    if (tid == 0) {
        bcount[blockIdx.x] = 0;
        bfill[blockIdx.x] = 0;
    }
    if (cnt <= 1) return;
    // ranking tables of the HBM passes
    uint8_t* s_tab = s_raw;
    uint32_t(*s_wcnt)[256] = reinterpret_cast<uint32_t(*)[256]>(s_tab);
    uint32_t* s_dstart = reinterpret_cast<uint32_t*>(s_tab) + kWaves * 256;
    uint32_t* s_dcnt = s_dstart + 256;
    uint32_t* s_run = s_dcnt + 256;
    uint32_t* s_scan = s_run + 256;
    uint64_t* s_or = reinterpret_cast<uint64_t*>(s_tab + (kWaves * 256 + 3 * 256 + kWaves) * 4);
    const uint64_t ref = keys[base];
    if (cnt <= min(cap, (uint32_t)kBucketCap)) {
        uint64_t* s_bk = reinterpret_cast<uint64_t*>(s_raw);
        uint32_t* s_bv = reinterpret_cast<uint32_t*>(s_raw + kBucketCap * 8);
        uint32_t heads = 0;
        const bool slots =
            cnt <= kSlots &&
            bucket_sort_slots(keys, vals, base, cnt, s_raw,
                              sc->quantiles ? quant + (uint64_t)blockIdx.x * kQuantiles : nullptr,
                              heads);
        if (!slots) {
            __syncthreads();  // (the slot sort's tables share the space)
            if (cnt <= 2u * kBucketThreads) bucket_sort_regs<2>(keys, vals, base, cnt, s_bk, s_bv);
            else if (cnt <= 4u * kBucketThreads) bucket_sort_regs<4>(keys, vals, base, cnt, s_bk, s_bv);
            else bucket_sort_regs<8>(keys, vals, base, cnt, s_bk, s_bv);
        }
        __syncthreads();
        if (slots) {
            // only the heads the slot sort marked start runs of equal keys
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if ((heads >> e) & 1u)
                    order_tie_run<DIM>(desc, total, keys, vals, deep, nseg, segbits, lmax,
                                       base + (uint32_t)e * kBucketThreads + (uint32_t)tid);
        } else {
            for (uint32_t i = tid; i < cnt; i += kBucketThreads)
                order_tie_run<DIM>(desc, total, keys, vals, deep, nseg, segbits, lmax, base + i);
        }
        return;
    }
    // HBM: stable LSD passes over the range, chunk by chunk in order
    uint64_t diff = 0;
    for (uint32_t i = tid; i < cnt; i += kBucketThreads) diff |= keys[base + i] ^ ref;
    const int hb = block_top_diff_bit(diff, s_or);
    uint64_t *ka = keys + base, *kb = tkeys + base;
    uint32_t *va = vals + base, *vb = tvals + base;
    for (int shift = 0; shift <= hb; shift += 8) {
        // digit totals of the range, then running starts
        if (tid < 256) s_run[tid] = 0;
        __syncthreads();
        for (uint32_t i = tid; i < cnt; i += kBucketThreads)
            atomicAdd(&s_run[(uint32_t)((ka[i] >> shift) & 255u)], 1u);
        __syncthreads();
        uint32_t total;
        const uint32_t start = block_exclusive_scan<kBucketThreads>(tid < 256 ? s_run[tid] : 0u,
                                                                    s_scan, &total);
        if (tid < 256) s_run[tid] = start;
        __syncthreads();
        for (uint32_t c0 = 0; c0 < cnt; c0 += kBucketChunk) {
            uint64_t key[4];
            uint32_t val[4], dig[4], pos[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t i = c0 + (uint32_t)wave * 256u + (uint32_t)k * 64u + (uint32_t)lane;
                const bool valid = i < cnt;
                key[k] = valid ? ka[i] : 0ull;
                val[k] = valid ? va[i] : 0u;
                dig[k] = valid ? (uint32_t)((key[k] >> shift) & 255u) : 256u;
            }
            chunk_rank(dig, pos, s_wcnt, s_dstart, s_dcnt, s_scan);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (dig[k] < 256u) {
                    const uint32_t o = s_run[dig[k]] + (pos[k] - s_dstart[dig[k]]);
                    kb[o] = key[k];
                    vb[o] = val[k];
                }
            }
            __syncthreads();
            if (tid < 256) s_run[tid] += s_dcnt[tid];
            __syncthreads();
        }
        uint64_t* tk = ka; ka = kb; kb = tk;
        uint32_t* tv = va; va = vb; vb = tv;
    }
    if (ka != keys + base) {
        for (uint32_t i = tid; i < cnt; i += kBucketThreads) {
            keys[base + i] = ka[i];
            vals[base + i] = va[i];
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < cnt; i += kBucketThreads)
        order_tie_run<DIM>(desc, total, keys, vals, deep, nseg, segbits, lmax, base + i);
}

// Per sorted body: its common-prefix lengths with its neighbours give the
// tree levels it owns (lo..hi, hi = its leaf; lo = 1 + common prefix with the
// previous body, 0 at a segment start). The internal nodes (levels lo..hi-1)
// are stored, in preorder: the exclusive scan of the per-body internal node
// counts lays them out; leaves are the bodies themselves, in sorted order.
// This kernel writes each body's offset within its tile (node_local) and the
// tile's total; k_tree_scan adds the tiles up. Also permutes the state into
// the new Morton order -- positions and masses as one record per body
// (leafpos), the rest in the state's arrays (velocities only while they are
// not derivable from the two positions) -- and flags equal keys (depth).
//
// Node ranges: the node at level L containing body i starts at the last body
// j <= i with lo[j] <= L and ends before the first body j > i with lo[j] <= L.
// Per tile and level this kernel records the first such body (~0 = none);
// k_tree_scan turns that into a suffix minimum over the tiles, so the node
// builder finds the end of a node that leaves its tile in one load.
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_scan_gather(
    const SegDesc* __restrict__ desc, const uint32_t* __restrict__ tile_start,
    uint64_t* __restrict__ ext_partial, uint32_t nchunks, const uint64_t* __restrict__ skeys,
    const uint32_t* __restrict__ svals, uint32_t nseg, int segbits, int lmax,
    uint32_t* __restrict__ node_local, uint32_t* __restrict__ tile_tot,
    uint8_t* __restrict__ lo_level, uint32_t* __restrict__ lvl_first,
    const DeepKey* __restrict__ deep, uint32_t* err, uint8_t* __restrict__ leaf_lv,
    double4* __restrict__ leafpos) {
    __shared__ uint32_t s_scan[kBlock / 64];
    __shared__ uint32_t s_cnt[kScanTile];
    __shared__ uint32_t s_firstlo[kTopSlots];
    const int tid = threadIdx.x;
    // the blocks of one XCD (workgroups go to the 8 XCDs in turn) take one
    // contiguous eighth of the tiles: the source lines that neighbouring
    // tiles share are then fetched once, into that XCD's L2
    const uint32_t nbk = gridDim.x, xcd = blockIdx.x & 7u, per = nbk >> 3, rem = nbk & 7u;
    const uint32_t tile = xcd * per + min(xcd, rem) + (blockIdx.x >> 3);
    const uint64_t tbase = tile_start[tile], tend = tile_start[tile + 1];
    if (tid < kTopSlots) s_firstlo[tid] = ~0u;
    // the extent partials were read by k_keys: cleared for the walk's maxima
    if (tile == 0)
        for (uint32_t c = tid; c < nchunks; c += kBlock) ext_partial[c] = 0;
    __syncthreads();

    // A node tile lies in one segment (tiles start at every segment's first
    // body): the segment's fields are read once, for the segment named by
    // the tile's first key, and each body's own key is checked against it.
    // Every thread requests its bodies' keys, neighbour keys and source
    // indices together, then their state records together (indices clamped
    // into range rather than guarded), so a body costs two memory round
    // trips rather than one per dependent field.
    const uint64_t total = tile_start[gridDim.x];  // (one block per tile)
    uint64_t kc[kScanItems], kl[kScanItems], kr[kScanItems];
    uint32_t sv[kScanItems];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint64_t i = min(tbase + (uint32_t)k * kBlock + tid, tend - 1);
        kc[k] = skeys[i];
        kl[k] = skeys[i > 0 ? i - 1 : 0];
        kr[k] = skeys[i + 1 < total ? i + 1 : i];
        sv[k] = svals[i];
    }
    const uint32_t s_t = uniform_u32(seg_of_key(skeys[tbase], segbits));
    const SegDesc& d = desc[min(s_t, nseg - 1)];
    const uint64_t off = d.off, n = d.n, end = off + n;
    const bool seg_ok = s_t < nseg && n > 0;  // (else every body is flagged)
    const bool near = seg_ok && d.near > 0.0;
    const bool vperm = seg_ok && gload(&d.st->vperm) != 0;
    bool bad[kScanItems];
    double4 lp[kScanItems];
    if (seg_ok) {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            const uint64_t i = min(tbase + (uint32_t)k * kBlock + tid, tend - 1);
            const uint64_t src = (uint64_t)sv[k] - off;
            bad[k] = seg_of_key(kc[k], segbits) != s_t || src >= n || i < off || i >= end;
            lp[k] = gload(d.spos + (bad[k] ? 0 : src));
        }
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) bad[k] = true;
    }
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t li = (uint32_t)k * kBlock + tid;
        const uint64_t i = tbase + li;
        uint32_t c = 0;
        if (i < tend) {
            int lo = 0, leaf = 0;
            if (bad[k]) {
                atomicOr(err, kErrCapacity);
            } else {
                // two bodies share at most kMaxLcp levels (the reference
                // panics past depth 64, octree.rs:60-62)
                const int depth = kMaxLcp + 1;
                int lcpL = i > off ? key_lcp2<DIM>(kl[k], kc[k], deep, i - 1, i, lmax) : -1;
                int lcpR = i + 1 < end ? key_lcp2<DIM>(kc[k], kr[k], deep, i, i + 1, lmax) : -1;
                if (lcpR >= depth) {
                    atomicOr(&d.st->err, kErrDepth);
                    lcpR = depth - 1;
                }
                if (lcpL >= depth) lcpL = depth - 1;
                // the internal nodes the body owns: levels lcpL + 1 .. lcpR
                // (its leaf, at level max(lcpL, lcpR) + 1, is implicit)
                c = (uint32_t)max(0, lcpR - lcpL);
                lo = lcpL + 1;
                leaf = max(lcpL, lcpR) + 1;  // the level of the body's own leaf
                // position and mass as one record in sorted order: the node
                // builder's input and the walk's leaves (the walk writes the
                // state's own arrays, and moves each body's record of previous
                // position and tag itself: it is that record's only reader)
                leafpos[i] = lp[k];
                if (vperm) {
                    const uint64_t j = i - off, src = (uint64_t)sv[k] - off;
                    gstore(d.dvx + j, gload(d.svx + src));
                    gstore(d.dvy + j, gload(d.svy + src));
                    gstore(d.dvz + j, gload(d.svz + src));
                }
            }
            lo_level[i] = (uint8_t)lo;
            // by sorted position, read by a walk with a near field
            if (near) leaf_lv[i] = (uint8_t)leaf;
            // (a body below every internal level, lo = kMaxLcp + 1, starts none)
            if (lo < kTopSlots) atomicMin(&s_firstlo[lo], (uint32_t)i);
        }
        s_cnt[li] = c;
    }
    __syncthreads();
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) sum += s_cnt[tid * kScanItems + k];
    uint32_t tile_total;
    uint32_t run = block_exclusive_scan(sum, s_scan, &tile_total);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t li = tid * kScanItems + k;
        const uint32_t c = s_cnt[li];
        s_cnt[li] = run;
        run += c;
    }
    if (tid < kTopSlots) {
        uint32_t mn = ~0u;
        for (int v = 0; v <= tid; ++v) mn = min(mn, s_firstlo[v]);
        lvl_first[(uint64_t)tile * kTopSlots + tid] = mn;
    }
    if (tid == 0) tile_tot[tile] = tile_total;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t li = (uint32_t)k * kBlock + tid;
        if (tbase + li < tend) node_local[tbase + li] = s_cnt[li];
    }
}

// Inclusive scan over a block of NT threads with an associative operation
// of identity id.
template <int NT, class Op>
__device__ __forceinline__ uint32_t block_inclusive_scan(uint32_t v, uint32_t* s_wave, uint32_t id,
                                                         Op op) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    v = wave_incl_scan_u32(v, id, op);
    if (lane == 63) s_wave[wave] = v;
    __syncthreads();
    for (int w = 0; w < wave; ++w) v = op(s_wave[w], v);
    __syncthreads();
    return v;
}

// Block 0: exclusive scan of the tile node totals into tile_base[0..tiles]
// (tile_base[tiles] = all nodes) and first_node[total] = all nodes. Block
// 1 + L: suffix minimum over the tiles of lvl_first[., L], in place.
__global__ __launch_bounds__(kScanThreads) void k_tree_scan(
    const uint32_t* __restrict__ tile_tot, uint32_t* __restrict__ tile_base, uint32_t tiles,
    uint32_t* __restrict__ lvl_first, uint32_t* __restrict__ first_node, uint64_t total) {
    __shared__ uint32_t s_wave[kScanThreads / 64];
    // a thread's first kKeep values stay in registers between the scan's
    // two passes (up to kKeep * kScanThreads tiles: no second read)
    constexpr uint32_t kKeep = 8;
    const uint32_t per = (tiles + kScanThreads - 1) / kScanThreads;
    const uint32_t b = threadIdx.x * per;
    uint32_t v[kKeep];
    if (blockIdx.x == 0) {
        uint32_t s = 0;
#pragma unroll
        for (uint32_t k = 0; k < kKeep; ++k) {
            v[k] = k < per && b + k < tiles ? tile_tot[b + k] : 0u;
            s += v[k];
        }
        for (uint32_t k = kKeep; k < per; ++k)
            if (b + k < tiles) s += tile_tot[b + k];
        uint32_t all;
        uint32_t run = block_exclusive_scan<kScanThreads>(s, s_wave, &all);
#pragma unroll
        for (uint32_t k = 0; k < kKeep; ++k) {
            if (k < per && b + k < tiles) {
                tile_base[b + k] = run;
                run += v[k];
            }
        }
        for (uint32_t k = kKeep; k < per; ++k) {
            if (b + k < tiles) {
                tile_base[b + k] = run;
                run += tile_tot[b + k];
            }
        }
        if (threadIdx.x == 0) {
            tile_base[tiles] = all;
            first_node[total] = all;
        }
        return;
    }
    const uint32_t L = blockIdx.x - 1;
    // suffix minimum: thread r walks the tiles from the end
    const uint32_t rb = threadIdx.x * per;
    uint32_t mn = ~0u;
#pragma unroll
    for (uint32_t k = 0; k < kKeep; ++k) {
        v[k] = k < per && rb + k < tiles ? lvl_first[(uint64_t)(tiles - 1 - (rb + k)) * kTopSlots + L]
                                         : ~0u;
        mn = min(mn, v[k]);
    }
    for (uint32_t k = kKeep; k < per; ++k)
        if (rb + k < tiles) mn = min(mn, lvl_first[(uint64_t)(tiles - 1 - (rb + k)) * kTopSlots + L]);
    uint32_t rrun = block_inclusive_scan<kScanThreads>(mn, s_wave, ~0u, OpMin{});
    // exclusive: the thread before's value (lane 0: the earlier waves')
    rrun = dpp_move<kDppWaveShr1, 0xF>(rrun, ~0u);
    {
        const int wave = threadIdx.x >> 6;
        for (int w = 0; w < wave; ++w) rrun = min(rrun, s_wave[w]);
    }
#pragma unroll
    for (uint32_t k = 0; k < kKeep; ++k) {
        if (k < per && rb + k < tiles) {
            rrun = min(rrun, v[k]);
            lvl_first[(uint64_t)(tiles - 1 - (rb + k)) * kTopSlots + L] = rrun;
        }
    }
    for (uint32_t k = kKeep; k < per; ++k) {
        if (rb + k < tiles) {
            uint32_t* p = lvl_first + (uint64_t)(tiles - 1 - (rb + k)) * kTopSlots + L;
            rrun = min(rrun, *p);
            *p = rrun;
        }
    }
}

// Last index j >= i (j < end) whose key shares key i's prefix at `level`,
// over keys staged in LDS (galloping, then bisection).
template <int DIM>
__device__ __forceinline__ uint32_t range_end_lds(const uint64_t* s, uint32_t i, uint32_t end,
                                                  int lmax, int level) {
    const uint64_t P = key_prefix<DIM>(s[i], lmax, level);
    uint32_t last = i, step = 1;
    for (;;) {
        const uint32_t probe = last + step;
        if (probe >= end || key_prefix<DIM>(s[probe], lmax, level) != P) break;
        last = probe;
        step <<= 1;
    }
    uint32_t hi = min(last + step, end);
    while (hi - last > 1) {
        const uint32_t mid = (last + hi) >> 1;
        if (key_prefix<DIM>(s[mid], lmax, level) == P) last = mid; else hi = mid;
    }
    return last;
}

// Mass sums (m, m x, m y, m z) of body ranges, FP64.
struct Sum4 {
    double m, x, y, z;
    __device__ __forceinline__ void add(const Sum4& o) {
        m += o.m;
        x += o.x;
        y += o.y;
        z += o.z;
    }
};

// dpp_move of a double, old +0.0.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_f64(double v) {
    const uint64_t b = f64_bits(v);
    const uint32_t lo = dpp_move<CTRL, ROWS>((uint32_t)b, 0u);
    const uint32_t hi = dpp_move<CTRL, ROWS>((uint32_t)(b >> 32), 0u);
    return bits_f64(((uint64_t)hi << 32) | lo);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ Sum4 dpp_sum4(const Sum4& v) {
    return Sum4{dpp_f64<CTRL, ROWS>(v.m), dpp_f64<CTRL, ROWS>(v.x), dpp_f64<CTRL, ROWS>(v.y),
                dpp_f64<CTRL, ROWS>(v.z)};
}
// Inclusive sum and minimum over a wave's lanes in lane order: steps of 1, 2,
// 4 and 8 lanes within rows of 16 (row_shr), then rows 1 and 3 take the last
// lane of the row before (row_bcast:15), rows 2 and 3 lane 31 (row_bcast:31).
// The association is fixed.
template <int CTRL, int ROWS>
__device__ __forceinline__ void scan_step(Sum4& v, uint32_t& m) {
    v.add(dpp_sum4<CTRL, ROWS>(v));
    m = min(m, dpp_move<CTRL, ROWS>(m, ~0u));
}
__device__ __forceinline__ void wave_incl_scan(Sum4& v, uint32_t& m) {
    scan_step<0x111, 0xF>(v, m);
    scan_step<0x112, 0xF>(v, m);
    scan_step<0x114, 0xF>(v, m);
    scan_step<0x118, 0xF>(v, m);
    scan_step<0x142, 0xA>(v, m);
    scan_step<0x143, 0xC>(v, m);
}

// Exclusive scans of two (Sum4, u32) pairs per thread over a block of NT
// threads in thread order: sums (fixed association) and minimums (~0u for
// the first thread), and the first pair's sum over the block.
template <int NT>
__device__ __forceinline__ void block_excl_scan2(Sum4& a, uint32_t& ma, Sum4& b, uint32_t& mb,
                                                 Sum4 (*s_w)[NT / 64], uint32_t (*s_m)[NT / 64],
                                                 Sum4* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    wave_incl_scan(a, ma);
    wave_incl_scan(b, mb);
    if (lane == 63) {
        s_w[0][wave] = a;
        s_w[1][wave] = b;
        s_m[0][wave] = ma;
        s_m[1][wave] = mb;
    }
    // exclusive within the wave: the lane before's inclusive value (wave_shr:1)
    Sum4 xa = dpp_sum4<kDppWaveShr1, 0xF>(a), xb = dpp_sum4<kDppWaveShr1, 0xF>(b);
    uint32_t ya = dpp_move<kDppWaveShr1, 0xF>(ma, ~0u), yb = dpp_move<kDppWaveShr1, 0xF>(mb, ~0u);
    __syncthreads();
    Sum4 ba{0.0, 0.0, 0.0, 0.0}, bb{0.0, 0.0, 0.0, 0.0}, tot{0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
    for (int w = 0; w < NT / 64; ++w) {
        const Sum4 ca = s_w[0][w];
        if (w < wave) {
            ba.add(ca);
            bb.add(s_w[1][w]);
            ya = min(ya, s_m[0][w]);
            yb = min(yb, s_m[1][w]);
        }
        tot.add(ca);
    }
    __syncthreads();
    ba.add(xa);
    bb.add(xb);
    a = ba;
    b = bb;
    ma = ya;
    mb = yb;
    *total = tot;
}

// Mass and centre of mass of a node from the mass sums of its bodies.
__device__ __forceinline__ Sum4 com_of(const Sum4& s) {
    const double inv = 1.0 / s.m;
    return Sum4{s.m, s.x * inv, s.y * inv, s.z * inv};
}

// Mass sums of the staged bodies [a, e), in body order.
__device__ __forceinline__ Sum4 range_sum(const double* s_m, const double* s_x, const double* s_y,
                                          const double* s_z, uint32_t a, uint32_t e) {
    Sum4 r{0.0, 0.0, 0.0, 0.0};
    for (uint32_t b = a; b < e; ++b) {
        const double m = s_m[b];
        r.m += m;
        r.x += m * s_x[b];
        r.y += m * s_y[b];
        r.z += m * s_z[b];
    }
    return r;
}

// r += r of lane ^ M.
template <int M>
__device__ __forceinline__ void fold_lanes(Sum4& r) {
    r.m += bits_f64(xor_lanes64<M>(f64_bits(r.m)));
    r.x += bits_f64(xor_lanes64<M>(f64_bits(r.x)));
    r.y += bits_f64(xor_lanes64<M>(f64_bits(r.y)));
    r.z += bits_f64(xor_lanes64<M>(f64_bits(r.z)));
}

// The same sums by a group of 16 lanes (a DPP row): lane l of the group
// folds bodies a + l, a + l + 16, ... in order, then a butterfly over the
// group; the group's lane 0 holds a fixed association.
__device__ __forceinline__ Sum4 range_sum_row(const double* s_m, const double* s_x,
                                              const double* s_y, const double* s_z, uint32_t a,
                                              uint32_t e) {
    Sum4 r{0.0, 0.0, 0.0, 0.0};
    for (uint32_t b = a + (threadIdx.x & 15u); b < e; b += 16u) {
        const double m = s_m[b];
        r.m += m;
        r.x += m * s_x[b];
        r.y += m * s_y[b];
        r.z += m * s_z[b];
    }
    fold_lanes<8>(r);
    fold_lanes<4>(r);
    fold_lanes<2>(r);
    fold_lanes<1>(r);
    return r;
}

// The reference's descent to a cell (child centres, octree.rs:160-215 /
// quadtree.rs:161-194):
// n levels down from a cell of half-width ext and centre c, the half-width
// halves and the centre moves by it along each axis, up where the key's
// octant bit is set, else down. w holds the digits to follow inverted and
// left-aligned (the next digit's DIM bits on top, x lowest), so a digit's
// inverted bit is the sign of that axis' move; c + (-h) is c - h exactly.
template <int DIM>
__device__ __forceinline__ void descend_w(uint64_t w, int n, double& ext, double& cx, double& cy,
                                          double& cz) {
    for (int l = 0; l < n; ++l) {
        const uint32_t top = (uint32_t)(w >> 32);
        ext = half(ext);
        const uint64_t hb = f64_bits(ext);
        const uint32_t lo = (uint32_t)hb, hi = (uint32_t)(hb >> 32);
        cx += __hiloint2double((int)(hi | ((top << (DIM - 1)) & 0x80000000u)), (int)lo);
        cy += __hiloint2double((int)(hi | ((top << (DIM - 2)) & 0x80000000u)), (int)lo);
        if (DIM == 3) cz += __hiloint2double((int)(hi | (top & 0x80000000u)), (int)lo);
        w <<= DIM;
    }
}
// Levels [from, to) of a body's path: the first key word's digits, then the
// deep key's (levels past lmax, bodies whose first words tie).
template <int DIM>
__device__ __forceinline__ void descend(uint64_t key, const DeepKey& dk, int lmax, int from, int to,
                                        double& ext, double& cx, double& cy, double& cz) {
    const int a = min(to, lmax);
    if (from < a) descend_w<DIM>(~key << (64 - DIM * (lmax - from)), a - from, ext, cx, cy, cz);
    constexpr int W = kWordLevels<DIM>;
    for (int l = max(from, lmax); l < to;) {
        const int t = l - lmax, p = t % W;
        const int n = min(to - l, W - p);
        descend_w<DIM>(~dk.w[t / W] << (64 - DIM * (W - p)), n, ext, cx, cy, cz);
        l += n;
    }
}
// Levels of a tile's common path kept per tile (the rest descend from there).
constexpr int kPathLevels = 12;

// Internal nodes of more bodies than this (one thread per node) or than
// kSmallNode (one thread per body) are queued and summed by lane groups.
constexpr uint32_t kTinyNode = 8;
constexpr uint32_t kSmallNode = 64;
// The queue: one thread per node, a node-parallel tile has at most
// kNodeQueue internal nodes; an entry is the owning body (10 bits), the
// range's last body (10 bits) and the node's place among the owner's (6
// bits). One thread per body: internal nodes of more than
// kSmallNode bodies hold disjoint body sets at one level, so a tile has at
// most kNodeTile / (kSmallNode + 1) per level; an entry (16-bit, two per
// word) is the owning body (10 bits) and the level (6 bits: an internal node
// lies above a leaf, so below the deepest level, 64).
constexpr int kNodeQueue = kNodeLds - kNodeTile;
static_assert(kNodeTile <= 1024 && kTopSlots <= 65, "queue entries: 10-bit bodies, 6-bit places");
static_assert(kTopSlots * (kNodeTile / (kSmallNode + 1)) <= 2 * kNodeQueue && kTopSlots <= 65,
              "queue bound of the body-parallel build");

// A node record as four 16-byte chunks: (cx, cy), (cz, mx), (my, mz),
// (mass, skip, level | octant << 16), or as eight 8-byte words.
static_assert(sizeof(NodeHot) == 64, "one node record per 64-byte line");
static_assert(offsetof(NodeHot, mass) == 48 && offsetof(NodeHot, skip) == 56 &&
                  offsetof(NodeHot, lvl_oct) == 60,
              "chunk layout of a node record");
__device__ __forceinline__ uint4 pack2(double a, double b) {
    const uint64_t x = f64_bits(a), y = f64_bits(b);
    return make_uint4((uint32_t)x, (uint32_t)(x >> 32), (uint32_t)y, (uint32_t)(y >> 32));
}
__device__ __forceinline__ uint2 pack1(double a) {
    const uint64_t x = f64_bits(a);
    return make_uint2((uint32_t)x, (uint32_t)(x >> 32));
}
__device__ __forceinline__ uint4 pack_tail(double mass, uint32_t skip, uint32_t level,
                                           uint32_t octant) {
    const uint64_t x = f64_bits(mass);
    return make_uint4((uint32_t)x, (uint32_t)(x >> 32), skip, level | (octant << 16));
}

// Builds the tree's internal nodes and their masses and centres of mass,
// one block per node tile (up to kNodeTile sorted bodies of one segment; the
// tiles of k_scan_gather). The leaves are the bodies themselves: the walk
// reads a leaf's position and mass from leafpos in sorted order, so no leaf
// record is written.
//
// The tile's bodies own the preorder range [tile_base[t], tile_base[t+1]) of
// internal nodes (tile base + offsets within the tile, also written out as
// first_node). Every internal node gets its geometric centre by the
// reference's descent, its level, octant and skip (= first internal node of
// the body after its range) in its record, and its body range (first body,
// body after the range) in node_span. A node's range end is found by a
// galloping search over the staged keys (keys carry their segment, so no
// range crosses one); a node that runs past the tile ends before the first
// later body starting a node at its level or above (the scanned per-level
// table), so no search leaves the tile.
//
// Bodies (keys, levels, positions, masses) are staged in LDS. A tile of at
// most kNodeQueue internal nodes is built one thread per node (the bodies
// list them first); a larger one (clustered bodies), body by body. Each
// record is written by the thread that finds its range; an internal node
// whose bodies all lie in the tile takes its mass and centre of mass from the
// sums of those bodies in body order (by its thread up to kTinyNode bodies,
// or kSmallNode body by body; else by a group of lanes after the node
// passes).
//
// An internal node whose range leaves the tile is left to k_com_span: it is
// recorded (generation-tagged) at slot tile * kTopSlots + level, and
// sc->top_levels bounds the levels k_com_span visits. For those nodes the
// block also writes the tile's partial mass sums (below).
template <int DIM>
__global__ __launch_bounds__(kNodeThreads) __attribute__((amdgpu_waves_per_eu(6))) void k_nodes(
    const SegDesc* __restrict__ desc, const uint64_t* __restrict__ seg_off,
    const uint32_t* __restrict__ tile_start, uint64_t total, const uint64_t* __restrict__ skeys_g,
    uint32_t nseg, int segbits, int lmax, const uint32_t* __restrict__ node_local,
    const uint32_t* __restrict__ tile_base, const uint32_t* __restrict__ lvl_first,
    uint32_t* __restrict__ first_node, const uint8_t* __restrict__ lo_level,
    NodeHot* __restrict__ hot, uint64_t cap, uint64_t* __restrict__ top_slot,
    uint2* __restrict__ top_last, Sum4* __restrict__ out_part, Sum4* __restrict__ in_part,
    Sum4* __restrict__ tot_part, const DeepKey* __restrict__ deep, StepCounters* __restrict__ sc,
    const double4* __restrict__ leafpos, uint2* __restrict__ node_span) {
    static_assert(kNodeTile <= 65536, "owner indices are 16-bit");
    // s_mask[L][c]: bit b set iff body 64 c + b starts a cell at level L
    // (lo <= L), for cmin[c] <= L < cmax[c]; no body of chunk c starts one
    // below cmin[c], and every body does from cmax[c] up
    __shared__ uint64_t s_mask[kTopSlots][kNodeTile / 64];
    __shared__ uint8_t s_cmin[kNodeTile / 64], s_cmax[kNodeTile / 64];
    __shared__ uint32_t s_first[kNodeTile + 1];
    __shared__ uint8_t s_lo[kNodeTile];
    __shared__ double s_bm[kNodeTile], s_bx[kNodeTile], s_by[kNodeTile], s_bz[kNodeTile];
    __shared__ uint16_t s_ilist[kNodeQueue];  // a node-parallel tile's internal nodes
    __shared__ uint32_t s_q[kNodeQueue];  // queued internal nodes (see above)
    __shared__ double s_path[kPathLevels][4];  // half-width and centre along the common path
    __shared__ uint32_t s_lonext;  // lo of the body just after the tile (0: none, or another segment)
    // per level L < s_lonext: the last body and the skip of the tile-spanning node
    __shared__ uint32_t s_span_last[kTopSlots], s_span_skip[kTopSlots];
    __shared__ uint32_t s_topl, s_nq;
    __shared__ int s_lt;  // levels of the path shared by all the tile's bodies (capped)
    __shared__ Sum4 s_w4[2][kNodeThreads / 64];
    __shared__ uint32_t s_wmin[2][kNodeThreads / 64];
    const uint64_t gen = (uint64_t)sc->gen << 32;
    const uint32_t tile = blockIdx.x, tiles = gridDim.x;
    const uint64_t b0 = tile_start[tile], b1 = tile_start[tile + 1];
    const uint32_t nb = (uint32_t)(b1 - b0);
    const uint32_t n0 = tile_base[tile];
    const uint32_t nend = tile_base[tile + 1];
    const uint32_t n1 = (uint32_t)min((uint64_t)nend, cap);
    const uint32_t nloc = n1 > n0 ? n1 - n0 : 0u;
    // node-parallel where the tile's internal nodes fit the list
    const bool node_par = nloc <= (uint32_t)kNodeQueue;
    // a node tile lies in one segment
    const SegDesc& d = desc[seg_of(seg_off, nseg, b0)];
    if (threadIdx.x == 0) {
        s_first[nb] = nend;
        s_topl = 0;
        s_nq = 0;
    }
    // Bodies: staging, checks; in a node-parallel tile each body also lists
    // its internal nodes at their tile-local places (node order). A thread
    // stages bodies t and t + kNodeThreads: every load is issued before any
    // is used. Whole waves take part (a wave holds one chunk of 64 bodies in
    // each round): its start masks.
    static_assert(kNodeTile == 2 * kNodeThreads, "two staged bodies per thread");
    const uint32_t nch = (nb + 63u) >> 6;
    uint32_t st_lo[2], st_f[2], st_fn[2];
    double4 st_p[2];
    // (indices clamped into the tile, a tile holds a body: unconditional
    // loads, so none waits on another at a branch)
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const uint32_t lc = min((uint32_t)r * kNodeThreads + threadIdx.x, nb - 1u);
        const uint64_t i = b0 + lc;
        st_lo[r] = lo_level[i];
        st_f[r] = node_local[i];
        st_fn[r] = node_local[lc + 1u < nb ? i + 1 : i];
        st_p[r] = leafpos[i];
    }
    const uint32_t lo_after = lo_level[min((uint64_t)b1, total - 1)];
    if (threadIdx.x == kNodeThreads - 1u) s_lonext = b1 < total ? lo_after : 0u;
    // Loads for the work between the staging and the build, issued now so
    // they arrive during the staging: the next tile's row of the scanned
    // per-level table (tile-spanning nodes), the tile's first and last keys
    // and the extent (the common path), and the body after the tile (the
    // coincidence check of the tile's last body).
    const bool span_t = threadIdx.x < (uint32_t)kTopSlots && b1 < total;
    const uint32_t span_nx = span_t ? lvl_first[(uint64_t)(tile + 1) * kTopSlots + threadIdx.x] : ~0u;
    constexpr uint32_t kPathThread = 2u * 64u;
    uint64_t path_ka = 0, path_kz = 0, path_e = 0;
    if (threadIdx.x == kPathThread) {
        path_ka = skeys_g[b0];
        path_kz = skeys_g[b1 - 1];
        path_e = gload(&d.st->extent_bits);
    }
    const uint32_t last_li = nb - 1u;  // (a tile holds a body)
    double4 after = make_double4(0.0, 0.0, 0.0, 0.0);
    if (threadIdx.x == (last_li & (kNodeThreads - 1u)) && b1 < d.off + d.n) after = leafpos[b1];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const uint32_t li = (uint32_t)r * kNodeThreads + threadIdx.x;
        uint32_t lo_eff = 255u;  // body 0 of the tile never ends a range in it
        if (li < nb) {
            const uint64_t i = b0 + li;
            const uint32_t lo = st_lo[r], f = st_f[r] + n0;
            const uint32_t fn = li + 1u < nb ? st_fn[r] + n0 : nend;
            if (li > 0) lo_eff = lo;
            s_lo[li] = (uint8_t)lo;
            s_first[li] = f;
            first_node[i] = f;
            if (node_par)
                for (uint32_t nd = f; nd < min(fn, n1); ++nd)
                    s_ilist[nd - n0] = (uint16_t)(li | ((nd - f) << 10));
            s_bm[li] = st_p[r].w;
            s_bx[li] = st_p[r].x;
            s_by[li] = st_p[r].y;
            s_bz[li] = st_p[r].z;
            const uint32_t hi = lo + (fn - f);  // the leaf's level
            if (fn > cap) atomicOr(&d.st->err, kErrCapacity);
            // read first: once the running maximum is reached almost no lane
            // needs the atomic (a stale read only costs an extra atomic)
            if (d.count && hi > __hip_atomic_load(&d.st->max_level, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT))
                atomicMax(&d.st->max_level, hi);
        }
        const uint32_t c = li >> 6;
        if ((c << 6) < nb) {
            uint32_t mn = lo_eff, mx = lo_eff == 255u ? 0u : lo_eff;
            for (int o = 32; o; o >>= 1) {
                mn = min(mn, (uint32_t)__shfl_xor((int)mn, o));
                mx = max(mx, (uint32_t)__shfl_xor((int)mx, o));
            }
            mn = uniform_u32(mn);
            mx = uniform_u32(mx);
            for (uint32_t L = mn; L < mx; ++L) {
                const uint64_t m = __ballot(lo_eff <= L);
                if ((threadIdx.x & 63u) == 0) s_mask[L][c] = m;
            }
            if ((threadIdx.x & 63u) == 0) {
                s_cmin[c] = (uint8_t)mn;
                s_cmax[c] = (uint8_t)mx;
            }
        }
    }
    // The node of level L < s_lonext that holds the tile's last body runs
    // into the next tile: it ends before the first later body starting a
    // node at its level or above (the scanned per-level table), and its skip
    // is that body's first node. That body lies in this segment or is the
    // next segment's first body (lo 0), whose tile follows this segment's
    // last. One thread per level, before the build.
    if (span_t && threadIdx.x < lo_after) {
        const uint64_t nxt = span_nx == ~0u ? total : (uint64_t)span_nx;
        s_span_last[threadIdx.x] = (uint32_t)(nxt - 1);
        uint32_t skip = tile_base[tiles];
        if (nxt < total) {
            const uint64_t rel = nxt - d.off;
            const uint32_t t2 =
                d.tile0 + (uint32_t)(rel < d.n ? rel / kNodeTile : (d.n + kNodeTile - 1) / kNodeTile);
            skip = node_local[nxt] + tile_base[t2];
        }
        s_span_skip[threadIdx.x] = skip;
    }
    if (threadIdx.x == kPathThread) {
        // every node of the tile lies on or below this path: its cells are
        // the ancestors of every body
        const int lt = min(key_lcp<DIM>(path_ka, path_kz, lmax), kPathLevels - 1);
        double e = bits_f64(path_e), x = 0.0, y = 0.0, z = 0.0;
        for (int l = 0;; ++l) {
            s_path[l][0] = e;
            s_path[l][1] = x;
            s_path[l][2] = y;
            s_path[l][3] = z;
            if (l == lt) break;
            descend<DIM>(path_ka, DeepKey{}, lmax, l, l + 1, e, x, y, z);  // (lt < lmax)
        }
        s_lt = lt;
    }
    // (this barrier also publishes the common path and the tile-spanning
    // levels: the build reads nothing written after it)
    __syncthreads();
    // the key of this thread's first node's owner (node-parallel build)
    const uint64_t key0 =
        node_par && threadIdx.x < nloc ? skeys_g[b0 + (s_ilist[threadIdx.x] & 1023u)] : 0ull;
    // the reference merges bodies within 1e-9 on every axis
    // (octree.rs:69-80); the device reports them instead (the tile's last
    // body against the next body of its segment, loaded above)
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const uint32_t li = (uint32_t)r * kNodeThreads + threadIdx.x;
        if (li < nb && b0 + li + 1u < d.off + d.n) {
            const double4 q = li + 1u < nb ? make_double4(s_bx[li + 1], s_by[li + 1], s_bz[li + 1], 0.0)
                                           : after;
            if (fabs(st_p[r].x - q.x) < 1e-9 && fabs(st_p[r].y - q.y) < 1e-9 &&
                fabs(st_p[r].z - q.z) < 1e-9)
                atomicOr(&d.st->err, kErrCoincident);
        }
    }
    const int lt = s_lt;

    // Chunk c's start mask at level L.
    auto chunk_mask = [&](uint32_t c, int L) -> uint64_t {
        if (L < (int)s_cmin[c]) return 0ull;
        if (L >= (int)s_cmax[c]) return ~0ull;
        return s_mask[L][c];
    };
    // The first body after local body li that starts a cell at level L
    // (lo <= L), or nb: the end of li's level-L cell in the tile.
    auto next_start = [&](uint32_t li, int L) -> uint32_t {
        const uint32_t j = li + 1u;
        if (j >= nb) return nb;
        uint32_t c = j >> 6;
        uint64_t m = chunk_mask(c, L) & (~0ull << (j & 63u));
        while (m == 0ull) {
            if (++c >= nch) return nb;
            m = chunk_mask(c, L);
        }
        return min((c << 6) + (uint32_t)__ffsll((long long)m) - 1u, nb);
    };
    // Skip of the internal node at level L owned by body i: the first node
    // after its range; its last body, and whether it leaves the tile (the
    // body after the tile shares the cell).
    auto internal_skip = [&](uint64_t i, int L, uint64_t& last, bool& spans) -> uint32_t {
        const uint32_t e = next_start((uint32_t)(i - b0), L);
        last = b0 + e - 1u;
        spans = false;
        if (e == nb && b1 < total && (int)s_lonext > L) {
            spans = true;
            last = s_span_last[L];
            return s_span_skip[L];
        }
        return s_first[e];
    };
    // The record of internal node nd (level L, owned by local body li,
    // geometric centre c, octant oct), and its body range. A tile-spanning
    // node (the only one of its level that starts in the tile and leaves it)
    // gets zero mass here and its own from k_com_span; a larger internal
    // node is queued, its mass words written by the group pass.
    auto emit = [&](uint32_t nd, uint32_t li, int L, double cx, double cy, double cz,
                    uint32_t oct) {
        Sum4 c{0.0, 0.0, 0.0, 0.0};
        bool queued = false;
        uint64_t last;
        bool spans;
        const uint32_t skip = internal_skip(b0 + li, L, last, spans);
        if (spans) {
            top_slot[(uint64_t)tile * kTopSlots + L] = gen | (uint64_t)nd;
            // its last body and that body's tile (the node lies in this segment)
            top_last[(uint64_t)tile * kTopSlots + L] =
                make_uint2((uint32_t)last, d.tile0 + (uint32_t)((last - d.off) / kNodeTile));
            atomicMax(&s_topl, (uint32_t)L + 1u);
        } else {
            const uint32_t e = (uint32_t)(last - b0) + 1u;
            if (e - li > (node_par ? kTinyNode : kSmallNode)) {
                // (the bounds above hold unless the node capacity cut the
                // tile short, an error; then the thread sums)
                const uint32_t q = atomicAdd(&s_nq, 1u);
                if (node_par && q < (uint32_t)kNodeQueue) {
                    s_q[q] = li | ((e - 1u) << 10) | ((nd - s_first[li]) << 20);
                    queued = true;
                } else if (!node_par && q < 2u * kNodeQueue) {
                    reinterpret_cast<uint16_t*>(s_q)[q] = (uint16_t)(li | ((uint32_t)L << 10));
                    queued = true;
                }
            }
            if (!queued) c = com_of(range_sum(s_bm, s_bx, s_by, s_bz, li, e));
        }
        // first body and the body after the range (global sorted indices)
        node_span[nd] = make_uint2((uint32_t)(b0 + li), (uint32_t)(last + 1));
        uint2* w = reinterpret_cast<uint2*>(hot + nd);
        const uint4 tail = pack_tail(c.m, skip, (uint32_t)L, oct);
        if (queued) {
            reinterpret_cast<uint4*>(w)[0] = pack2(cx, cy);
            w[2] = pack1(cz);
            w[7] = make_uint2(tail.z, tail.w);
        } else {
            uint4* rec = reinterpret_cast<uint4*>(w);
            rec[0] = pack2(cx, cy);
            rec[1] = pack2(cz, c.x);
            rec[2] = pack2(c.y, c.z);
            rec[3] = tail;
        }
    };

    // node nd (level L) of local body li: its centre by the reference's
    // descent (from the tile's common path), then its record
    auto build = [&](uint32_t nd, uint32_t li, int L, uint64_t key) {
        DeepKey kx{};
        if (L > lmax) kx = deep[b0 + li];
        const int f = min(L, lt);
        double ext = s_path[f][0], cx = s_path[f][1], cy = s_path[f][2], cz = s_path[f][3];
        descend<DIM>(key, kx, lmax, f, L, ext, cx, cy, cz);
        const uint32_t oct = L > 0 ? key_digit<DIM>(key, kx, lmax, L - 1) : 0u;
        emit(nd, li, L, cx, cy, cz, oct);
    };
    if (node_par) {
        // one thread per internal node (listed in node order by the bodies:
        // at most kNodeQueue of them in a node-parallel tile)
        for (uint32_t q = threadIdx.x; q < nloc; q += kNodeThreads) {
            const uint32_t ent = s_ilist[q];
            const uint32_t li = ent & 1023u, j = ent >> 10;
            build(n0 + q, li, (int)s_lo[li] + (int)j, q == threadIdx.x ? key0 : skeys_g[b0 + li]);
        }
    } else {
        // one thread per body: one descent through the body's internal nodes
        for (uint32_t li = threadIdx.x; li < nb; li += kNodeThreads) {
            const uint64_t key = skeys_g[b0 + li];
            const int lo = s_lo[li];
            const uint32_t nbase = s_first[li];
            const int cnt = (int)(s_first[li + 1] - nbase);
            const int hi = lo + cnt - 1;
            DeepKey kx{};
            if (hi > lmax) kx = deep[b0 + li];
            const int f = min(lo, lt);
            double ext = s_path[f][0], cx = s_path[f][1], cy = s_path[f][2], cz = s_path[f][3];
            descend<DIM>(key, kx, lmax, f, lo, ext, cx, cy, cz);
            for (int k = 0; k < cnt; ++k) {
                const int L = lo + k;
                const uint64_t nd = (uint64_t)nbase + (uint64_t)k;
                if (nd >= cap) break;  // reported above
                const uint32_t oct = L > 0 ? key_digit<DIM>(key, kx, lmax, L - 1) : 0u;
                emit((uint32_t)nd, li, L, cx, cy, cz, oct);
                if (k + 1 < cnt) descend<DIM>(key, kx, lmax, L, L + 1, ext, cx, cy, cz);
            }
        }
    }

    // Partial mass sums for the tile-spanning nodes (k_com_span): the tile
    // total; per level L, the suffix from the first body of the level-L
    // ancestor of the tile's last body, where that ancestor starts in the
    // tile and runs into the next one (out_part); and the prefix up to the
    // end of the level-L ancestor of the tile's first body, where that one
    // started before the tile and ends in it (in_part). Bodies come in
    // contiguous pairs per thread; every sum has a fixed association. (The
    // scans' barriers also order the node pass before the group pass below.)
    // mass sums and level of local body li (zero / ~0 when absent)
    auto body_sum = [&](uint32_t li) -> Sum4 {
        if (li >= nb) return Sum4{0.0, 0.0, 0.0, 0.0};
        const double m = s_bm[li];
        return Sum4{m, m * s_bx[li], m * s_by[li], m * s_bz[li]};
    };
    auto body_lo = [&](uint32_t li) -> uint32_t { return li < nb ? (uint32_t)s_lo[li] : ~0u; };
    // the prefix side takes bodies l0, l1 per thread, the suffix side the
    // mirrored pair r0, r1, so both are scans in thread order
    const uint32_t l0 = 2u * threadIdx.x, l1 = l0 + 1u;
    const uint32_t r1 = (uint32_t)kNodeTile - 1u - 2u * threadIdx.x, r0 = r1 - 1u;
    Sum4 pre = body_sum(l0), suf = body_sum(r0), total4;
    pre.add(body_sum(l1));
    suf.add(body_sum(r1));
    uint32_t mnp = min(body_lo(l0), body_lo(l1)), mns = min(body_lo(r0), body_lo(r1));
    block_excl_scan2<kNodeThreads>(pre, mnp, suf, mns, s_w4, s_wmin, &total4);
    {
        // prefix side: tile total, and in_part at the end of each ancestor
        // of the first body (levels <= lam0) that ends in the tile
        // (levels are at most 64: clamp "none" before comparing as int)
        const int mn = (int)min(mnp, 255u);
        if (threadIdx.x == 0) tot_part[tile] = total4;
        const int lam0 = nb ? (int)s_lo[0] - 1 : -1;
        // body li ends the level-L ancestor iff lo[li] <= L < every lo
        // between the first body and li
        if (l0 > 0 && l0 < nb)
            for (int L = (int)body_lo(l0); L <= lam0 && L < mn; ++L)
                in_part[(uint64_t)tile * kTopSlots + L] = pre;
        if (l1 < nb) {
            const int mn1 = min(mn, (int)body_lo(l0));
            if ((int)body_lo(l1) <= lam0 && (int)body_lo(l1) < mn1) {
                Sum4 p = pre;
                p.add(body_sum(l0));
                for (int L = (int)body_lo(l1); L <= lam0 && L < mn1; ++L)
                    in_part[(uint64_t)tile * kTopSlots + L] = p;
            }
        }
    }
    {
        // suffix side: out_part from the first body of each ancestor of the
        // last body (levels <= lam1) that starts in the tile and runs on
        const int mn = (int)min(mns, 255u);
        int lam1 = -1;
        if (nb && s_lonext > 0u) lam1 = min((int)s_lonext - 1, kTopSlots - 1);
        // body li starts the level-L ancestor iff lo[li] <= L < every lo
        // after it
        if (r1 < nb && (int)body_lo(r1) <= lam1 && (int)body_lo(r1) < mn) {
            Sum4 sf = suf;
            sf.add(body_sum(r1));
            for (int L = (int)body_lo(r1); L <= lam1 && L < mn; ++L)
                out_part[(uint64_t)tile * kTopSlots + L] = sf;
        }
        if (r0 < nb) {
            const int mn0 = min(mn, (int)min(body_lo(r1), 255u));
            if ((int)body_lo(r0) <= lam1 && (int)body_lo(r0) < mn0) {
                Sum4 sf = suf;
                sf.add(body_sum(r1));
                sf.add(body_sum(r0));
                for (int L = (int)body_lo(r0); L <= lam1 && L < mn0; ++L)
                    out_part[(uint64_t)tile * kTopSlots + L] = sf;
            }
        }
    }
    if (threadIdx.x == 0 && s_topl) atomicMax(&sc->top_levels, s_topl);

    // Queued internal nodes, one group of 16 lanes each: mass, centre of
    // mass.
    const uint32_t nq = min(s_nq, node_par ? (uint32_t)kNodeQueue : 2u * kNodeQueue);
    for (uint32_t q = threadIdx.x >> 4; q < nq; q += kNodeThreads / 16) {
        uint32_t li, e, nd;
        if (node_par) {
            const uint32_t ent = s_q[q];
            li = ent & 1023u;
            e = ((ent >> 10) & 1023u) + 1u;
            nd = s_first[li] + (ent >> 20);
        } else {
            const uint32_t ent = reinterpret_cast<const uint16_t*>(s_q)[q];
            li = ent & 1023u;
            const int L = (int)(ent >> 10);
            uint64_t last;
            bool spans;
            internal_skip(b0 + li, L, last, spans);  // (within the tile)
            e = (uint32_t)(last - b0) + 1u;
            nd = s_first[li] + (uint32_t)(L - (int)s_lo[li]);
        }
        const Sum4 c = com_of(range_sum_row(s_bm, s_bx, s_by, s_bz, li, e));
        if ((threadIdx.x & 15u) == 0) {
            uint2* w = reinterpret_cast<uint2*>(hot + nd);
            w[3] = pack1(c.x);
            reinterpret_cast<uint4*>(w)[2] = pack2(c.y, c.z);
            w[6] = pack1(c.m);
        }
    }
}

// Mass and centre of mass of every tile-spanning node (blocks: its owner
// tile x groups of levels, one wave per level): its owner tile's suffix from
// its first body, the totals of the tiles in between, and the prefix of the
// tile where it ends (or that tile's total when it ends with the tile).
__global__ __launch_bounds__(kBlock) void k_com_span(
    const SegDesc* __restrict__ desc, const uint64_t* __restrict__ seg_off, uint32_t nseg,
    const uint32_t* __restrict__ tile_start, NodeHot* __restrict__ hot,
    const uint2* __restrict__ top_last, const uint64_t* __restrict__ top_slot,
    const Sum4* __restrict__ out_part, const Sum4* __restrict__ in_part,
    const Sum4* __restrict__ tot_part, const StepCounters* __restrict__ sc) {
    const uint32_t t = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t gen = sc->gen;
    const int levels = (int)min(sc->top_levels, (uint32_t)kTopSlots);
    for (int L = (int)blockIdx.y * (kBlock / 64) + wave; L < levels;
         L += (int)gridDim.y * (kBlock / 64)) {
        // the slot, the node's last body and the owner tile's suffix sum
        // load together (a stale slot's are read and not used)
        const uint64_t sv = top_slot[(uint64_t)t * kTopSlots + L];
        const uint2 lt = top_last[(uint64_t)t * kTopSlots + L];
        const Sum4 own = out_part[(uint64_t)t * kTopSlots + L];
        if ((uint32_t)(sv >> 32) != gen) continue;
        const uint32_t nd = (uint32_t)sv;
        const uint64_t last = lt.x;
        const uint32_t t1 = lt.y;
        // the end tile's values, all loaded before the one needed is chosen
        const uint32_t end_start = tile_start[t1 + 1];
        const Sum4 end_tot = tot_part[t1];
        const Sum4 end_in = in_part[(uint64_t)t1 * kTopSlots + L];
        Sum4 a{0.0, 0.0, 0.0, 0.0};
        for (uint32_t tt = t + 1u + (uint32_t)lane; tt < t1; tt += 64u) a.add(tot_part[tt]);
        // butterfly over the lanes (lane 0's association is fixed)
        fold_lanes<32>(a);
        fold_lanes<16>(a);
        fold_lanes<8>(a);
        fold_lanes<4>(a);
        fold_lanes<2>(a);
        fold_lanes<1>(a);
        if (lane == 0) {
            Sum4 r = own;
            r.add(a);
            r.add(last + 1 == end_start ? end_tot : end_in);
            const double inv = 1.0 / r.m;
            hot[nd].mx = r.x * inv;
            hot[nd].my = r.y * inv;
            hot[nd].mz = r.z * inv;
            hot[nd].mass = r.m;
        }
    }
}

// ---------------------------------------------------------------- forces
struct Kahan3 {
    float sx, sy, sz, cx, cy, cz;
    __device__ __forceinline__ void zero() { sx = sy = sz = cx = cy = cz = 0.0f; }
    __device__ __forceinline__ static void add1(float& s, float& c, float v) {
        const float y = v - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ void add(float x, float y, float z) {
        add1(sx, cx, x);
        add1(sy, cy, y);
        add1(sz, cz, z);
    }
};

// The reference's opening test, `half_width / r < theta` with
// r = sqrt(dx*dx + dy*dy + dz*dz) in FP64 (octree.rs:137-140), decided
// exactly. An FP32 estimate hw * rsq(s) is within 3.3e-7 (relative) of the
// exact quotient (conversions 6e-8 each, rsq 1.2e-7, product 6e-8), and the
// FP64 quotient of the reference within 2.3e-16 of it, so outside a band of
// 1e-6 * theta the estimate decides as the reference does. Inside the band,
// or with operands outside FP32's normal range, the FP64 square root and
// division decide.
// Lanes for which `act` holds take the decision; the FP32 estimate is formed
// for every lane and only the FP64 test (lanes inside the band) is behind a
// branch, so the common path has no exec-mask handling.
__device__ __forceinline__ bool accept_cell(double hw, double s, double theta, bool act) {
    const bool inr = s > 1e-30 && s < 1e30 && hw > 1e-30 && hw < 1e30;
    const float q = (float)hw * __builtin_amdgcn_rsqf((float)s);
    const float t = (float)theta;
    const float band = 1e-6f * fabsf(t);
    const bool yes = q < t - band, no = q > t + band;
    bool take = yes;
    if (act && !(inr && (yes || no))) take = hw / sqrt(s) < theta;
    return take;
}


// Pair acceleration on body a from (b, mb), mixed precision: the separation
// b - a is taken in FP64 and rounded once; |F|/m_a = mb / (s + e) along the
// unsoftened direction (astro/src/lib.rs:84-113, with m_a cancelled). Where
// `use` is false (or the separation rounds to zero) the term is zero; where
// `take` is false the lane takes nothing from this source and its sums are
// left as they were (not even a zero is added: a compensated add of zero
// moves the sum's bits). Both by selects rather than branches: no exec-mask
// handling in the walk's loop.
__device__ __forceinline__ void pair_mixed(double ax, double ay, double az, double bx, double by,
                                              double bz, double mb, float ef, Kahan3& acc,
                                              bool take, bool use) {
    const float fx = (float)(bx - ax), fy = (float)(by - ay), fz = (float)(bz - az);
    const float s = fmaf(fz, fz, fmaf(fy, fy, fx * fx));
    const float w = use && s > 0.0f
                        ? (float)mb * __builtin_amdgcn_rsqf(s) * __builtin_amdgcn_rcpf(s + ef)
                        : 0.0f;
    Kahan3 n = acc;
    n.add(fx * w, fy * w, fz * w);
    acc.sx = take ? n.sx : acc.sx;
    acc.sy = take ? n.sy : acc.sy;
    acc.sz = take ? n.sz : acc.sz;
    acc.cx = take ? n.cx : acc.cx;
    acc.cy = take ? n.cy : acc.cy;
    acc.cz = take ? n.cz : acc.cz;
}

// Pair force on a from b exactly as astro/src/lib.rs:84-113 computes it (FP64,
// same operation order; powi(2) is x*x and powf(0.5) lowers to sqrt).
__device__ __forceinline__ void pair_fp64(double ax, double ay, double az, double am, double bx,
                                          double by, double bz, double bm, double e, double f[3]) {
    const double d0 = ax - bx, d1 = ay - by, d2 = az - bz;
    const double rn = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    const double re = d0 * d0 + d1 * d1 + d2 * d2 + e;
    const double r0 = (bx - ax) / rn, r1 = (by - ay) / rn, r2 = (bz - az) / rn;
    f[0] += r0 * 1.0 * am * bm / re;
    f[1] += r1 * 1.0 * am * bm / re;
    f[2] += r2 * 1.0 * am * bm / re;
}

// Barnes-Hut walk of one sorted body per lane, wave-coherent and stackless:
// the wave steps through the preorder node array together (one node address
// per wave, so node loads are scalar); each lane applies the reference's own
// opening test (octree.rs:137-157: accept iff half_width / |x - centre| <
// theta, at every node, own cell included) and skips an accepted subtree
// through its skip index; the wave descends into a node iff some lane opened
// it. Then the Stormer update (verlet.rs:24-82) in place on the permuted
// state, or the acceleration into acc_out (upload order) without it.
// (a waves-per-EU hint of 6: the three records in flight fit the scalar
// registers with almost no spills, at seven waves per SIMD, eight for the
// mixed walk without the near-field test; a hint of 8 spilled 34 scalar
// registers into vector lanes, whose reloads are VALU instructions in every
// iteration and cost more than the eighth wave hid)
// Levels below a body's leaf at which its near field still opens cells.
constexpr int kNearLevels = 2;

template <int DIM, bool FP64, bool INTEGRATE, bool NEAR, bool COUNT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6))) void k_walk(const SegDesc* __restrict__ desc, uint64_t total,
                                                 const uint64_t* __restrict__ skeys, int segbits,
                                                 const uint32_t* __restrict__ first_node,
                                                 const NodeHot* __restrict__ hot, uint64_t cap,
                                                 uint32_t nseg, double* __restrict__ acc_out,
                                                 uint64_t* __restrict__ ext_partial,
                                                 StepCounters* __restrict__ sc, uint64_t tag,
                                                 const uint8_t* __restrict__ leaf_lv,
                                                 uint32_t run_log2,
                                                 const double4* __restrict__ leafpos,
                                                 const uint2* __restrict__ node_span,
                                                 const uint32_t* __restrict__ svals) {
    // Workgroups go to the 8 XCDs in turn. Each XCD takes runs of 2^run_log2
    // consecutive blocks (16 by default: 4096 sorted bodies), so neighbouring
    // walks share their node records in one L2, while the runs of a batch's
    // costly segments still spread over all XCDs (whole groups of 8 runs; the
    // blocks past them keep their own index: a bijection for any grid; 0 is
    // the plain round-robin order)
    uint32_t blk = blockIdx.x;
    const uint32_t group = 8u << run_log2;
    if (blockIdx.x < gridDim.x / group * group) {
        const uint32_t x = blockIdx.x & 7u, k = blockIdx.x >> 3;
        blk = (k >> run_log2) * group + (x << run_log2) + (k & ((1u << run_log2) - 1u));
    }
    const uint64_t g = (uint64_t)blk * blockDim.x + threadIdx.x;
    bool valid = g < total;
    uint32_t s = 0;
    double x = 0.0, y = 0.0, z = 0.0, m = 0.0;
    bool fixed = true;
    uint64_t j = 0, src = 0;
    double4 pr = make_double4(0.0, 0.0, 0.0, 0.0);
    if (valid) {
        const uint64_t key = skeys[g];
        const uint32_t v = svals[g];
        s = seg_of_key(key, segbits);
        valid = s < nseg;
        src = v;
    }
    if (valid) {
        const SegDesc& d = desc[s];
        j = g - d.off;
        // the body's index in the previous order (k_scan_gather flags one
        // outside its segment; such a body is not stepped)
        src -= d.off;
        valid = src < d.n;
    }
    if (valid) {
        const SegDesc& d = desc[s];
        const double4 p = leafpos[g];
        x = p.x; y = p.y; z = p.z;
        m = p.w;
        pr = gload(d.sprv + src);  // previous position and tag, previous order
        fixed = tag_fixed(pr.w);
    }
    Kahan3 ka;
    ka.zero();
    double f64[3] = {0.0, 0.0, 0.0};
    uint32_t count = 0;
    bool pending = valid;
    while (__ballot(pending)) {
        const int lead = __ffsll((long long)__ballot(pending)) - 1;
        const uint32_t s0 = uniform_u32(__shfl(s, lead));
        const bool mine = pending && s == s0;
        const SegDesc& d0 = desc[s0];
        // the segment's bodies [b, bend) and internal nodes [k, kend) in
        // preorder (global sorted indices)
        const uint32_t bend = uniform_u32((uint32_t)(d0.off + d0.n));
        const uint32_t kend = uniform_u32(min((uint64_t)first_node[bend], cap));
        uint32_t k = uniform_u32(first_node[d0.off]);
        uint32_t b = uniform_u32((uint32_t)d0.off);
        const double E = bits_f64(d0.st->extent_bits);
        const double theta = d0.theta;
        const double e = d0.e;
        const float ef = (float)e;
        const bool walker = mine && !fixed;
        // near field (no reference counterpart; off at 0): a cell at the
        // scale of the body's own leaf cell (its level and the kNearLevels
        // below) is opened, never taken whole, when the body lies within
        // `near` widths of that leaf cell of the cell's box, so the bodies
        // that close at that scale add their own pair forces. Larger and
        // much smaller cells stay under the opening test alone: a body alone
        // in a large leaf (far out in a halo) would otherwise take the finely
        // split core beside it body by body.
        // (NEAR: some segment of the batch has one; the walk without the
        // test holds fewer registers)
        const bool nearf = NEAR && d0.near > 0.0;
        const int lleaf = nearf && walker ? (int)leaf_lv[g] : 0;
        const double margin = nearf && walker ? d0.near * ldexp(E, 1 - lleaf) : 0.0;
        uint32_t skip_until = 0;  // a lane skips the items of the bodies below this
        // The walk visits the tree in preorder as two merged sequences: the
        // next item is internal node k when k starts at body b (its span's
        // first body), else the leaf of body b. After node k the next item is
        // (b, k + 1) when some lane opened it, else (the body after its
        // range, its skip). The next record is loaded once the lanes have
        // decided (index clamped into the node array; a record past kend is
        // never used). Each step raises b or k (guarded against a corrupt
        // record), so the loop ends.
        const uint32_t clamp = (uint32_t)(cap - 1);
        NodeHot h = at32(hot, min(k, clamp));
        uint2 sp = at32(node_span, min(k, clamp));
        while (b < bend) {
            // this item's source (a taken node's centre of mass and mass, or
            // a leaf's body) and whether this lane adds it
            double bx, by, bz, bm;
            bool add = false;
            if (k < kend && sp.x == b) {
                bx = vcopy(h.mx); by = vcopy(h.my); bz = vcopy(h.mz); bm = vcopy(h.mass);
                // the opening test for every lane, its result kept by the
                // lanes still walking (no branch on the common path)
                const bool act = walker && b >= skip_until;
                const double d0x = x - h.cx, d0y = y - h.cy, d0z = z - h.cz;
                // (the level is taken from the record where it is used: left
                // to itself the compiler extracts it right after the next
                // record's load, which makes that load wait at once)
                int lvw;
                asm("s_and_b32 %0, %1, 0xffff" : "=s"(lvw) : "s"(h.lvl_oct) : "scc");
                const double hw = ldexp(E, -lvw);
                bool take = accept_cell(hw, d0x * d0x + d0y * d0y + d0z * d0z, theta, act);
                if (nearf) {
                    const int lv = lvw;
                    const double reach = hw + margin;
                    const bool close = fabs(d0x) < reach && fabs(d0y) < reach &&
                                       (DIM == 2 || fabs(d0z) < reach);
                    take = take && !(lv >= lleaf && lv <= lleaf + kNearLevels && close);
                }
                take = take && act;
                const bool open = act && !take;
                skip_until = take ? sp.y : skip_until;
                add = take;
                // the next item's index by selects (one pair of record loads,
                // no branch: each branch cost scalar instructions)
                const bool go_open = __ballot(open) || h.skip <= k || sp.y <= b;
                const uint32_t kn = go_open ? k + 1u : h.skip;
                b = uniform_u32(go_open ? b : sp.y);
                k = uniform_u32(kn);
                h = at32(hot, min(k, clamp));
                sp = at32(node_span, min(k, clamp));
            } else {
                // the leaf of body b: its exact position and mass
                const double4 lp = at32(leafpos, b);
                bx = lp.x; by = lp.y; bz = lp.z; bm = lp.w;
                add = walker && b >= skip_until;
                b = uniform_u32(b + 1u);
            }
            // (the reference skips a source at the body's own position)
            const bool other = !(bx == x && by == y && bz == z);
            if (FP64) {
                if (add) {
                    if (COUNT) ++count;
                    if (other) pair_fp64(x, y, z, m, bx, by, bz, bm, e, f64);
                }
            } else {
                // (every lane forms the term; the lanes that add keep it)
                if (COUNT) count += add ? 1u : 0u;
                pair_mixed(x, y, z, bx, by, bz, bm, ef, ka, add, other);
            }
        }
        pending = pending && !mine;
    }
    // interaction totals (debug counting): one atomic per wave and segment,
    // every lane converged here
    bool cpend = COUNT && valid && desc[s].count;
    while (__ballot(cpend)) {
        const int lead = __ffsll((long long)__ballot(cpend)) - 1;
        const uint32_t s0 = uniform_u32(__shfl(s, lead));
        const bool mine = cpend && s == s0;
        unsigned long long c = mine ? (unsigned long long)count : 0ull;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
        if ((int)(threadIdx.x & 63) == lead)
            atomicAdd((unsigned long long*)&desc[s0].st->interactions, c);
        cpend = cpend && !mine;
    }
    uint64_t emax = 0;  // the new position's max |coord| (bit pattern)
    uint32_t chunk = ~0u;
    bool nonfinite = false;
    if (valid) {
        const SegDesc& d = desc[s];
        double ax = 0.0, ay = 0.0, az = 0.0;
        if (!fixed) {
            if (FP64) {
                ax = f64[0] / m; ay = f64[1] / m; az = f64[2] / m;
            } else if (m == 0.0) {
                ax = ay = az = __builtin_nan("");
            } else {
                ax = (double)ka.sx - (double)ka.cx;
                ay = (double)ka.sy - (double)ka.cy;
                az = (double)ka.sz - (double)ka.cz;
            }
        }
        if (COUNT && d.count) gstore(d.counts + tag_id(pr.w), (int64_t)count);
        if (INTEGRATE) {
            // verlet.rs:24-82. After the Taylor first step the velocity is
            // (x' - x) / dt of the stored positions and is formed on download.
            const double dt = d.dt;
            const double dt2 = dt * dt;
            double nx, ny, nz;
            if (d.st->history == 0) {
                const double ux = gload(d.dvx + j), uy = gload(d.dvy + j), uz = gload(d.dvz + j);
                nx = x + ux * dt + 0.5 * ax * dt2;
                ny = y + uy * dt + 0.5 * ay * dt2;
                nz = z + uz * dt + 0.5 * az * dt2;
                gstore(d.dvx + j, ux + ax * dt);
                gstore(d.dvy + j, uy + ay * dt);
                gstore(d.dvz + j, uz + az * dt);
            } else {
                nx = 2.0 * x - pr.x + ax * dt2;
                ny = 2.0 * y - pr.y + ay * dt2;
                nz = 2.0 * z - pr.z + az * dt2;
            }
            // (global stores: a flat store would hold up the LDS waits of the
            // extent merge below until HBM acknowledged it)
            gstore(d.dpos + j, make_double4(nx, ny, nz, m));
            gstore(d.dprv + j, make_double4(x, y, z, pr.w));
            emax = f64_bits(fmax(fmax(fabs(nx), fabs(ny)), fabs(nz)));
            nonfinite = !isfinite(nx) || !isfinite(ny) || !isfinite(nz);
            chunk = d.chunk0 + (uint32_t)(j / kExtentChunk);
            if (nonfinite) atomicOr(&d.st->err, kErrNonFinite);
        } else {
            double* o = acc_out + 3 * (d.off + tag_id(pr.w));
            o[0] = ax; o[1] = ay; o[2] = az;
            // the permuted state keeps its positions, masses and history
            gstore(d.dpos + j, make_double4(x, y, z, m));
            gstore(d.dprv + j, pr);
        }
    }
    if (INTEGRATE) {
        // the next step's extent partials (what k_extent_partial computes):
        // per chunk the wave's maximum; a wave's first chunk is merged with
        // the other waves' in LDS (a block lies in one chunk unless a segment
        // starts in it), so a block issues one atomic per chunk
        __shared__ unsigned long long s_pv[kBlock / 64];
        __shared__ uint32_t s_pc[kBlock / 64];
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        if (lane == 0) s_pc[wave] = ~0u;
        bool pend = chunk != ~0u, first = true;
        while (__ballot(pend)) {
            const int lead = __ffsll((long long)__ballot(pend)) - 1;
            const uint32_t c0 = uniform_u32(__shfl(chunk, lead));
            const bool mine = pend && chunk == c0;
            unsigned long long v = mine ? (unsigned long long)emax : 0ull;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                const unsigned long long t = __shfl_xor(v, o);
                v = t > v ? t : v;
            }
            if (lane == lead) {
                if (first) {
                    s_pc[wave] = c0;
                    s_pv[wave] = v;
                } else {
                    atomicMax((unsigned long long*)&ext_partial[c0], v);
                }
            }
            first = false;
            pend = pend && !mine;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 0; w < kBlock / 64; ++w) {
                const uint32_t c = s_pc[w];
                if (c == ~0u) continue;
                unsigned long long v = s_pv[w];
                for (int w2 = w + 1; w2 < kBlock / 64; ++w2) {
                    if (s_pc[w2] == c) {
                        v = s_pv[w2] > v ? s_pv[w2] : v;
                        s_pc[w2] = ~0u;
                    }
                }
                atomicMax((unsigned long long*)&ext_partial[c], v);
            }
            if (blockIdx.x == 0) sc->ext_tag = tag;
        }
    }
}

// The end of a direct-sum evaluation for body i of segment d (position x, y,
// z, mass m): its acceleration from the summed forces f (FP64: sums of
// m_i m_j terms, divided by m here; mixed: accelerations), then the Stormer
// update (verlet.rs:24-82) into dst, or the acceleration into acc_out.
template <bool FP64, bool INTEGRATE>
__device__ __forceinline__ void direct_store(const SegDesc& d, uint64_t i, double x, double y,
                                             double z, double m, bool fixed, double fx,
                                             double fy, double fz, double* acc_out) {
    const double4 pr = d.sprv[i];  // previous position and tag
    double ax = 0.0, ay = 0.0, az = 0.0;
    if (!fixed) {
        if (FP64) {
            ax = fx / m; ay = fy / m; az = fz / m;
        } else if (m == 0.0) {
            ax = ay = az = __builtin_nan("");
        } else {
            ax = fx; ay = fy; az = fz;
        }
    }
    if (INTEGRATE) {
        // velocities are stored only by the Taylor first step
        const double dt = d.dt;
        const double dt2 = dt * dt;
        double nx, ny, nz;
        if (d.st->history == 0) {
            const double ux = d.svx[i], uy = d.svy[i], uz = d.svz[i];
            nx = x + ux * dt + 0.5 * ax * dt2;
            ny = y + uy * dt + 0.5 * ay * dt2;
            nz = z + uz * dt + 0.5 * az * dt2;
            d.dvx[i] = ux + ax * dt;
            d.dvy[i] = uy + ay * dt;
            d.dvz[i] = uz + az * dt;
        } else {
            nx = 2.0 * x - pr.x + ax * dt2;
            ny = 2.0 * y - pr.y + ay * dt2;
            nz = 2.0 * z - pr.z + az * dt2;
        }
        d.dpos[i] = make_double4(nx, ny, nz, m);
        // a non-finite new position fails the simulation (as the tree
        // step's walk reports it)
        if (!isfinite(nx) || !isfinite(ny) || !isfinite(nz)) atomicOr(&d.st->err, kErrNonFinite);
        d.dprv[i] = make_double4(x, y, z, pr.w);
    } else {
        double* o = acc_out + 3 * (d.off + tag_id(pr.w));
        o[0] = ax; o[1] = ay; o[2] = az;
    }
}

// simple_astro (transformers.rs:219-244): LDS-tiled all-pairs sum. Each block
// holds 256 targets of one segment and streams a range of the segment's
// sources through LDS: all of them, or one of d.nsplit contiguous ranges of
// source tiles (the tile word's top 8 bits), whose partial sums k_direct_finish
// adds in range order. Mixed precision sums each 256-source tile in FP32 and
// adds the tile partial into a compensated FP32 sum. Reads src, writes dst
// (same order).
template <bool FP64, bool INTEGRATE>
__global__ __launch_bounds__(kBlock) void k_direct(const SegDesc* __restrict__ desc,
                                                   const uint32_t* __restrict__ blk_seg,
                                                   const uint32_t* __restrict__ blk_tile,
                                                   double* __restrict__ part,
                                                   double* __restrict__ acc_out) {
    __shared__ double4 sj[kBlock];
    const uint32_t s = blk_seg[blockIdx.x];
    const SegDesc& d = desc[s];
    const uint32_t tw = blk_tile[blockIdx.x];
    const uint64_t i = (uint64_t)(tw & 0xFFFFFFu) * kBlock + threadIdx.x;
    const uint32_t sp = tw >> 24, ns = d.nsplit;
    const uint64_t src_tiles = (d.n + kBlock - 1) / kBlock;
    const uint64_t j0 = src_tiles * sp / ns * kBlock;
    const uint64_t j1 = min(src_tiles * (sp + 1) / ns * kBlock, d.n);
    const bool valid = i < d.n;
    double x = 0.0, y = 0.0, z = 0.0, m = 0.0;
    bool fixed = true;
    if (valid) {
        const double4 p = d.spos[i];
        x = p.x; y = p.y; z = p.z;
        m = p.w;
        fixed = tag_fixed(d.sprv[i].w);
    }
    const double e = d.e;
    const float ef = (float)e;
    Kahan3 ka;
    ka.zero();
    double f64[3] = {0.0, 0.0, 0.0};
    for (uint64_t t0 = j0; t0 < j1; t0 += kBlock) {
        const uint64_t jj = t0 + threadIdx.x;
        // mixed precision stages the mass as FP32 in the fourth word
        if (jj < j1) {
            const double4 p = d.spos[jj];
            sj[threadIdx.x] =
                make_double4(p.x, p.y, p.z, FP64 ? p.w : bits_f64((uint64_t)f32_bits((float)p.w)));
        }
        __syncthreads();
        const int cnt = (int)min((uint64_t)kBlock, j1 - t0);
        if (FP64) {
            for (int k = 0; k < cnt; ++k) {
                const double4 b = sj[k];
                if (!(b.x == x && b.y == y && b.z == z))
                    pair_fp64(x, y, z, m, b.x, b.y, b.z, b.w, e, f64);
            }
        } else {
            float px = 0.0f, py = 0.0f, pz = 0.0f;
#pragma unroll 4
            for (int k = 0; k < cnt; ++k) {
                const double4 b = sj[k];
                const float fx = (float)(b.x - x), fy = (float)(b.y - y), fz = (float)(b.z - z);
                const float s2 = fmaf(fz, fz, fmaf(fy, fy, fx * fx));
                // |F|/m_a = m_b / (r (r^2 + e)) as 1/r and 1/(r^2 + e)
                // separately (their product r^2 (r^2 + e)^2 leaves FP32's
                // range near r = 2.6e6 and below r = 5e-7 with e = 0); a
                // body's own entry (and any equal position, which the
                // reference skips) is selected out, not weighted
                const float w = s2 > 0.0f ? bits_f32((uint32_t)f64_bits(b.w)) *
                                                __builtin_amdgcn_rsqf(s2) *
                                                __builtin_amdgcn_rcpf(s2 + ef)
                                          : 0.0f;
                px = fmaf(fx, w, px);
                py = fmaf(fy, w, py);
                pz = fmaf(fz, w, pz);
            }
            ka.add(px, py, pz);
        }
        __syncthreads();
    }
    if (!valid) return;
    double fx, fy, fz;
    if (FP64) {
        fx = f64[0]; fy = f64[1]; fz = f64[2];
    } else {
        fx = (double)ka.sx - (double)ka.cx;
        fy = (double)ka.sy - (double)ka.cy;
        fz = (double)ka.sz - (double)ka.cz;
    }
    if (ns > 1) {
        double* o = part + 3 * (d.part_off + (uint64_t)sp * d.n + i);
        o[0] = fx; o[1] = fy; o[2] = fz;
        return;
    }
    direct_store<FP64, INTEGRATE>(d, i, x, y, z, m, fixed, fx, fy, fz, acc_out);
}

// Bodies of segments whose sources k_direct split: the partial sums added in
// range order, then the store.
template <bool FP64, bool INTEGRATE>
__global__ __launch_bounds__(kBlock) void k_direct_finish(const SegDesc* __restrict__ desc,
                                                          const uint64_t* __restrict__ seg_off,
                                                          uint32_t nseg, uint64_t total,
                                                          const double* __restrict__ part,
                                                          double* __restrict__ acc_out) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const SegDesc& d = desc[seg_of(seg_off, nseg, g)];
    const uint32_t ns = d.nsplit;
    if (ns <= 1) return;
    const uint64_t i = g - d.off;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (uint32_t sp = 0; sp < ns; ++sp) {
        const double* o = part + 3 * (d.part_off + (uint64_t)sp * d.n + i);
        fx += o[0]; fy += o[1]; fz += o[2];
    }
    const double4 p = d.spos[i];
    direct_store<FP64, INTEGRATE>(d, i, p.x, p.y, p.z, p.w, tag_fixed(d.sprv[i].w), fx, fy, fz,
                                  acc_out);
}

// Velocity of body j: stored (before and right after the Taylor first step)
// or the reference's backward difference (x - x_prev) / dt (verlet.rs:68-70).
struct VelocitySource {
    const double *vx, *vy, *vz;
    const double4* prv;
    int derive;
    double dt;
    __device__ __forceinline__ void at(uint64_t j, double x, double y, double z, double& ox,
                                       double& oy, double& oz) const {
        if (derive) {
            const double4 p = prv[j];
            ox = (x - p.x) / dt;
            oy = (y - p.y) / dt;
            oz = (z - p.z) / dt;
        } else {
            ox = vx[j];
            oy = vy[j];
            oz = vz[j];
        }
    }
};

// FP64 energy terms of one state: per block partial sums of
// {V (half of the all-pairs sum), K, Px, Py, Pz}. U(r) = -(ma mb / sqrt(e))
// atan(sqrt(e) / r) is the potential of |F| = ma mb / (r^2 + e).
__global__ __launch_bounds__(kBlock) void k_energy(const double4* __restrict__ pos,
                                                   VelocitySource vs, uint64_t n,
                                                   double e, double* __restrict__ partial) {
    __shared__ double4 sj[kBlock];
    __shared__ double red[5][kBlock];
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool valid = i < n;
    double xi = 0, yi = 0, zi = 0, mi = 0;
    if (valid) {
        const double4 p = pos[i];
        xi = p.x; yi = p.y; zi = p.z; mi = p.w;
    }
    const double se = sqrt(e);
    double V = 0.0;
    for (uint64_t t0 = 0; t0 < n; t0 += kBlock) {
        const uint64_t jj = t0 + threadIdx.x;
        if (jj < n) sj[threadIdx.x] = pos[jj];
        __syncthreads();
        const int cnt = (int)min((uint64_t)kBlock, n - t0);
        if (valid) {
            for (int k = 0; k < cnt; ++k) {
                const double4 b = sj[k];
                if (b.x == xi && b.y == yi && b.z == zi) continue;
                const double d0 = xi - b.x, d1 = yi - b.y, d2 = zi - b.z;
                const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
                const double mm = mi * b.w;
                V += e > 0.0 ? -(mm / se) * atan(se / r) : -mm / r;
            }
        }
        __syncthreads();
    }
    double K = 0, Px = 0, Py = 0, Pz = 0;
    if (valid) {
        double vx, vy, vz;
        vs.at(i, xi, yi, zi, vx, vy, vz);
        K = 0.5 * mi * (vx * vx + vy * vy + vz * vz);
        Px = mi * vx; Py = mi * vy; Pz = mi * vz;
    }
    red[0][threadIdx.x] = 0.5 * V;
    red[1][threadIdx.x] = K;
    red[2][threadIdx.x] = Px;
    red[3][threadIdx.x] = Py;
    red[4][threadIdx.x] = Pz;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int c = 0; c < 5; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x < 5) partial[5 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// AoS <-> SoA at the host boundary.
__global__ void k_unpack(const Entity* __restrict__ aos, uint64_t n, double4* __restrict__ pos,
                         double4* __restrict__ prv, double* vx, double* vy, double* vz) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Entity e = aos[i];
    pos[i] = make_double4(e.x, e.y, e.z, e.mass);
    prv[i] = make_double4(e.x, e.y, e.z, make_tag((uint32_t)i, e.fixed != 0));
    vx[i] = e.vx; vy[i] = e.vy; vz[i] = e.vz;
}

// The previous positions (Stormer history) of an uploaded state.
__global__ void k_unpack_prev(const Entity* __restrict__ aos, uint64_t n,
                              double4* __restrict__ prv) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Entity e = aos[i];
    prv[i] = make_double4(e.x, e.y, e.z, prv[i].w);
}

// The Stormer history as Entity records in upload order: the current
// record (radius, mass, id, fixed, velocity) with the previous position.
__global__ void k_pack_prev(Entity* __restrict__ out, const Entity* __restrict__ cur, uint64_t n,
                            const double4* __restrict__ prv) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double4 p = prv[j];
    const uint32_t u = tag_id(p.w);
    Entity e = cur[u];
    e.x = p.x; e.y = p.y; e.z = p.z;
    out[u] = e;
}

__global__ void k_pack(Entity* __restrict__ aos, uint64_t n, const double4* __restrict__ pos,
                       VelocitySource v, const double4* __restrict__ prv) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    Entity& e = aos[tag_id(prv[j].w)];
    const double4 p = pos[j];
    e.x = p.x; e.y = p.y; e.z = p.z;
    v.at(j, e.x, e.y, e.z, e.vx, e.vy, e.vz);
}

// Streaming copy (16-byte vectors, grid-stride): the bandwidth probe.
__global__ __launch_bounds__(kBlock) void k_copy(const uint4* __restrict__ src,
                                                 uint4* __restrict__ dst, uint64_t n16) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
        dst[i] = src[i];
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------ euler / rk4 stages
// The euler and rk4 integrators (integrators/src/euler.rs:17-49,
// rk4.rs:24-182) run their acceleration passes through the force-only
// tree or direct pass, which leaves the state in some body order (ids give
// each body's upload index) and the accelerations in upload order. These
// kernels work in the state's current order and keep the base state and the
// rk4 sums in upload order.
struct StageState {
    double4* pos;        // (x, y, z, m): the stages write x, y, z
    const double4* prv;  // the tags
    double *vx, *vy, *vz;
    uint64_t n;
};

// euler.rs:26-45: x' = x + v dt + 0.5 a dt^2 (dt.powi(2) = dt * dt),
// v' = v + a dt, every body (fixed ones have a = 0 and keep moving).
__global__ void k_euler_update(StageState st, const double* __restrict__ acc, double dt) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= st.n) return;
    const double* a = acc + 3 * (uint64_t)tag_id(st.prv[j].w);
    const double dt2 = dt * dt;
    const double4 p = st.pos[j];
    const double x = p.x, y = p.y, z = p.z;
    const double vx = st.vx[j], vy = st.vy[j], vz = st.vz[j];
    st.pos[j] = make_double4(x + vx * dt + 0.5 * a[0] * dt2, y + vy * dt + 0.5 * a[1] * dt2,
                             z + vz * dt + 0.5 * a[2] * dt2, p.w);
    st.vx[j] = vx + a[0] * dt;
    st.vy[j] = vy + a[1] * dt;
    st.vz[j] = vz + a[2] * dt;
}

// The base state (x, v) of an rk4 step, saved in upload order.
__global__ void k_rk4_begin(StageState st, double* __restrict__ base) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= st.n) return;
    double* b = base + 6 * (uint64_t)tag_id(st.prv[j].w);
    const double4 p = st.pos[j];
    b[0] = p.x; b[1] = p.y; b[2] = p.z;
    b[3] = st.vx[j]; b[4] = st.vy[j]; b[5] = st.vz[j];
}

// Stage s (0..3) of rk4 after its acceleration pass on the stage state:
// k_s = (dt * v_stage, dt * a_stage) (rk4.rs:40-51 and alike), the running
// sum k1 + 2.0 k2 + 2.0 k3 + k4 in the reference's left-to-right order
// (:162-167); then the next stage state base + w k_s (w = 0.5, 0.5, 1.0;
// :55-66, :89-100, :123-134, every body, fixed ones included), or after k4
// the result base + sum / 6.0, a fixed body keeping its position with zero
// velocity (:168-174).
__global__ void k_rk4_stage(StageState st, const double* __restrict__ acc,
                            const double* __restrict__ base, double* __restrict__ ksum, double dt,
                            int stage) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= st.n) return;
    const double4 pr = st.prv[j];
    const uint64_t u = tag_id(pr.w);
    const double* a = acc + 3 * u;
    const double* b = base + 6 * u;
    double* ks = ksum + 6 * u;
    const double k[6] = {dt * st.vx[j], dt * st.vy[j], dt * st.vz[j],
                         dt * a[0], dt * a[1], dt * a[2]};
    double sum[6];
    for (int c = 0; c < 6; ++c) {
        if (stage == 0) sum[c] = k[c];
        else if (stage == 3) sum[c] = ks[c] + k[c];
        else sum[c] = ks[c] + 2.0 * k[c];
        ks[c] = sum[c];
    }
    double o[6];
    if (stage < 3) {
        const double w = stage < 2 ? 0.5 : 1.0;
        for (int c = 0; c < 6; ++c) o[c] = b[c] + w * k[c];
    } else if (!tag_fixed(pr.w)) {
        for (int c = 0; c < 6; ++c) o[c] = b[c] + sum[c] / 6.0;
    } else {
        for (int c = 0; c < 3; ++c) o[c] = b[c];
        for (int c = 3; c < 6; ++c) o[c] = 0.0;
    }
    st.pos[j] = make_double4(o[0], o[1], o[2], st.pos[j].w);
    st.vx[j] = o[3]; st.vy[j] = o[4]; st.vz[j] = o[5];
}

// v = (x - x_prev) / dt (the Stormer velocity, verlet.rs:68-70) written into
// the stored velocity arrays, for a change of integrator.
__global__ void k_store_velocity(StageState st, double dt) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= st.n) return;
    const double4 p = st.pos[j], q = st.prv[j];
    st.vx[j] = (p.x - q.x) / dt;
    st.vy[j] = (p.y - q.y) / dt;
    st.vz[j] = (p.z - q.z) / dt;
}

// dst[k] += src[k], FP64.
__global__ void k_add_f64(double* __restrict__ dst, const double* __restrict__ src, uint64_t n) {
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (uint64_t)gridDim.x * blockDim.x)
        dst[k] += src[k];
}

}  // namespace pg

