// physim_gpu engine: device memory, batched step sequencing, hipGraph replay,
// per-kernel event timing and host-boundary conversions. Kernels live in
// pg_kernels.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "pg_common.h"
#include "pg_engine.h"
#include "pg_kernels.h"

using namespace pg;

// ------------------------------------------------------------------ errors
static thread_local std::string g_err;

int pg_set_error(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

extern "C" const char* pg_last_error(void) { return g_err.c_str(); }
extern "C" const char* pg_version(void) { return "physim_gpu 0.1.0 (gfx950)"; }

#define HIP_TRY(call)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return pg_set_error(PG_ERR_HIP, "%s failed: %s (%s:%d)", #call,                 \
                                hipGetErrorString(e_), __FILE__, __LINE__);                 \
    } while (0)

#define PG_TRY(call)              \
    do {                          \
        int rc_ = (call);         \
        if (rc_ != PG_OK) return rc_; \
    } while (0)

// ------------------------------------------------------------ engine types
struct EventPair {
    hipEvent_t a, b;
};

struct pg_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t budget = 0;
    std::vector<pg_sim*> sims;  // live simulations; freed with the engine
    uint64_t sims_bytes = 0;
    uint64_t scratch_bytes = 0;
    std::recursive_mutex mu;
    bool graphs = true;
    bool profiling = false;

    // batch scratch
    uint64_t cap_bodies = 0, cap_nodes = 0;
    uint32_t cap_segs = 0, cap_blocks = 0;
    void* scratch = nullptr;
    uint64_t* keys[2] = {nullptr, nullptr};
    uint32_t* vals[2] = {nullptr, nullptr};
    uint32_t* first_node = nullptr;
    uint8_t* lo_level = nullptr;
    uint8_t* leaf_lv = nullptr;    // leaf level of each sorted body (near-field walks)
    double4* leafpos = nullptr;    // position and mass of each sorted body (the leaves)
    uint2* node_span = nullptr;    // per internal node: first body, body after its range
    uint32_t* sort_cnt = nullptr;  // per-tile digit counts, digit-major
    uint32_t* sort_tot = nullptr;  // per-digit totals of one pass
    uint16_t* bkt = nullptr;       // sample sort: bucket of each key
    uint64_t* spl = nullptr;       // sample sort: splitters
    uint64_t* spc = nullptr;       // sample sort: coarse splitter table (many buckets)
    uint64_t* quant = nullptr;     // sample sort: finer quantiles per bucket
    uint32_t* bcount = nullptr;    // sample sort: keys per bucket
    uint32_t* bfill = nullptr;     // sample sort: reserved keys per bucket
    uint32_t* bbase = nullptr;     // sample sort: first key of each bucket
    uint32_t max_buckets = kMaxWideBuckets;  // larger batches use the radix sort
    uint32_t lds_buckets = kMaxBuckets;  // more buckets: splitters in HBM (coarse table in LDS)
    uint32_t bucket_keys = kBucketKeys;  // keys per sample-sort bucket (mean)
    uint32_t walk_run_log2 = 4;          // walk blocks per XCD run: 2^walk_run_log2
    uint64_t ext_valid = 0;  // batch whose extent partials the last step left (0: none)
    uint64_t sorted_valid = 0;  // batch whose sample-sorted keys keys[1] holds (0: none)
    uint32_t bucket_cap = kBucketCap;    // larger buckets are sorted in HBM
    uint32_t* node_local = nullptr;  // per-body node offset within its tile
    uint32_t* tile_tot = nullptr;    // nodes per tile
    uint32_t* tile_base = nullptr;   // first node of each tile (+ total)
    uint32_t* tile_start = nullptr;  // first body of each node tile (+ total)
    uint32_t* lvl_first = nullptr;   // per tile and level: first node-starting body, scanned
    DeepKey* deep = nullptr;         // deep keys of bodies whose first key words tie
    NodeHot* hot = nullptr;
    double* acc_out = nullptr;
    SegDesc* desc[2] = {nullptr, nullptr};
    uint64_t* seg_off = nullptr;
    uint32_t* blk_seg = nullptr;
    uint32_t* blk_tile = nullptr;
    uint32_t* ext_seg = nullptr;
    uint32_t* ext_chunk = nullptr;
    uint64_t* ext_partial = nullptr;
    uint64_t* top_slot = nullptr;  // generation-tagged tile-spanning node per (tile, level)
    uint2* top_last = nullptr;     // its last body and that body's tile
    Sum4* out_part = nullptr;      // per tile and level: suffix sums of spanning nodes
    Sum4* in_part = nullptr;       // per tile and level: prefix sums of spanning nodes
    Sum4* tot_part = nullptr;      // per tile: mass sums of all its bodies
    StepCounters* sc = nullptr;
    uint32_t* err_global = nullptr;
    uint32_t cap_chunks = 0;
    uint32_t cap_tiles = 0;       // node tiles the scratch holds
    uint64_t cap_top = 0;

    // host mirrors
    std::vector<SegDesc> h_desc[2];
    std::vector<uint64_t> h_seg_off;
    std::vector<uint32_t> h_blk_seg, h_blk_tile, h_ext_seg, h_ext_chunk, h_tile_start;

    // graph cache (two entries: the two parities of an odd step count)
    struct GraphEntry {
        std::vector<uint64_t> sig;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<GraphEntry> graph_cache;

    // per-kernel timing
    std::vector<EventPair> ev_pool;
    std::vector<std::pair<const char*, EventPair>> ev_pending;
    std::map<std::string, std::pair<double, uint64_t>> prof;

    // last build (for tree / Morton export)
    pg_sim* last_sim = nullptr;
    uint64_t last_total = 0;
    int last_sorted = 0;   // key buffer holding the sorted keys
    int last_dim = 3;
    int last_lmax = 21;

    // transient simulation of the transform path
    pg_sim* xform = nullptr;

    // the scratch is the library's static device array (allocation-free
    // transform): never grown or freed
    bool fixed_scratch = false;
    // per-call host tables of run_batch, kept so that a call whose sizes fit
    // what earlier calls held allocates nothing on the host either
    std::vector<pg_sim*> h_sims;
    std::vector<uint32_t> h_nsplit, h_chunk0, h_nchunk, h_tile0;
    std::vector<uint64_t> h_part_off;
};

struct pg_sim {
    pg_engine* eng = nullptr;
    uint64_t n = 0;
    pg_params p{};
    void* block = nullptr;
    uint64_t bytes = 0;
    double4* pos[2];  // (x, y, z, m) per body, in the last build's Morton order
    double4* prv[2];  // (x_prev, y_prev, z_prev, tag: upload index and fixed flag)
    double *vx[2], *vy[2], *vz[2];
    Entity* aos = nullptr;
    SegState* st = nullptr;
    int64_t* counts = nullptr;
    bool counting = false;
    int cur = 0;
    bool uploaded = false;
    double dt_last = 0.0;  // dt of the last integrating step
    int32_t integ = PG_INTEG_VERLET;
    double near = 0.0;  // near field in leaf widths (0: the reference's walk)
    double* stage = nullptr;   // euler / rk4: base state and rk4 sums (12 doubles per body)
    uint64_t stage_steps = 0;  // steps taken by the euler / rk4 stage loop
};

// ---------------------------------------------------------------- sizing
static inline int key_dim(int kind) { return kind == PG_KIND_QUADTREE ? 2 : 3; }

static inline int ceil_log2(uint32_t v) {
    int b = 0;
    while ((1u << b) < v) ++b;
    return b;
}

static inline int key_lmax(int dim, int segbits) {
    int l = (64 - segbits) / dim;
    int cap = dim == 3 ? 21 : 32;
    return l < cap ? l : cap;
}

static inline uint64_t align256(uint64_t v) { return (v + 255) & ~255ull; }

static uint64_t state_bytes(uint64_t n) {
    // two copies (position and previous-position records, three velocity
    // arrays), the uploaded AoS, counts, SegState
    return 2 * (align256(n * 32) * 2 + align256(n * 8) * 3) + align256(n * 80) +
           align256(n * 8) + 256;
}

// Nodes of a tree over n bodies whose neighbours share at most kMaxLcp
// levels: n leaves plus the total ascent of the common-prefix sequence, at
// most half its total variation.
static uint64_t nodes_bound(uint64_t n, int dim, int lmax) {
    (void)dim;
    (void)lmax;
    return n + (n * (uint64_t)(kMaxLcp + 1) + 1) / 2 + 1;
}

extern "C" uint64_t pg_sim_bytes(uint64_t n, int32_t kind) {
    const int dim = key_dim(kind);
    const uint64_t scratch =
        kind == PG_KIND_DIRECT
            ? n * 24
            : n * (8 * 2 + 4 * 2 + 4 + 4 + 1 + 8 + 24) + nodes_bound(n, dim, key_lmax(dim, 0)) *
                                                      sizeof(NodeHot);
    return state_bytes(n) + scratch;
}

// --------------------------------------------------------------- engine
extern "C" int pg_engine_create(int device, uint64_t budget, pg_engine** out) {
    if (!out) return pg_set_error(PG_ERR_ARG, "out is NULL");
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev)
        return pg_set_error(PG_ERR_ARG, "device %d out of range (%d devices)", device, ndev);
    HIP_TRY(hipSetDevice(device));
    pg_engine* e = new pg_engine();
    e->device = device;
    hipError_t he = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (he != hipSuccess) {
        delete e;
        return pg_set_error(PG_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(he));
    }
    size_t freeb = 0, totalb = 0;
    HIP_TRY(hipMemGetInfo(&freeb, &totalb));
    e->budget = budget ? budget : (uint64_t)(0.9 * (double)totalb);
    *out = e;
    return PG_OK;
}

static void free_scratch(pg_engine* e) {
    for (auto& g : e->graph_cache)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    e->graph_cache.clear();
    if (e->scratch && !e->fixed_scratch) (void)hipFree(e->scratch);
    e->scratch = nullptr;
    e->scratch_bytes = 0;
    e->cap_bodies = e->cap_nodes = 0;
    e->cap_segs = e->cap_blocks = 0;
}

extern "C" void pg_engine_destroy(pg_engine* e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    (void)hipStreamSynchronize(e->stream);
    e->xform = nullptr;
    while (!e->sims.empty()) pg_sim_destroy(e->sims.back());
    free_scratch(e);
    for (auto& ev : e->ev_pool) {
        (void)hipEventDestroy(ev.a);
        (void)hipEventDestroy(ev.b);
    }
    for (auto& ev : e->ev_pending) {
        (void)hipEventDestroy(ev.second.a);
        (void)hipEventDestroy(ev.second.b);
    }
    (void)hipStreamDestroy(e->stream);
    delete e;
}

extern "C" int pg_engine_memory(pg_engine* e, uint64_t* used, uint64_t* budget) {
    if (!e) return pg_set_error(PG_ERR_ARG, "engine is NULL");
    if (used) *used = e->sims_bytes + e->scratch_bytes;
    if (budget) *budget = e->budget;
    return PG_OK;
}

extern "C" int pg_engine_set_graphs(pg_engine* e, int enable) {
    if (!e) return pg_set_error(PG_ERR_ARG, "engine is NULL");
    e->graphs = enable != 0;
    return PG_OK;
}

extern "C" int pg_engine_set_sort(pg_engine* e, uint32_t max_buckets, uint32_t bucket_cap) {
    if (!e) return pg_set_error(PG_ERR_ARG, "engine is NULL");
    if (max_buckets > kMaxWideBuckets || bucket_cap > kBucketCap)
        return pg_set_error(PG_ERR_ARG, "sample sort limits are %d buckets of %d keys in LDS",
                            kMaxWideBuckets, kBucketCap);
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    (void)hipSetDevice(e->device);
    (void)hipStreamSynchronize(e->stream);
    // captured steps hold the previous choice
    for (auto& g : e->graph_cache)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    e->graph_cache.clear();
    e->sorted_valid = 0;  // the bucket layout may change
    e->max_buckets = max_buckets;
    e->bucket_cap = bucket_cap;
    return PG_OK;
}

// Drops captured steps and the sort/extent tags after a change that alters
// the launch sequence (kernel arguments are fixed at capture).
static void invalidate_launches(pg_engine* e) {
    (void)hipStreamSynchronize(e->stream);
    for (auto& g : e->graph_cache)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    e->graph_cache.clear();
    e->sorted_valid = 0;
    e->ext_valid = 0;
}

extern "C" int pg_engine_set_tuning(pg_engine* e, const pg_tuning* t) {
    if (!e || !t) return pg_set_error(PG_ERR_ARG, "engine or tuning is NULL");
    if (t->max_buckets > kMaxWideBuckets || t->bucket_keys < kMinBucketKeys ||
        t->bucket_keys > kBucketCap ||
        t->walk_run_log2 > 8 || t->lds_buckets > kMaxBuckets)
        return pg_set_error(PG_ERR_ARG,
                            "tuning out of range (max_buckets <= %d, bucket_keys in [256, %d], "
                            "walk_run_log2 <= 8, lds_buckets <= %d)",
                            kMaxWideBuckets, kBucketCap, kMaxBuckets);
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    HIP_TRY(hipSetDevice(e->device));
    invalidate_launches(e);
    e->max_buckets = t->max_buckets;
    e->bucket_keys = t->bucket_keys;
    e->walk_run_log2 = t->walk_run_log2;
    e->lds_buckets = t->lds_buckets ? t->lds_buckets : (uint32_t)kMaxBuckets;
    return PG_OK;
}

extern "C" int pg_engine_get_tuning(pg_engine* e, pg_tuning* t) {
    if (!e || !t) return pg_set_error(PG_ERR_ARG, "engine or tuning is NULL");
    t->max_buckets = e->max_buckets;
    t->bucket_keys = e->bucket_keys;
    t->walk_run_log2 = e->walk_run_log2;
    t->lds_buckets = e->lds_buckets;
    return PG_OK;
}

extern "C" int pg_engine_set_profiling(pg_engine* e, int enable) {
    if (!e) return pg_set_error(PG_ERR_ARG, "engine is NULL");
    e->profiling = enable != 0;
    return PG_OK;
}

extern "C" int pg_engine_profile_reset(pg_engine* e) {
    if (!e) return pg_set_error(PG_ERR_ARG, "engine is NULL");
    e->prof.clear();
    return PG_OK;
}

extern "C" int pg_engine_profile_read(pg_engine* e, char* names, size_t cap, double* ms,
                                      uint64_t* launches, uint32_t max_kernels, uint32_t* count) {
    if (!e) return pg_set_error(PG_ERR_ARG, "engine is NULL");
    std::string all;
    uint32_t k = 0;
    for (auto& kv : e->prof) {
        if (k < max_kernels) {
            if (ms) ms[k] = kv.second.first;
            if (launches) launches[k] = kv.second.second;
            all += kv.first;
            all += '\n';
        }
        ++k;
    }
    if (names && cap) {
        size_t len = std::min(cap - 1, all.size());
        memcpy(names, all.data(), len);
        names[len] = 0;
    }
    if (count) *count = k;
    return PG_OK;
}

extern "C" int pg_engine_synchronize(pg_engine* e) {
    if (!e) return pg_set_error(PG_ERR_ARG, "engine is NULL");
    HIP_TRY(hipStreamSynchronize(e->stream));
    return PG_OK;
}

extern "C" int pg_engine_copy_probe(pg_engine* e, uint64_t bytes, uint32_t reps, double* gbps) {
    if (!e || bytes < 16 || reps == 0) return pg_set_error(PG_ERR_ARG, "bad copy probe arguments");
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    HIP_TRY(hipSetDevice(e->device));
    const uint64_t n16 = bytes / 16;
    void *a = nullptr, *b = nullptr;
    HIP_TRY(hipMalloc(&a, n16 * 16));
    hipError_t he = hipMalloc(&b, n16 * 16);
    if (he != hipSuccess) {
        (void)hipFree(a);
        return pg_set_error(PG_ERR_NOMEM, "copy probe: %s", hipGetErrorString(he));
    }
    hipEvent_t t0, t1;
    (void)hipEventCreate(&t0);
    (void)hipEventCreate(&t1);
    const unsigned grid = 4096;
    (void)hipMemsetAsync(a, 1, n16 * 16, e->stream);
    k_copy<<<grid, kBlock, 0, e->stream>>>((const uint4*)a, (uint4*)b, n16);  // warm-up
    (void)hipEventRecord(t0, e->stream);
    for (uint32_t r = 0; r < reps; ++r)
        k_copy<<<grid, kBlock, 0, e->stream>>>((const uint4*)a, (uint4*)b, n16);
    (void)hipEventRecord(t1, e->stream);
    he = hipEventSynchronize(t1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, t0, t1);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (he != hipSuccess) return pg_set_error(PG_ERR_HIP, "copy probe: %s", hipGetErrorString(he));
    if (gbps) *gbps = 2.0 * (double)(n16 * 16) * reps / (ms * 1e-3) / 1e9;
    return PG_OK;
}

// The batch scratch layout for `bodies` bodies, `nodes` tree nodes, `segs`
// segments and `blocks` direct-sum blocks, carved from one block at `base`
// (sizes only when base is NULL). Returns the block's size in bytes.
static uint64_t carve_scratch(pg_engine* e, char* b, uint64_t bodies, uint64_t nodes,
                              uint32_t segs, uint32_t blocks) {
    const uint64_t sort_tiles = (bodies + kSortTile - 1) / kSortTile;
    // node tiles start at every segment's first body: at most bodies/T + segs
    const uint64_t node_tiles = bodies / kNodeTile + segs + 1;
    uint64_t off = 0;
    auto take = [&](uint64_t bytes) {
        uint64_t o = off;
        off += align256(bytes);
        return o;
    };
    const uint64_t o_k0 = take(bodies * 8), o_k1 = take(bodies * 8);
    const uint64_t o_v0 = take(bodies * 4), o_v1 = take(bodies * 4);
    const uint64_t o_fn = take((bodies + 1) * 4), o_lo = take(bodies), o_ll = take(bodies);
    const uint64_t o_ss = take(sort_tiles * 256 * 4), o_st = take(256 * 4);
    // (a batch of `bodies` keys has at most bodies / kMinBucketKeys + 1 buckets)
    const uint64_t maxb = std::min<uint64_t>(kMaxWideBuckets, bodies / kMinBucketKeys + 2);
    const uint64_t o_bk = take(bodies * 2), o_sp = take(maxb * 8);
    const uint64_t o_spc = take(maxb / kWideGroup * 8 + 8);
    const uint64_t o_bc = take(maxb * 4), o_bf = take(maxb * 4), o_bb = take(maxb * 4);
    const uint64_t o_qt = take(maxb * kQuantiles * 8);
    const uint64_t o_nl = take((bodies + 1) * 4), o_tt = take((node_tiles + 2) * 4);
    const uint64_t o_tb = take((node_tiles + 2) * 4), o_ts = take((node_tiles + 2) * 4);
    const uint64_t o_op = take((node_tiles + 2) * kTopSlots * sizeof(Sum4));
    const uint64_t o_ip = take((node_tiles + 2) * kTopSlots * sizeof(Sum4));
    const uint64_t o_tp = take((node_tiles + 2) * sizeof(Sum4));
    const uint64_t o_lf = take((node_tiles + 2) * kTopSlots * 4);
    const uint64_t o_dk = take(bodies * sizeof(DeepKey));
    const uint64_t o_hot = take(nodes * sizeof(NodeHot));
    const uint64_t o_span = take(nodes * sizeof(uint2)), o_lp = take(bodies * sizeof(double4));
    const uint64_t o_acc = take(bodies * 24);
    const uint64_t o_d0 = take(segs * sizeof(SegDesc)), o_d1 = take(segs * sizeof(SegDesc));
    const uint64_t o_so = take((segs + 1) * 8);
    const uint64_t o_bs = take(blocks * 4), o_bt = take(blocks * 4);
    const uint64_t o_sc = take(sizeof(StepCounters)), o_eg = take(256);
    // extent chunks never straddle segments: at most bodies/C + segs of them
    const uint64_t chunks = bodies / kExtentChunk + segs + 1;
    const uint64_t o_es = take(chunks * 4), o_ec = take(chunks * 4), o_ep = take(chunks * 8);
    // a tile boundary is crossed by at most one node per level (levels 0..32)
    const uint64_t top = (node_tiles + 2) * kTopSlots;
    const uint64_t o_tl = take(top * 8), o_tla = take(top * 8);
    if (!b) return off;
    e->keys[0] = (uint64_t*)(b + o_k0);
    e->keys[1] = (uint64_t*)(b + o_k1);
    e->vals[0] = (uint32_t*)(b + o_v0);
    e->vals[1] = (uint32_t*)(b + o_v1);
    e->first_node = (uint32_t*)(b + o_fn);
    e->lo_level = (uint8_t*)(b + o_lo);
    e->leaf_lv = (uint8_t*)(b + o_ll);
    e->sort_cnt = (uint32_t*)(b + o_ss);
    e->sort_tot = (uint32_t*)(b + o_st);
    e->bkt = (uint16_t*)(b + o_bk);
    e->spl = (uint64_t*)(b + o_sp);
    e->spc = (uint64_t*)(b + o_spc);
    e->quant = (uint64_t*)(b + o_qt);
    e->bcount = (uint32_t*)(b + o_bc);
    e->bfill = (uint32_t*)(b + o_bf);
    e->bbase = (uint32_t*)(b + o_bb);
    e->node_local = (uint32_t*)(b + o_nl);
    e->tile_tot = (uint32_t*)(b + o_tt);
    e->tile_base = (uint32_t*)(b + o_tb);
    e->tile_start = (uint32_t*)(b + o_ts);
    e->lvl_first = (uint32_t*)(b + o_lf);
    e->deep = (DeepKey*)(b + o_dk);
    e->hot = (NodeHot*)(b + o_hot);
    e->node_span = (uint2*)(b + o_span);
    e->leafpos = (double4*)(b + o_lp);
    e->acc_out = (double*)(b + o_acc);
    e->desc[0] = (SegDesc*)(b + o_d0);
    e->desc[1] = (SegDesc*)(b + o_d1);
    e->seg_off = (uint64_t*)(b + o_so);
    e->blk_seg = (uint32_t*)(b + o_bs);
    e->blk_tile = (uint32_t*)(b + o_bt);
    e->sc = (StepCounters*)(b + o_sc);
    e->err_global = (uint32_t*)(b + o_eg);
    e->ext_seg = (uint32_t*)(b + o_es);
    e->ext_chunk = (uint32_t*)(b + o_ec);
    e->ext_partial = (uint64_t*)(b + o_ep);
    e->top_slot = (uint64_t*)(b + o_tl);
    e->top_last = (uint2*)(b + o_tla);
    e->out_part = (Sum4*)(b + o_op);
    e->in_part = (Sum4*)(b + o_ip);
    e->tot_part = (Sum4*)(b + o_tp);
    e->cap_chunks = (uint32_t)chunks;
    e->cap_tiles = (uint32_t)node_tiles;
    e->cap_top = top;
    e->cap_bodies = bodies;
    e->cap_nodes = nodes;
    e->cap_segs = segs;
    e->cap_blocks = blocks;
    e->ext_valid = 0;
    e->sorted_valid = 0;
    e->last_sim = nullptr;
    return off;
}

// Zeroes the scratch's versioned status words once, so no stale tag can
// match (the fill happens on the engine's stream).
static int clear_scratch_status(pg_engine* e) {
    HIP_TRY(hipMemsetAsync(e->sc, 0, sizeof(StepCounters), e->stream));
    HIP_TRY(hipMemsetAsync(e->err_global, 0, 256, e->stream));
    HIP_TRY(hipMemsetAsync(e->top_slot, 0, e->cap_top * 8, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return PG_OK;
}

// Grows the batch scratch to hold `bodies` bodies, `nodes` tree nodes, `segs`
// segments and `blocks` direct-sum blocks. One allocation, carved in place.
// An engine on a fixed scratch (the allocation-free path) never grows: a
// batch beyond its capacities is an error.
static int ensure_scratch(pg_engine* e, uint64_t bodies, uint64_t nodes, uint32_t segs,
                          uint32_t blocks) {
    if (bodies <= e->cap_bodies && nodes <= e->cap_nodes && segs <= e->cap_segs &&
        blocks <= e->cap_blocks && e->scratch)
        return PG_OK;
    if (e->fixed_scratch)
        return pg_set_error(PG_ERR_ARG,
                            "batch exceeds the fixed scratch (%llu bodies, %llu nodes, %u blocks)",
                            (unsigned long long)e->cap_bodies, (unsigned long long)e->cap_nodes,
                            e->cap_blocks);
    bodies = std::max<uint64_t>(bodies, e->cap_bodies);
    nodes = std::max<uint64_t>(nodes, e->cap_nodes);
    segs = std::max<uint32_t>(segs, e->cap_segs);
    blocks = std::max<uint32_t>(blocks, e->cap_blocks);
    bodies = std::max<uint64_t>(bodies, 1024);
    segs = std::max<uint32_t>(segs, 4);
    blocks = std::max<uint32_t>(blocks, 16);
    nodes = std::max<uint64_t>(nodes, 1024);
    const uint64_t total = carve_scratch(e, nullptr, bodies, nodes, segs, blocks);
    if (e->sims_bytes + total > e->budget)
        return pg_set_error(PG_ERR_BUDGET, "scratch of %llu bytes exceeds the HBM budget",
                            (unsigned long long)total);
    (void)hipStreamSynchronize(e->stream);
    free_scratch(e);
    void* p = nullptr;
    hipError_t he = hipMalloc(&p, total);
    if (he != hipSuccess)
        return pg_set_error(PG_ERR_NOMEM, "hipMalloc(%llu) for scratch: %s",
                            (unsigned long long)total, hipGetErrorString(he));
    e->scratch = p;
    e->scratch_bytes = total;
    carve_scratch(e, (char*)p, bodies, nodes, segs, blocks);
    return clear_scratch_status(e);
}

// ----------------------------------------------------------- kernel timing
static EventPair ev_get(pg_engine* e) {
    if (!e->ev_pool.empty()) {
        EventPair p = e->ev_pool.back();
        e->ev_pool.pop_back();
        return p;
    }
    EventPair p;
    (void)hipEventCreate(&p.a);
    (void)hipEventCreate(&p.b);
    return p;
}

template <class F>
static void launch(pg_engine* e, bool capturing, const char* name, F&& f) {
    if (e->profiling && !capturing) {
        EventPair p = ev_get(e);
        (void)hipEventRecord(p.a, e->stream);
        f();
        (void)hipEventRecord(p.b, e->stream);
        e->ev_pending.push_back({name, p});
    } else {
        f();
    }
}

static void ev_collect(pg_engine* e) {
    for (auto& pe : e->ev_pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, pe.second.a, pe.second.b) == hipSuccess) {
            auto& slot = e->prof[pe.first];
            slot.first += ms;
            slot.second += 1;
        }
        e->ev_pool.push_back(pe.second);
    }
    e->ev_pending.clear();
}

// ---------------------------------------------------------- step sequence
// Source splits of a direct-sum segment of n bodies: small segments are too
// few target tiles to fill the GPU, so up to kMaxDirectSplit blocks share a
// target tile, each summing a contiguous range of source tiles. The count
// follows from the segment's own size, so a segment sums alike alone and in
// a batch.
constexpr uint32_t kMaxDirectSplit = 4;
static uint32_t direct_splits(uint64_t n) {
    const uint64_t tiles = (n + kBlock - 1) / kBlock;
    uint64_t s = n ? ((1ull << 19) + n - 1) / n : 1;
    s = std::min<uint64_t>(std::min<uint64_t>(s, kMaxDirectSplit), tiles);
    return (uint32_t)std::max<uint64_t>(s, 1);
}

struct BatchPlan {
    int kind = PG_KIND_OCTREE;
    int precision = PG_PREC_MIXED;
    int dim = 3, segbits = 0, lmax = 21, passes = 8;
    uint32_t buckets = 0;  // sample sort with this many buckets (0: radix sort)
    uint64_t sort_tag = 0; // identifies the batch (sims, sizes) for the sample sort's splitters
    uint32_t nseg = 0;
    uint64_t total = 0;
    uint32_t nblocks = 0;
    uint32_t nchunks = 0;
    uint32_t ntiles = 0;  // node tiles
    uint32_t max_split = 1;  // direct sum: the most source splits of a segment
    bool integrate = true;
    bool near = false;   // some segment walks with a near field (the walk that tests for it)
    bool count = false;  // some segment counts interactions (the walk that counts)
};

static inline unsigned grid_for(uint64_t n, unsigned cap) {
    uint64_t g = (n + kBlock - 1) / kBlock;
    if (g < 1) g = 1;
    return (unsigned)std::min<uint64_t>(g, cap);
}

// `extent`: the extent partials may be stale (the last step did not leave
// them), so the extent pass runs. `fused`: keys[1] holds this batch's
// previous sample-sorted keys, so k_keys takes the splitters from them and
// counts the buckets itself (no k_splitters, no k_bucket_count).
template <int DIM, bool FP64, bool INTEG>
static void enqueue_tree_step(pg_engine* e, const BatchPlan& bp, int q, bool cap, bool extent,
                              bool fused) {
    hipStream_t st = e->stream;
    SegDesc* desc = e->desc[q];
    const uint64_t total = bp.total;
    if (extent)
        launch(e, cap, "extent", [&] {
            k_extent_partial<<<std::min<uint32_t>(bp.nchunks, kExtentBlocks), kBlock, 0, st>>>(
                desc, e->ext_seg, e->ext_chunk, e->ext_partial, e->sc, bp.sort_tag, bp.nchunks);
        });
    const unsigned sort_tiles = (unsigned)((total + kSortTile - 1) / kSortTile);
    // more buckets than a key tile holds splitters for: the splitters in HBM
    // behind a coarse table (k_splitters_wide), counted by k_keys where the
    // previous step's sorted keys give them (fused), else by k_bucket_count_wide
    const bool wide = bp.buckets > e->lds_buckets;
    const bool fuse = fused && bp.buckets;
    if (fuse && wide)
        launch(e, cap, "splitters", [&] {
            k_splitters_wide<<<1, kSplitThreads, 0, st>>>(e->keys[1], e->keys[0], total,
                                                          bp.buckets, e->spl, e->spc, e->bcount,
                                                          e->bfill, e->sc, 1);
        });
    launch(e, cap, "keys", [&] {
#define PG_KEYS(F)                                                                             \
    k_keys<DIM, F><<<sort_tiles, kKeyThreads, 0, st>>>(                                        \
        desc, e->seg_off, bp.nseg, total, bp.segbits, bp.lmax, e->ext_partial, e->sc,          \
        INTEG ? 1 : 0, bp.buckets ? bp.sort_tag : 0ull, e->keys[0], e->vals[0], e->sort_cnt,   \
        sort_tiles, e->keys[1], bp.buckets, e->bkt, e->bcount, e->quant, e->spl, e->spc)
        if (!fuse) PG_KEYS(0);
        else if (!wide) PG_KEYS(1);
        else PG_KEYS(2);
#undef PG_KEYS
    });
    if (bp.buckets && wide) {
        const uint32_t nb = bp.buckets;
        if (!fuse) {
            launch(e, cap, "splitters", [&] {
                k_splitters_wide<<<1, kSplitThreads, 0, st>>>(e->keys[1], e->keys[0], total, nb,
                                                              e->spl, e->spc, e->bcount, e->bfill,
                                                              e->sc, 0);
            });
            launch(e, cap, "bucket_count", [&] {
                k_bucket_count_wide<<<sort_tiles, kSortThreads, 0, st>>>(
                    e->keys[0], total, e->spl, e->spc, nb, e->bkt, e->bcount, e->keys[1],
                    e->quant, e->sc);
            });
        }
        launch(e, cap, "bucket_base", [&] {
            k_bucket_base<<<1, kSplitThreads, 0, st>>>(e->bcount, nb, e->bbase);
        });
        launch(e, cap, "bucket_scatter", [&] {
            k_bucket_scatter_wide<<<sort_tiles, kSortThreads, 0, st>>>(
                e->keys[0], e->vals[0], e->bkt, total, e->bbase, e->bfill, e->keys[1], e->vals[1]);
        });
    } else if (bp.buckets) {
        const uint32_t nb = bp.buckets;
        if (!fused) {
            launch(e, cap, "splitters", [&] {
                k_splitters<<<1, kSplitThreads, 0, st>>>(e->keys[1], e->keys[0], total, nb, e->spl,
                                                         e->bcount, e->bfill, e->sc);
            });
            launch(e, cap, "bucket_count", [&] {
                k_bucket_count<<<sort_tiles, kSortThreads, 0, st>>>(e->keys[0], total, e->spl, nb,
                                                                    e->bkt, e->bcount, e->keys[1],
                                                                    e->quant, e->sc);
            });
        }
        launch(e, cap, "bucket_scatter", [&] {
            k_bucket_scatter<<<sort_tiles, kSortThreads, 0, st>>>(
                e->keys[0], e->vals[0], e->bkt, total, e->bcount, e->bfill, e->bbase, nb,
                e->keys[1], e->vals[1]);
        });
    }
    if (bp.buckets) {
        const uint32_t nb = bp.buckets;
        launch(e, cap, "bucket_sort", [&] {
            k_bucket_sort<DIM><<<nb, kBucketThreads, 0, st>>>(
                e->keys[1], e->vals[1], e->keys[0], e->vals[0], e->bcount, e->bfill, e->bbase,
                e->bucket_cap, desc, total, e->deep, bp.nseg, bp.segbits, bp.lmax, e->quant, e->sc);
        });
    } else {
        for (int p = 0; p < bp.passes; ++p) {
            const int a = p & 1, b = a ^ 1;
            if (p > 0)
                launch(e, cap, "sort_count", [&] {
                    k_sort_count<<<sort_tiles, kSortThreads, 0, st>>>(e->keys[a], total, 8 * p,
                                                                      e->sort_cnt, sort_tiles);
                });
            launch(e, cap, "sort_scan", [&] {
                k_sort_scan<<<256, kScanThreads, 0, st>>>(e->sort_cnt, sort_tiles, e->sort_tot);
            });
            launch(e, cap, "sort_scatter", [&] {
                k_sort_scatter<<<sort_tiles, kSortThreads, 0, st>>>(e->keys[a], e->vals[a], e->keys[b],
                                                              e->vals[b], total, 8 * p, e->sort_cnt,
                                                              sort_tiles, e->sort_tot, e->err_global);
            });
        }
    }
    const int sorted = bp.buckets ? 1 : bp.passes & 1;
    const uint64_t* skeys = e->keys[sorted];
    uint32_t* svals = e->vals[sorted];
    if (!bp.buckets)
        launch(e, cap, "ties", [&] {
            k_ties<DIM><<<grid_for(total, 1u << 30), kBlock, 0, st>>>(
                desc, total, skeys, svals, e->deep, bp.nseg, bp.segbits, bp.lmax);
        });
    const unsigned tiles = bp.ntiles;
    launch(e, cap, "scan_gather", [&] {
        k_scan_gather<DIM><<<tiles, kBlock, 0, st>>>(
            desc, e->tile_start, e->ext_partial, bp.nchunks, skeys, svals, bp.nseg, bp.segbits, bp.lmax, e->node_local,
            e->tile_tot, e->lo_level, e->lvl_first, e->deep, e->err_global, e->leaf_lv, e->leafpos);
    });
    launch(e, cap, "tree_scan", [&] {
        k_tree_scan<<<1 + kTopSlots, kScanThreads, 0, st>>>(e->tile_tot, e->tile_base, tiles,
                                                             e->lvl_first, e->first_node, total);
    });
    launch(e, cap, "nodes", [&] {
        k_nodes<DIM><<<tiles, kNodeThreads, 0, st>>>(
            desc, e->seg_off, e->tile_start, total, skeys, bp.nseg, bp.segbits, bp.lmax,
            e->node_local, e->tile_base, e->lvl_first, e->first_node, e->lo_level, e->hot,
            e->cap_nodes, e->top_slot, e->top_last, e->out_part, e->in_part, e->tot_part, e->deep,
            e->sc, e->leafpos, e->node_span);
    });
    launch(e, cap, "com_span", [&] {
        k_com_span<<<dim3(tiles, kComSpanGroups), kBlock, 0, st>>>(
            desc, e->seg_off, bp.nseg, e->tile_start, e->hot, e->top_last, e->top_slot,
            e->out_part, e->in_part, e->tot_part, e->sc);
    });
    launch(e, cap, "walk", [&] {
#define PG_WALK(NEAR, COUNT)                                                                   \
    k_walk<DIM, FP64, INTEG, NEAR, COUNT><<<grid_for(total, 1u << 30), kBlock, 0, st>>>(       \
        desc, total, skeys, bp.segbits, e->first_node, e->hot, e->cap_nodes, bp.nseg,         \
        e->acc_out, e->ext_partial, e->sc, bp.sort_tag, e->leaf_lv, e->walk_run_log2,          \
        e->leafpos, e->node_span, svals)
        // (walks without the near-field test or the interaction counters
        // where no segment asks for them hold fewer registers)
        if (bp.near) {
            if (bp.count) PG_WALK(true, true);
            else PG_WALK(true, false);
        } else {
            if (bp.count) PG_WALK(false, true);
            else PG_WALK(false, false);
        }
#undef PG_WALK
    });
}

template <bool FP64, bool INTEG>
static void enqueue_direct_step(pg_engine* e, const BatchPlan& bp, int q, bool cap) {
    hipStream_t st = e->stream;
    SegDesc* desc = e->desc[q];
    launch(e, cap, "prep", [&] {
        k_prep<<<1, kBlock, 0, st>>>(desc, bp.nseg, e->sc, INTEG ? 1 : 0);
    });
    // (one timed bracket: the sum and, where sources are split, the pass
    // that adds the partial sums and stores)
    launch(e, cap, "direct", [&] {
        double* part = reinterpret_cast<double*>(e->hot);
        k_direct<FP64, INTEG><<<bp.nblocks, kBlock, 0, st>>>(desc, e->blk_seg, e->blk_tile,
                                                            part, e->acc_out);
        if (bp.max_split > 1)
            k_direct_finish<FP64, INTEG><<<grid_for(bp.total, 1u << 30), kBlock, 0, st>>>(
                desc, e->seg_off, bp.nseg, bp.total, part, e->acc_out);
    });
}

static void enqueue_step(pg_engine* e, const BatchPlan& bp, int q, bool cap, bool extent,
                         bool fused) {
    const bool f64 = bp.precision == PG_PREC_FP64;
    if (bp.kind == PG_KIND_DIRECT) {
        if (bp.integrate) {
            if (f64) enqueue_direct_step<true, true>(e, bp, q, cap);
            else enqueue_direct_step<false, true>(e, bp, q, cap);
        } else {
            if (f64) enqueue_direct_step<true, false>(e, bp, q, cap);
            else enqueue_direct_step<false, false>(e, bp, q, cap);
        }
        return;
    }
    if (bp.dim == 3) {
        if (bp.integrate) {
            if (f64) enqueue_tree_step<3, true, true>(e, bp, q, cap, extent, fused);
            else enqueue_tree_step<3, false, true>(e, bp, q, cap, extent, fused);
        } else {
            if (f64) enqueue_tree_step<3, true, false>(e, bp, q, cap, extent, fused);
            else enqueue_tree_step<3, false, false>(e, bp, q, cap, extent, fused);
        }
    } else {
        if (bp.integrate) {
            if (f64) enqueue_tree_step<2, true, true>(e, bp, q, cap, extent, fused);
            else enqueue_tree_step<2, false, true>(e, bp, q, cap, extent, fused);
        } else {
            if (f64) enqueue_tree_step<2, true, false>(e, bp, q, cap, extent, fused);
            else enqueue_tree_step<2, false, false>(e, bp, q, cap, extent, fused);
        }
    }
}

static uint32_t map_err(uint32_t bits) {
    if (bits & kErrNonFinite) return PG_ERR_NONFINITE;
    if (bits & kErrCoincident) return PG_ERR_COINCIDENT;
    if (bits & kErrDepth) return PG_ERR_DEPTH;
    if (bits & kErrCapacity) return PG_ERR_INTERNAL;
    return PG_OK;
}

static int err_text(uint32_t bits) {
    const uint32_t code = map_err(bits);
    switch (code) {
        case PG_ERR_NONFINITE:
            return pg_set_error(PG_ERR_NONFINITE, "non-finite body coordinate");
        case PG_ERR_COINCIDENT:
            return pg_set_error(PG_ERR_COINCIDENT,
                                "bodies within 1e-9 on every axis (the reference merges them; "
                                "not supported on the device)");
        case PG_ERR_DEPTH:
            return pg_set_error(PG_ERR_DEPTH,
                                "two bodies share every level the Morton key holds");
        case PG_ERR_INTERNAL:
            return pg_set_error(PG_ERR_INTERNAL,
                                "device-side capacity exceeded (bits 0x%x)", bits);
        default:
            return PG_OK;
    }
}

// Core of every device computation: `steps` steps of all sims together
// (integrate) or one force evaluation into e->acc_out (no integrate).
static int run_batch(pg_engine* e, pg_sim* const* sims_in, uint32_t count, uint32_t steps,
                     bool integrate) {
    std::vector<pg_sim*>& sims = e->h_sims;
    sims.clear();
    for (uint32_t i = 0; i < count; ++i) {
        pg_sim* s = sims_in[i];
        if (!s) return pg_set_error(PG_ERR_ARG, "sim %u is NULL", i);
        if (s->eng != e) return pg_set_error(PG_ERR_ARG, "sim %u belongs to another engine", i);
        if (!s->uploaded) return pg_set_error(PG_ERR_STATE, "sim %u has no uploaded state", i);
        if (s->n) sims.push_back(s);
    }
    if (sims.empty() || steps == 0) return PG_OK;
    BatchPlan bp;
    bp.kind = sims[0]->p.kind;
    bp.precision = sims[0]->p.precision;
    for (pg_sim* s : sims)
        if (s->p.kind != bp.kind || s->p.precision != bp.precision)
            return pg_set_error(PG_ERR_ARG, "a batch must share kind and precision");
    bp.integrate = integrate;
    bp.nseg = (uint32_t)sims.size();
    bp.dim = key_dim(bp.kind);
    bp.segbits = bp.nseg > 1 ? ceil_log2(bp.nseg) : 0;
    bp.lmax = key_lmax(bp.dim, bp.segbits);
    bp.passes = (bp.segbits + bp.dim * bp.lmax + 7) / 8;
    for (pg_sim* s : sims) {
        bp.total += s->n;
        bp.near = bp.near || s->near > 0.0;
        bp.count = bp.count || s->counting;
    }
    if (bp.total >= (1ull << 32))
        return pg_set_error(PG_ERR_ARG, "a batch holds at most 2^32-1 bodies");
    {
        // a batch past max_buckets of bucket_keys grows its buckets up to
        // the slot sort's reach (kSlots keys) before it takes the radix sort
        uint64_t keys_per = e->bucket_keys;
        if (e->max_buckets && (bp.total + keys_per - 1) / keys_per > e->max_buckets)
            keys_per = std::max<uint64_t>(keys_per, (bp.total + e->max_buckets - 1) / e->max_buckets);
        const uint64_t nb = (bp.total + keys_per - 1) / keys_per;
        bp.buckets = nb >= 1 && nb <= e->max_buckets && keys_per <= kSlots ? (uint32_t)nb : 0u;
    }
    uint64_t nodes = bp.kind == PG_KIND_DIRECT ? 0 : nodes_bound(bp.total, bp.dim, bp.lmax) + bp.nseg;
    if (nodes >= (1ull << 32)) return pg_set_error(PG_ERR_ARG, "tree too large for 32-bit nodes");
    e->h_blk_seg.clear();
    e->h_blk_tile.clear();
    std::vector<uint32_t>& nsplit = e->h_nsplit;
    std::vector<uint64_t>& part_off = e->h_part_off;
    nsplit.assign(bp.nseg, 1u);
    part_off.assign(bp.nseg, 0ull);
    if (bp.kind == PG_KIND_DIRECT) {
        // Blocks: per segment, target tile and source split (split in the
        // tile word's top 8 bits). A segment's partial sums (3 doubles per
        // body and split) live in the node array, which the direct sum does
        // not otherwise use.
        uint64_t part = 0;
        for (uint32_t s = 0; s < bp.nseg; ++s) {
            const uint32_t nb = (uint32_t)((sims[s]->n + kBlock - 1) / kBlock);
            const uint32_t S = direct_splits(sims[s]->n);
            for (uint32_t t = 0; t < nb; ++t)
                for (uint32_t p = 0; p < S; ++p) {
                    e->h_blk_seg.push_back(s);
                    e->h_blk_tile.push_back(t | (p << 24));
                }
            nsplit[s] = S;
            if (S > 1) {
                part_off[s] = part;
                part += (uint64_t)S * sims[s]->n;
                bp.max_split = std::max(bp.max_split, S);
            }
        }
        bp.nblocks = (uint32_t)e->h_blk_seg.size();
        nodes = (part * 24 + sizeof(NodeHot) - 1) / sizeof(NodeHot);
        if (nodes >= (1ull << 32)) return pg_set_error(PG_ERR_ARG, "direct-sum partials too large");
    }
    e->h_ext_seg.clear();
    e->h_ext_chunk.clear();
    std::vector<uint32_t>& chunk0 = e->h_chunk0;
    std::vector<uint32_t>& nchunk = e->h_nchunk;
    chunk0.assign(bp.nseg, 0u);
    nchunk.assign(bp.nseg, 0u);
    for (uint32_t s = 0; s < bp.nseg; ++s) {
        const uint32_t nc = (uint32_t)((sims[s]->n + kExtentChunk - 1) / kExtentChunk);
        chunk0[s] = (uint32_t)e->h_ext_seg.size();
        nchunk[s] = nc;
        for (uint32_t c = 0; c < nc; ++c) {
            e->h_ext_seg.push_back(s);
            e->h_ext_chunk.push_back(c);
        }
    }
    bp.nchunks = (uint32_t)e->h_ext_seg.size();
    // node tiles of kNodeTile bodies from each segment's first body, so that a
    // segment's tree arithmetic does not depend on where the batch puts it
    e->h_tile_start.clear();
    std::vector<uint32_t>& tile0 = e->h_tile0;
    tile0.assign(bp.nseg, 0u);
    {
        uint64_t o = 0;
        for (uint32_t s = 0; s < bp.nseg; ++s) {
            tile0[s] = (uint32_t)e->h_tile_start.size();
            for (uint64_t t = 0; t < sims[s]->n; t += kNodeTile)
                e->h_tile_start.push_back((uint32_t)(o + t));
            o += sims[s]->n;
        }
        bp.ntiles = (uint32_t)e->h_tile_start.size();
        e->h_tile_start.push_back((uint32_t)o);
    }
    // A fixed scratch holds fewer nodes than the worst-case bound; the node
    // passes check every node index against the capacity and report a tree
    // that would exceed it (PG_ERR_INTERNAL) instead of writing past it.
    if (e->fixed_scratch && bp.kind != PG_KIND_DIRECT) nodes = std::min(nodes, e->cap_nodes);
    PG_TRY(ensure_scratch(e, bp.total, nodes, bp.nseg, bp.nblocks));
    if (bp.nchunks > e->cap_chunks)
        return pg_set_error(PG_ERR_INTERNAL, "extent chunk table overflow");
    if (bp.ntiles > e->cap_tiles)
        return pg_set_error(PG_ERR_INTERNAL, "node tile table overflow");

    // descriptors for both parities
    e->h_seg_off.assign(bp.nseg + 1, 0);
    for (int q = 0; q < 2; ++q) e->h_desc[q].assign(bp.nseg, SegDesc{});
    uint64_t off = 0;
    for (uint32_t s = 0; s < bp.nseg; ++s) {
        pg_sim* sm = sims[s];
        e->h_seg_off[s] = off;
        for (int q = 0; q < 2; ++q) {
            const int src = sm->cur ^ q;
            const int dst = bp.kind == PG_KIND_DIRECT && !integrate ? src : src ^ 1;
            SegDesc& d = e->h_desc[q][s];
            d.spos = sm->pos[src]; d.sprv = sm->prv[src];
            d.svx = sm->vx[src]; d.svy = sm->vy[src]; d.svz = sm->vz[src];
            d.dpos = sm->pos[dst]; d.dprv = sm->prv[dst];
            d.dvx = sm->vx[dst]; d.dvy = sm->vy[dst]; d.dvz = sm->vz[dst];
            d.off = off;
            d.n = sm->n;
            d.theta = sm->p.theta;
            d.e = std::fabs(sm->p.e);
            d.dt = sm->p.dt;
            d.seg = s;
            d.count = sm->counting ? 1u : 0u;
            d.counts = sm->counts;
            d.st = sm->st;
            d.chunk0 = chunk0[s];
            d.nchunks = nchunk[s];
            d.tile0 = tile0[s];
            d.nsplit = nsplit[s];
            d.part_off = part_off[s];
            d.near = sm->near;
        }
        off += sm->n;
    }
    e->h_seg_off[bp.nseg] = off;
    HIP_TRY(hipMemcpyAsync(e->desc[0], e->h_desc[0].data(), bp.nseg * sizeof(SegDesc),
                           hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(e->desc[1], e->h_desc[1].data(), bp.nseg * sizeof(SegDesc),
                           hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(e->seg_off, e->h_seg_off.data(), (bp.nseg + 1) * 8,
                           hipMemcpyHostToDevice, e->stream));
    if (bp.nblocks) {
        HIP_TRY(hipMemcpyAsync(e->blk_seg, e->h_blk_seg.data(), bp.nblocks * 4,
                               hipMemcpyHostToDevice, e->stream));
        HIP_TRY(hipMemcpyAsync(e->blk_tile, e->h_blk_tile.data(), bp.nblocks * 4,
                               hipMemcpyHostToDevice, e->stream));
    }
    HIP_TRY(hipMemcpyAsync(e->ext_seg, e->h_ext_seg.data(), bp.nchunks * 4,
                           hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(e->ext_chunk, e->h_ext_chunk.data(), bp.nchunks * 4,
                           hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(e->tile_start, e->h_tile_start.data(), (bp.ntiles + 1) * 4,
                           hipMemcpyHostToDevice, e->stream));

    // the batch's identity for the sample sort: its splitters may come from
    // the keys its previous step sorted only when the same batch ran last
    {
        uint64_t h = 1469598103934665603ull;
        auto mix = [&](uint64_t v) {
            h ^= v;
            h *= 1099511628211ull;
        };
        mix(bp.total);
        mix(bp.nseg);
        for (pg_sim* s : sims) {
            mix((uint64_t)(uintptr_t)s);
            mix(s->n);
        }
        bp.sort_tag = h ? h : 1;
    }
    const bool flips = bp.kind != PG_KIND_DIRECT || integrate;
    uint32_t done = 0;
    // An integrating tree step's walk leaves the next step's extent partials:
    // the extent pass runs only where the last step did not (after uploads,
    // other batches, force-only or direct steps). Graphs hold steps that
    // follow such a step; a call that starts without one runs its first two
    // steps eagerly (keeping the graph's parities).
    // Likewise the sample sort takes its splitters from the batch's previous
    // sorted keys where the last step left them (fused into k_keys).
    const bool tree = bp.kind != PG_KIND_DIRECT;
    bool ext_ok = tree && e->ext_valid != 0 && e->ext_valid == bp.sort_tag;
    bool srt_ok = tree && bp.buckets && e->sorted_valid != 0 && e->sorted_valid == bp.sort_tag;
    const bool warm_needs_sort = tree && bp.buckets;
    auto warm = [&] { return !tree || (ext_ok && (srt_ok || !warm_needs_sort)); };
    if (integrate && e->graphs && !e->profiling && steps >= 4 && !warm()) {
        enqueue_step(e, bp, 0, false, !ext_ok, srt_ok);
        enqueue_step(e, bp, 1, false, false, warm_needs_sort);
        done = 2;
        ext_ok = true;
        srt_ok = warm_needs_sort;
    }
    if (integrate && e->graphs && !e->profiling && steps >= 2 && warm()) {
        // signature: sims, parities, sizes, flags and the scratch it was captured on
        std::vector<uint64_t> sig;
        sig.push_back((uint64_t)(uintptr_t)e->scratch);
        sig.push_back(((uint64_t)bp.kind * 16 + bp.precision) * 2 + (bp.near ? 1 : 0));
        for (pg_sim* s : sims) {
            sig.push_back((uint64_t)(uintptr_t)s);
            sig.push_back(s->n * 4 + (uint64_t)s->cur * 2 + (s->counting ? 1 : 0));
        }
        hipGraphExec_t exec = nullptr;
        for (auto& g : e->graph_cache)
            if (g.sig == sig) exec = g.exec;
        if (!exec) {
            hipGraph_t graph = nullptr;
            HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
            enqueue_step(e, bp, 0, true, false, warm_needs_sort);
            enqueue_step(e, bp, 1, true, false, warm_needs_sort);
            HIP_TRY(hipStreamEndCapture(e->stream, &graph));
            HIP_TRY(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            (void)hipGraphDestroy(graph);
            if (e->graph_cache.size() >= 4) {
                (void)hipGraphExecDestroy(e->graph_cache.front().exec);
                e->graph_cache.erase(e->graph_cache.begin());
            }
            e->graph_cache.push_back({sig, exec});
        }
        for (; done + 2 <= steps; done += 2) HIP_TRY(hipGraphLaunch(exec, e->stream));
    }
    for (; done < steps; ++done) {
        enqueue_step(e, bp, (int)(done & 1), false, !ext_ok, srt_ok);
        ext_ok = true;
        srt_ok = warm_needs_sort;
        if (!integrate) break;
    }
    e->ext_valid = integrate && tree ? bp.sort_tag : 0;
    e->sorted_valid = warm_needs_sort ? bp.sort_tag : 0;
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (e->profiling) ev_collect(e);

    if (flips && (steps & 1))
        for (pg_sim* s : sims) s->cur ^= 1;
    if (integrate)
        for (pg_sim* s : sims) s->dt_last = s->p.dt;

    e->last_sim = bp.nseg == 1 && bp.kind != PG_KIND_DIRECT ? sims[0] : nullptr;
    e->last_total = bp.total;
    e->last_sorted = bp.buckets ? 1 : bp.passes & 1;
    e->last_dim = bp.dim;
    e->last_lmax = bp.lmax;

    uint32_t gerr = 0;
    HIP_TRY(hipMemcpy(&gerr, e->err_global, 4, hipMemcpyDeviceToHost));
    if (gerr) {
        uint32_t z = 0;
        HIP_TRY(hipMemcpy(e->err_global, &z, 4, hipMemcpyHostToDevice));
        return err_text(gerr);
    }
    int rc = PG_OK;
    for (pg_sim* s : sims) {
        SegState st;
        HIP_TRY(hipMemcpy(&st, s->st, sizeof st, hipMemcpyDeviceToHost));
        if (st.err && rc == PG_OK) rc = err_text(st.err);
    }
    return rc;
}

// ------------------------------------------------------------- simulations
static int validate_params(const pg_params* p) {
    if (!p) return pg_set_error(PG_ERR_ARG, "params is NULL");
    if (p->kind != PG_KIND_DIRECT && p->kind != PG_KIND_QUADTREE && p->kind != PG_KIND_OCTREE)
        return pg_set_error(PG_ERR_ARG, "unknown kind %d", p->kind);
    if (p->precision != PG_PREC_MIXED && p->precision != PG_PREC_FP64)
        return pg_set_error(PG_ERR_ARG, "unknown precision %d", p->precision);
    return PG_OK;
}

// A simulation's device arrays for n bodies, carved from one block of
// state_bytes(n) at b.
static void carve_state(pg_sim* s, char* b, uint64_t n) {
    uint64_t off = 0;
    auto take = [&](uint64_t sz) {
        char* q = b + off;
        off += align256(sz);
        return q;
    };
    for (int k = 0; k < 2; ++k) {
        // (contiguous: a copy holds an Entity array of n bodies when staged)
        s->pos[k] = (double4*)take(n * 32);
        s->prv[k] = (double4*)take(n * 32);
        s->vx[k] = (double*)take(n * 8); s->vy[k] = (double*)take(n * 8); s->vz[k] = (double*)take(n * 8);
    }
    s->aos = (Entity*)take(n * 80);
    s->counts = (int64_t*)take(n * 8);
    s->st = (SegState*)take(sizeof(SegState));
}

extern "C" int pg_sim_create(pg_engine* e, uint64_t n, const pg_params* p, pg_sim** out) {
    if (!e || !out) return pg_set_error(PG_ERR_ARG, "engine or out is NULL");
    PG_TRY(validate_params(p));
    if (n >= (1ull << 31)) return pg_set_error(PG_ERR_ARG, "n too large");
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    HIP_TRY(hipSetDevice(e->device));
    const uint64_t bytes = state_bytes(n);
    if (e->sims_bytes + e->scratch_bytes + bytes > e->budget)
        return pg_set_error(PG_ERR_BUDGET, "simulation of %llu bodies exceeds the HBM budget",
                            (unsigned long long)n);
    pg_sim* s = new pg_sim();
    s->eng = e;
    s->n = n;
    s->p = *p;
    hipError_t he = hipMalloc(&s->block, bytes);
    if (he != hipSuccess) {
        delete s;
        return pg_set_error(PG_ERR_NOMEM, "hipMalloc(%llu): %s", (unsigned long long)bytes,
                            hipGetErrorString(he));
    }
    s->bytes = bytes;
    carve_state(s, (char*)s->block, n);
    HIP_TRY(hipMemset(s->st, 0, sizeof(SegState)));
    e->sims_bytes += bytes;
    e->sims.push_back(s);
    *out = s;
    return PG_OK;
}

extern "C" void pg_sim_destroy(pg_sim* s) {
    if (!s) return;
    pg_engine* e = s->eng;
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    (void)hipSetDevice(e->device);
    (void)hipStreamSynchronize(e->stream);
    // graphs captured on this sim's buffers must not be replayed
    for (auto it = e->graph_cache.begin(); it != e->graph_cache.end();) {
        bool uses = false;
        for (uint64_t w : it->sig)
            if (w == (uint64_t)(uintptr_t)s) uses = true;
        if (uses) {
            (void)hipGraphExecDestroy(it->exec);
            it = e->graph_cache.erase(it);
        } else {
            ++it;
        }
    }
    if (e->last_sim == s) e->last_sim = nullptr;
    if (e->xform == s) e->xform = nullptr;
    e->sims.erase(std::remove(e->sims.begin(), e->sims.end(), s), e->sims.end());
    if (s->stage) (void)hipFree(s->stage);
    (void)hipFree(s->block);
    e->sims_bytes -= s->bytes;
    delete s;
}

extern "C" int pg_sim_set_params(pg_sim* s, const pg_params* p) {
    if (!s) return pg_set_error(PG_ERR_ARG, "sim is NULL");
    PG_TRY(validate_params(p));
    s->p = *p;
    return PG_OK;
}

extern "C" int pg_sim_set_near_field(pg_sim* s, double leaf_widths) {
    if (!s) return pg_set_error(PG_ERR_ARG, "sim is NULL");
    if (!(leaf_widths >= 0.0) || leaf_widths > 1e6)
        return pg_set_error(PG_ERR_ARG, "near field of %g leaf widths", leaf_widths);
    s->near = leaf_widths;
    return PG_OK;
}

extern "C" int pg_sim_set_counting(pg_sim* s, int enable) {
    if (!s) return pg_set_error(PG_ERR_ARG, "sim is NULL");
    s->counting = enable != 0;
    return PG_OK;
}

extern "C" int pg_sim_upload(pg_sim* s, const Entity* state, uint64_t n) {
    if (!s) return pg_set_error(PG_ERR_ARG, "sim is NULL");
    if (n != s->n) return pg_set_error(PG_ERR_ARG, "n %llu != sim n %llu", (unsigned long long)n,
                                       (unsigned long long)s->n);
    if (n && !state) return pg_set_error(PG_ERR_ARG, "state is NULL");
    pg_engine* e = s->eng;
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    HIP_TRY(hipSetDevice(e->device));
    s->cur = 0;
    if (n) {
        HIP_TRY(hipMemcpyAsync(s->aos, state, n * sizeof(Entity), hipMemcpyHostToDevice, e->stream));
        k_unpack<<<grid_for(n, 1u << 30), kBlock, 0, e->stream>>>(
            s->aos, n, s->pos[0], s->prv[0], s->vx[0], s->vy[0], s->vz[0]);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipMemsetAsync(s->st, 0, sizeof(SegState), e->stream));
    s->stage_steps = 0;
    // new positions: no extent partials are current
    if (e->sc) HIP_TRY(hipMemsetAsync(&e->sc->ext_tag, 0, sizeof(uint64_t), e->stream));
    e->ext_valid = 0;
    HIP_TRY(hipStreamSynchronize(e->stream));
    s->uploaded = true;
    if (e->last_sim == s) e->last_sim = nullptr;
    return PG_OK;
}

extern "C" int pg_sim_upload_state(pg_sim* s, const Entity* state, const Entity* previous,
                                   uint64_t n) {
    PG_TRY(pg_sim_upload(s, state, n));
    if (!previous || !n) return PG_OK;
    pg_engine* e = s->eng;
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    HIP_TRY(hipSetDevice(e->device));
    // the previous input is staged in the other state buffer (its records
    // and velocities hold n * 88 contiguous bytes; upload made buffer 0
    // current), its positions become the history of buffer 0
    Entity* stage = (Entity*)s->pos[1];
    HIP_TRY(hipMemcpyAsync(stage, previous, n * sizeof(Entity), hipMemcpyHostToDevice, e->stream));
    k_unpack_prev<<<grid_for(n, 1u << 30), kBlock, 0, e->stream>>>(stage, n, s->prv[0]);
    HIP_TRY(hipGetLastError());
    // one step taken: the next is a Stormer step (verlet.rs:104-106)
    SegState st;
    memset(&st, 0, sizeof st);
    st.steps_done = 1;
    HIP_TRY(hipMemcpyAsync(s->st, &st, sizeof st, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return PG_OK;
}

// Where the current state's velocities come from: the stored arrays up to and
// right after the Taylor first step, the backward difference afterwards.
static int velocity_source(pg_sim* s, VelocitySource* vs) {
    SegState st;
    HIP_TRY(hipMemcpy(&st, s->st, sizeof st, hipMemcpyDeviceToHost));
    const int c = s->cur;
    vs->vx = s->vx[c]; vs->vy = s->vy[c]; vs->vz = s->vz[c];
    vs->prv = s->prv[c];
    vs->derive = s->integ == PG_INTEG_VERLET && st.steps_done >= 2 ? 1 : 0;
    vs->dt = s->dt_last;
    return PG_OK;
}

extern "C" int pg_sim_download(pg_sim* s, Entity* out, uint64_t n) {
    if (!s) return pg_set_error(PG_ERR_ARG, "sim is NULL");
    if (n != s->n) return pg_set_error(PG_ERR_ARG, "n mismatch");
    if (!s->uploaded) return pg_set_error(PG_ERR_STATE, "no state uploaded");
    if (!n) return PG_OK;
    if (!out) return pg_set_error(PG_ERR_ARG, "out is NULL");
    pg_engine* e = s->eng;
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    HIP_TRY(hipSetDevice(e->device));
    const int c = s->cur;
    VelocitySource vs;
    PG_TRY(velocity_source(s, &vs));
    k_pack<<<grid_for(n, 1u << 30), kBlock, 0, e->stream>>>(s->aos, n, s->pos[c], vs, s->prv[c]);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, s->aos, n * sizeof(Entity), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return PG_OK;
}

static StageState stage_state(pg_sim* s) {
    const int c = s->cur;
    return StageState{s->pos[c], s->prv[c], s->vx[c], s->vy[c], s->vz[c], s->n};
}

// euler.rs:17-49 / rk4.rs:24-182 on the device: each acceleration pass is
// the force-only tree or direct pass over the current (stage) state, as the
// reference's acc_fn runs the transform on it; the stage kernels form the
// updates.
static int step_stages(pg_engine* e, pg_sim* s, uint32_t steps) {
    if (!s->uploaded) return pg_set_error(PG_ERR_STATE, "no state uploaded");
    if (!s->n) return PG_OK;
    const unsigned g = grid_for(s->n, 1u << 30);
    double* base = s->stage;
    double* ksum = s->stage + 6 * s->n;
    for (uint32_t k = 0; k < steps; ++k) {
        if (s->integ == PG_INTEG_EULER) {
            PG_TRY(run_batch(e, &s, 1, 1, false));
            k_euler_update<<<g, kBlock, 0, e->stream>>>(stage_state(s), e->acc_out, s->p.dt);
            HIP_TRY(hipGetLastError());
        } else {
            k_rk4_begin<<<g, kBlock, 0, e->stream>>>(stage_state(s), base);
            HIP_TRY(hipGetLastError());
            for (int st = 0; st < 4; ++st) {
                PG_TRY(run_batch(e, &s, 1, 1, false));
                k_rk4_stage<<<g, kBlock, 0, e->stream>>>(stage_state(s), e->acc_out, base, ksum,
                                                         s->p.dt, st);
                HIP_TRY(hipGetLastError());
            }
        }
        s->stage_steps += 1;
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->ext_valid = 0;  // the stage kernels moved the bodies
    return PG_OK;
}

extern "C" int pg_sim_checkpoint(pg_sim* s, Entity* state, Entity* previous, uint64_t n,
                                 int* has_history) {
    if (!s || !state) return pg_set_error(PG_ERR_ARG, "sim or state is NULL");
    if (s->integ != PG_INTEG_VERLET)
        return pg_set_error(PG_ERR_STATE, "checkpoints hold the verlet history (euler and rk4 "
                                          "resume from the state alone: pg_sim_download)");
    PG_TRY(pg_sim_download(s, state, n));
    pg_engine* e = s->eng;
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    HIP_TRY(hipSetDevice(e->device));
    SegState st;
    HIP_TRY(hipMemcpy(&st, s->st, sizeof st, hipMemcpyDeviceToHost));
    const bool hist = st.steps_done > 0;
    if (has_history) *has_history = hist ? 1 : 0;
    if (!previous || !n) return PG_OK;
    if (!hist) {
        memcpy(previous, state, n * sizeof(Entity));
        return PG_OK;
    }
    // the previous positions, staged in the other state buffer (dead
    // between steps; its records and velocities hold n * 88 bytes)
    const int c = s->cur;
    Entity* stage = (Entity*)s->pos[c ^ 1];
    k_pack_prev<<<grid_for(n, 1u << 30), kBlock, 0, e->stream>>>(stage, s->aos, n, s->prv[c]);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(previous, stage, n * sizeof(Entity), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return PG_OK;
}

extern "C" int pg_sim_step(pg_sim* s, uint32_t steps) {
    if (!s) return pg_set_error(PG_ERR_ARG, "sim is NULL");
    pg_engine* e = s->eng;
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    HIP_TRY(hipSetDevice(e->device));
    if (s->integ != PG_INTEG_VERLET) return step_stages(e, s, steps);
    return run_batch(e, &s, 1, steps, true);
}

extern "C" int pg_sim_set_integrator(pg_sim* s, int32_t integrator) {
    if (!s) return pg_set_error(PG_ERR_ARG, "sim is NULL");
    if (integrator != PG_INTEG_VERLET && integrator != PG_INTEG_EULER &&
        integrator != PG_INTEG_RK4)
        return pg_set_error(PG_ERR_ARG, "unknown integrator %d", integrator);
    pg_engine* e = s->eng;
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    HIP_TRY(hipSetDevice(e->device));
    if (integrator == s->integ) return PG_OK;
    if (integrator != PG_INTEG_VERLET && !s->stage && s->n) {
        hipError_t he = hipMalloc(&s->stage, s->n * 12 * sizeof(double));
        if (he != hipSuccess) {
            s->stage = nullptr;
            return pg_set_error(PG_ERR_NOMEM, "hipMalloc for the integrator stages: %s",
                                hipGetErrorString(he));
        }
    }
    if (s->uploaded && s->n) {
        // the current velocities into the stored arrays (the Stormer ones are
        // derived), then a fresh start: a new verlet element begins with its
        // Taylor step (verlet.rs:102-103)
        VelocitySource vs;
        PG_TRY(velocity_source(s, &vs));
        if (vs.derive)
            k_store_velocity<<<grid_for(s->n, 1u << 30), kBlock, 0, e->stream>>>(stage_state(s),
                                                                                 vs.dt);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemsetAsync(s->st, 0, sizeof(SegState), e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
    }
    s->integ = integrator;
    return PG_OK;
}

extern "C" int pg_batch_step(pg_engine* e, pg_sim* const* sims, uint32_t count, uint32_t steps) {
    if (!e || (count && !sims)) return pg_set_error(PG_ERR_ARG, "engine or sims is NULL");
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    HIP_TRY(hipSetDevice(e->device));
    for (uint32_t i = 0; i < count; ++i)
        if (sims[i] && sims[i]->integ != PG_INTEG_VERLET)
            return pg_set_error(PG_ERR_ARG, "batched steps take verlet simulations (sim %u)", i);
    return run_batch(e, sims, count, steps, true);
}

// The autotuner. Every candidate tuning leaves the results bit-identical
// (the sort's order, tree and arithmetic do not depend on them); only the
// time changes. Each candidate runs the batch from the same saved state: four
// warm-up steps (first-step passes, graph capture), then `trial_steps` timed
// steps; the fastest is kept, and the simulations get their saved state back.
extern "C" int pg_engine_autotune(pg_engine* e, pg_sim* const* sims, uint32_t count,
                                  uint32_t trial_steps, pg_tuning* best, pg_tuning* tried,
                                  double* ms_per_step, uint32_t cap, uint32_t* ntried) {
    if (!e || !sims || !count) return pg_set_error(PG_ERR_ARG, "engine or sims is NULL");
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    HIP_TRY(hipSetDevice(e->device));
    uint64_t total = 0;
    for (uint32_t i = 0; i < count; ++i) {
        if (!sims[i] || sims[i]->eng != e || !sims[i]->uploaded)
            return pg_set_error(PG_ERR_ARG, "sim %u is not an uploaded sim of this engine", i);
        if (sims[i]->integ != PG_INTEG_VERLET)
            return pg_set_error(PG_ERR_ARG, "the autotuner times verlet steps (sim %u)", i);
        total += sims[i]->n;
    }
    if (trial_steps == 0) trial_steps = 4;
    // candidates: the walk's XCD runs, and for a batch the sample sort can
    // take, its bucket size (or the radix sort)
    std::vector<pg_tuning> cand;
    const bool tree = sims[0]->p.kind != PG_KIND_DIRECT;
    std::vector<uint32_t> runs = tree ? std::vector<uint32_t>{0u, 2u, 4u, 6u}
                                      : std::vector<uint32_t>{e->walk_run_log2};
    std::vector<std::pair<uint32_t, uint32_t>> sorts;  // (max_buckets, bucket_keys)
    if (tree) {
        for (uint32_t bk : {1280u, 1664u, 2048u})
            if ((total + bk - 1) / bk <= (uint64_t)kMaxWideBuckets)
                sorts.push_back({kMaxWideBuckets, bk});
        sorts.push_back({0u, kBucketKeys});
    } else {
        sorts.push_back({e->max_buckets, e->bucket_keys});
    }
    for (auto so : sorts)
        for (uint32_t r : runs) cand.push_back(pg_tuning{so.first, so.second, r, e->lds_buckets});
    pg_tuning keep;
    PG_TRY(pg_engine_get_tuning(e, &keep));
    // the saved state: each simulation's device block and host step fields
    struct Saved {
        void* block = nullptr;
        int cur = 0;
        uint64_t stage_steps = 0;
        double dt_last = 0.0;
    };
    std::vector<Saved> saved(count);
    int rc = PG_OK;
    for (uint32_t i = 0; i < count && rc == PG_OK; ++i) {
        if (hipMalloc(&saved[i].block, sims[i]->bytes) != hipSuccess) {
            saved[i].block = nullptr;
            rc = pg_set_error(PG_ERR_NOMEM, "autotune: no memory to save sim %u", i);
            break;
        }
        if (hipMemcpyAsync(saved[i].block, sims[i]->block, sims[i]->bytes,
                           hipMemcpyDeviceToDevice, e->stream) != hipSuccess)
            rc = pg_set_error(PG_ERR_HIP, "autotune: saving sim %u failed", i);
        saved[i].cur = sims[i]->cur;
        saved[i].stage_steps = sims[i]->stage_steps;
        saved[i].dt_last = sims[i]->dt_last;
    }
    auto restore = [&]() -> int {
        for (uint32_t i = 0; i < count; ++i) {
            if (!saved[i].block) continue;
            HIP_TRY(hipMemcpyAsync(sims[i]->block, saved[i].block, sims[i]->bytes,
                                   hipMemcpyDeviceToDevice, e->stream));
            sims[i]->cur = saved[i].cur;
            sims[i]->stage_steps = saved[i].stage_steps;
            sims[i]->dt_last = saved[i].dt_last;
        }
        // the positions changed under the extent partials the last walk left:
        // the device-side tag that lets the extent pass skip goes too (as on
        // an upload)
        if (e->sc) HIP_TRY(hipMemsetAsync(&e->sc->ext_tag, 0, sizeof(uint64_t), e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
        e->ext_valid = 0;
        e->sorted_valid = 0;
        e->last_sim = nullptr;
        return PG_OK;
    };
    double best_ms = 1e300;
    pg_tuning best_t = keep;
    uint32_t k = 0;
    for (const pg_tuning& t : cand) {
        if (rc != PG_OK) break;
        rc = pg_engine_set_tuning(e, &t);
        if (rc == PG_OK) rc = restore();
        // warm-up: the first-step passes, then a captured graph
        if (rc == PG_OK) rc = run_batch(e, sims, count, 4, true);
        double ms = 0.0;
        if (rc == PG_OK) {
            const auto t0 = std::chrono::steady_clock::now();
            rc = run_batch(e, sims, count, trial_steps, true);
            ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                     .count() / trial_steps;
        }
        if (rc != PG_OK) {
            const std::string why = pg_last_error();
            pg_set_error(rc, "autotune candidate %u (max_buckets %u, bucket_keys %u, walk_run_log2 "
                         "%u): %s", k, t.max_buckets, t.bucket_keys, t.walk_run_log2, why.c_str());
            break;
        }
        if (k < cap) {
            if (tried) tried[k] = t;
            if (ms_per_step) ms_per_step[k] = ms;
        }
        ++k;
        if (ms < best_ms) {
            best_ms = ms;
            best_t = t;
        }
    }
    // the fastest tuning stays set; the simulations get their state back
    const std::string err = pg_last_error();
    int rc2 = pg_engine_set_tuning(e, rc == PG_OK ? &best_t : &keep);
    int rc3 = restore();
    if (rc != PG_OK) pg_set_error(rc, "%s", err.c_str());
    for (Saved& sv : saved)
        if (sv.block) (void)hipFree(sv.block);
    if (ntried) *ntried = k;
    if (best) *best = rc == PG_OK ? best_t : keep;
    if (rc != PG_OK) return rc;
    return rc2 != PG_OK ? rc2 : rc3;
}

extern "C" int pg_sim_accel(pg_sim* s, double* acc) {
    if (!s) return pg_set_error(PG_ERR_ARG, "sim is NULL");
    if (s->n && !acc) return pg_set_error(PG_ERR_ARG, "acc is NULL");
    pg_engine* e = s->eng;
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    HIP_TRY(hipSetDevice(e->device));
    PG_TRY(run_batch(e, &s, 1, 1, false));
    if (s->n) HIP_TRY(hipMemcpy(acc, e->acc_out, s->n * 24, hipMemcpyDeviceToHost));
    return PG_OK;
}

extern "C" int pg_transform_aos(pg_engine* e, const pg_params* p, const Entity* state, uint64_t n,
                                Acceleration* acc) {
    if (!e) return pg_set_error(PG_ERR_ARG, "engine is NULL");
    PG_TRY(validate_params(p));
    if (n == 0) return PG_OK;
    if (!state || !acc) return pg_set_error(PG_ERR_ARG, "state or acc is NULL");
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    HIP_TRY(hipSetDevice(e->device));
    if (e->xform && e->xform->n != n) {
        pg_sim_destroy(e->xform);
        e->xform = nullptr;
    }
    if (!e->xform) PG_TRY(pg_sim_create(e, n, p, &e->xform));
    PG_TRY(pg_sim_set_params(e->xform, p));
    PG_TRY(pg_sim_upload(e->xform, state, n));
    std::vector<double> a(n * 3);
    PG_TRY(pg_sim_accel(e->xform, a.data()));
    for (uint64_t i = 0; i < n; ++i) {
        acc[i].x += a[3 * i + 0];
        acc[i].y += a[3 * i + 1];
        acc[i].z += a[3 * i + 2];
    }
    return PG_OK;
}

extern "C" int pg_sim_energy(pg_sim* s, double out[6]) {
    if (!s || !out) return pg_set_error(PG_ERR_ARG, "sim or out is NULL");
    if (!s->uploaded) return pg_set_error(PG_ERR_STATE, "no state uploaded");
    pg_engine* e = s->eng;
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    HIP_TRY(hipSetDevice(e->device));
    const uint64_t n = s->n;
    for (int i = 0; i < 6; ++i) out[i] = 0.0;
    if (!n) return PG_OK;
    const unsigned nb = (unsigned)((n + kBlock - 1) / kBlock);
    double* part = nullptr;
    HIP_TRY(hipMalloc(&part, (size_t)nb * 5 * 8));
    const int c = s->cur;
    VelocitySource vs;
    int vrc = velocity_source(s, &vs);
    if (vrc != PG_OK) {
        (void)hipFree(part);
        return vrc;
    }
    k_energy<<<nb, kBlock, 0, e->stream>>>(s->pos[c], vs, n, std::fabs(s->p.e), part);
    hipError_t le = hipGetLastError();
    std::vector<double> h((size_t)nb * 5);
    hipError_t ce = hipMemcpyAsync(h.data(), part, h.size() * 8, hipMemcpyDeviceToHost, e->stream);
    hipError_t se = hipStreamSynchronize(e->stream);
    (void)hipFree(part);
    if (le != hipSuccess || ce != hipSuccess || se != hipSuccess)
        return pg_set_error(PG_ERR_HIP, "energy kernel failed");
    long double V = 0, K = 0, P[3] = {0, 0, 0};
    for (unsigned b = 0; b < nb; ++b) {
        V += h[5 * b + 0];
        K += h[5 * b + 1];
        P[0] += h[5 * b + 2];
        P[1] += h[5 * b + 3];
        P[2] += h[5 * b + 4];
    }
    out[0] = (double)K;
    out[1] = (double)V;
    out[2] = (double)(K + V);
    out[3] = (double)P[0];
    out[4] = (double)P[1];
    out[5] = (double)P[2];
    return PG_OK;
}

extern "C" int pg_sim_stats(pg_sim* s, pg_stats* out) {
    if (!s || !out) return pg_set_error(PG_ERR_ARG, "sim or out is NULL");
    pg_engine* e = s->eng;
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    HIP_TRY(hipSetDevice(e->device));
    SegState st;
    HIP_TRY(hipMemcpy(&st, s->st, sizeof st, hipMemcpyDeviceToHost));
    memset(out, 0, sizeof *out);
    out->steps_done = st.steps_done + s->stage_steps;
    out->interactions = st.interactions;
    out->max_level = st.max_level;
    out->error = map_err(st.err);
    out->key_levels = (uint32_t)kMaxLcp + 1u;
    if (e->last_sim == s) {
        // internal nodes stored, plus one leaf per body
        uint32_t ni = 0;
        HIP_TRY(hipMemcpy(&ni, e->first_node + e->last_total, 4, hipMemcpyDeviceToHost));
        out->nodes = (uint64_t)ni + s->n;
    }
    return PG_OK;
}

extern "C" int pg_sim_interaction_counts(pg_sim* s, int64_t* counts) {
    if (!s || (s->n && !counts)) return pg_set_error(PG_ERR_ARG, "sim or counts is NULL");
    if (!s->counting) return pg_set_error(PG_ERR_STATE, "counting is off");
    pg_engine* e = s->eng;
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    HIP_TRY(hipSetDevice(e->device));
    if (s->n) HIP_TRY(hipMemcpy(counts, s->counts, s->n * 8, hipMemcpyDeviceToHost));
    return PG_OK;
}

extern "C" int pg_sim_morton(pg_sim* s, uint64_t* keys, uint32_t* perm, uint64_t n) {
    if (!s) return pg_set_error(PG_ERR_ARG, "sim is NULL");
    pg_engine* e = s->eng;
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    if (e->last_sim != s) return pg_set_error(PG_ERR_STATE, "the last build was not of this sim");
    if (n != s->n) return pg_set_error(PG_ERR_ARG, "n mismatch");
    HIP_TRY(hipSetDevice(e->device));
    if (keys) HIP_TRY(hipMemcpy(keys, e->keys[e->last_sorted], n * 8, hipMemcpyDeviceToHost));
    if (perm) {
        // the upload indices: the tags of the permuted records
        std::vector<double4> pr(n);
        HIP_TRY(hipMemcpy(pr.data(), s->prv[s->cur], n * sizeof(double4), hipMemcpyDeviceToHost));
        for (uint64_t j = 0; j < n; ++j) perm[j] = tag_id(pr[j].w);
    }
    return PG_OK;
}

extern "C" int pg_sim_tree(pg_sim* s, int64_t* rec, double* fl, uint64_t cap, uint64_t* count) {
    if (!s) return pg_set_error(PG_ERR_ARG, "sim is NULL");
    pg_engine* e = s->eng;
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    if (e->last_sim != s) return pg_set_error(PG_ERR_STATE, "the last build was not of this sim");
    HIP_TRY(hipSetDevice(e->device));
    const uint64_t n = s->n;
    // the build stores the internal nodes in preorder; every body is a leaf
    std::vector<uint32_t> fn(n + 1);
    HIP_TRY(hipMemcpy(fn.data(), e->first_node, (n + 1) * 4, hipMemcpyDeviceToHost));
    const uint64_t ni = fn[n];
    const uint64_t nn = ni + n;
    if (count) *count = nn;
    if (!rec && !fl) return PG_OK;
    std::vector<NodeHot> hot(ni);
    std::vector<uint2> span(ni);
    std::vector<uint32_t> id(n);
    std::vector<uint8_t> lo(n);
    std::vector<double4> lp(n);
    std::vector<uint64_t> keys(n);
    std::vector<DeepKey> deep;
    if (ni) {
        HIP_TRY(hipMemcpy(hot.data(), e->hot, ni * sizeof(NodeHot), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(span.data(), e->node_span, ni * sizeof(uint2), hipMemcpyDeviceToHost));
    }
    // the permuted ids live in the buffer the build wrote: the current one
    {
        std::vector<double4> pr(n);
        HIP_TRY(hipMemcpy(pr.data(), s->prv[s->cur], n * sizeof(double4), hipMemcpyDeviceToHost));
        for (uint64_t j = 0; j < n; ++j) id[j] = tag_id(pr[j].w);
    }
    HIP_TRY(hipMemcpy(lo.data(), e->lo_level, n, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(lp.data(), e->leafpos, n * sizeof(double4), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(keys.data(), e->keys[e->last_sorted], n * 8, hipMemcpyDeviceToHost));
    SegState st;
    HIP_TRY(hipMemcpy(&st, s->st, sizeof st, hipMemcpyDeviceToHost));
    double E;
    memcpy(&E, &st.extent_bits, 8);
    const int lmax = e->last_lmax, dim = e->last_dim;
    const uint32_t mask = (1u << dim) - 1u;
    uint64_t out = 0, k = 0;
    auto put = [&](int64_t level, int64_t oct, int64_t body, int64_t first, double mass,
                   double mx, double my, double mz, double cx, double cy, double cz) {
        if (out < cap) {
            if (rec) {
                rec[4 * out + 0] = level;
                rec[4 * out + 1] = oct;
                rec[4 * out + 2] = body;
                rec[4 * out + 3] = first;
            }
            if (fl) {
                double* f = fl + 8 * out;
                f[0] = mass; f[1] = mx; f[2] = my; f[3] = mz;
                f[4] = cx; f[5] = cy; f[6] = cz;
                f[7] = std::ldexp(E, -(int)level);
            }
        }
        ++out;
    };
    for (uint64_t b = 0; b < n && out < cap; ++b) {
        // the internal nodes that start at body b (shallowest first), then its leaf
        for (; k < ni && span[k].x == b; ++k) {
            const NodeHot& h = hot[k];
            const int L = h.level();
            put(L, L > 0 ? (int64_t)h.octant() : -1, -1, (int64_t)b, h.mass, h.mx, h.my, h.mz,
                h.cx, h.cy, h.cz);
        }
        // the leaf's level, octant and centre: the reference's descent
        // (octree.rs:160-215 / quadtree.rs:161-194) along the body's key, the
        // same operations as the device's
        const int L = (int)lo[b] + (int)(fn[b + 1] - fn[b]);
        if (L > lmax && deep.empty()) {
            deep.resize(n);
            HIP_TRY(hipMemcpy(deep.data(), e->deep, n * sizeof(DeepKey), hipMemcpyDeviceToHost));
        }
        const uint64_t key = keys[b];
        const int wl = dim == 3 ? 21 : 32;  // deep key levels per word
        auto digit = [&](int l) {
            if (l < lmax) return (uint32_t)((key >> (dim * (lmax - 1 - l))) & mask);
            const int t = l - lmax;
            return (uint32_t)((deep[b].w[t / wl] >> (dim * (wl - 1 - t % wl))) & mask);
        };
        double ext = E, c[3] = {0.0, 0.0, 0.0};
        for (int l = 0; l < L; ++l) {
            const uint32_t o = digit(l);
            const double h2 = ext * 0.5;
            for (int a = 0; a < dim; ++a) c[a] = ((o >> a) & 1u) ? c[a] + h2 : c[a] - h2;
            ext = h2;
        }
        put(L, L > 0 ? (int64_t)digit(L - 1) : -1, (int64_t)id[b], (int64_t)b, lp[b].w, lp[b].x,
            lp[b].y, lp[b].z, c[0], c[1], c[2]);
    }
    return PG_OK;
}

// ------------------------------------------- allocation-free transform
// pg_transform_aos for callers that cannot allocate device memory. Every
// device buffer of the call -- the staged Entity AoS, the SoA state, keys,
// sort buffers, tree nodes, per-step tables, accelerations and the staged
// Acceleration AoS -- lies in one static device array of kNoAllocBytes,
// `scratch`, carved for at most kNoAllocMaxBodies bodies and
// kNoAllocNodesPerBody tree nodes per body. Each GPU holds its own copy of
// the array (until the process ends), found once per device with
// hipGetSymbolAddress; host code uses only that address. Calls on one GPU
// share the array and are serialised by a per-device lock (calls that
// overlap on one GPU would race on it); the array keeps the last call's
// contents, and every call rewrites or clears what it reads.
constexpr uint64_t kNoAllocMaxBodies = 65536;
constexpr uint64_t kNoAllocNodesPerBody = 4;
constexpr uint64_t kNoAllocBytes = 48ull << 20;
constexpr int kNoAllocDevices = 16;
static_assert(sizeof(SegState) <= 256, "state_bytes reserves 256 bytes for SegState");

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
struct NoAllocCtx {
    pg_engine eng;        // engine view whose scratch is `scratch`
    pg_sim sim;           // the transient simulation, carved from it too
    char* base = nullptr; // this device's address of `scratch`
    double* stage = nullptr;  // staged Acceleration AoS (n * 3 doubles)
    std::mutex mu;
    bool ready = false;
};
NoAllocCtx g_noalloc[kNoAllocDevices];
}  // namespace

// Once per device, with that device current: the array's address, the
// carve-up (state, staging, then the batch scratch) and the cleared status
// words; host tables reserved at their largest, so no later call allocates.
static int noalloc_setup(NoAllocCtx& c, int device, hipStream_t stream) {
    void* ptr = nullptr;
    HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    const uint64_t n = kNoAllocMaxBodies;
    const uint64_t state_sz = align256(state_bytes(n));
    const uint64_t stage_sz = align256(n * 24);
    const uint64_t nodes = n * kNoAllocNodesPerBody;
    const uint32_t blocks = (uint32_t)((n + kBlock - 1) / kBlock) * kMaxDirectSplit;
    pg_engine& e = c.eng;
    const uint64_t need = state_sz + stage_sz + carve_scratch(&e, nullptr, n, nodes, 1, blocks);
    if (need > kNoAllocBytes)
        return pg_set_error(PG_ERR_INTERNAL, "allocation-free layout needs %llu of %llu bytes",
                            (unsigned long long)need, (unsigned long long)kNoAllocBytes);
    c.base = (char*)ptr;
    c.stage = (double*)(c.base + state_sz);
    e.device = device;
    e.stream = stream;
    e.fixed_scratch = true;
    e.graphs = false;
    e.budget = kNoAllocBytes;
    e.scratch = c.base;
    e.scratch_bytes = kNoAllocBytes;
    carve_scratch(&e, c.base + state_sz + stage_sz, n, nodes, 1, blocks);
    PG_TRY(clear_scratch_status(&e));
    const uint64_t chunks = n / kExtentChunk + 2, tiles = n / kNodeTile + 2;
    e.h_sims.reserve(1);
    e.h_nsplit.reserve(1);
    e.h_part_off.reserve(1);
    e.h_chunk0.reserve(1);
    e.h_nchunk.reserve(1);
    e.h_tile0.reserve(1);
    e.h_desc[0].reserve(1);
    e.h_desc[1].reserve(1);
    e.h_seg_off.reserve(2);
    e.h_blk_seg.reserve(blocks);
    e.h_blk_tile.reserve(blocks);
    e.h_ext_seg.reserve(chunks);
    e.h_ext_chunk.reserve(chunks);
    e.h_tile_start.reserve(tiles);
    c.ready = true;
    return PG_OK;
}

extern "C" int pg_transform_aos_noalloc(pg_engine* eng, const pg_params* p, const Entity* state,
                                        uint64_t n, Acceleration* acc) {
    if (!eng) return pg_set_error(PG_ERR_ARG, "engine is NULL");
    PG_TRY(validate_params(p));
    if (n == 0) return PG_OK;
    if (!state || !acc) return pg_set_error(PG_ERR_ARG, "state or acc is NULL");
    if (n > kNoAllocMaxBodies)
        return pg_set_error(PG_ERR_ARG,
                            "n = %llu exceeds the allocation-free transform's maximum of %llu bodies",
                            (unsigned long long)n, (unsigned long long)kNoAllocMaxBodies);
    if (eng->device < 0 || eng->device >= kNoAllocDevices)
        return pg_set_error(PG_ERR_ARG, "device %d beyond the allocation-free path's %d devices",
                            eng->device, kNoAllocDevices);
    std::lock_guard<std::recursive_mutex> lk(eng->mu);
    HIP_TRY(hipSetDevice(eng->device));
    NoAllocCtx& c = g_noalloc[eng->device];
    std::lock_guard<std::mutex> lc(c.mu);
    c.eng.stream = eng->stream;
    if (!c.ready) PG_TRY(noalloc_setup(c, eng->device, eng->stream));
    pg_sim* s = &c.sim;
    s->eng = &c.eng;
    s->n = n;
    s->p = *p;
    s->counting = false;
    s->uploaded = false;
    carve_state(s, c.base, n);
    PG_TRY(pg_sim_upload(s, state, n));
    PG_TRY(run_batch(&c.eng, &s, 1, 1, false));
    // acc[i] += the device's accelerations, added on the device in FP64 (the
    // same sums as pg_transform_aos forms on the host)
    const uint64_t words = n * 3;
    HIP_TRY(hipMemcpyAsync(c.stage, acc, words * 8, hipMemcpyHostToDevice, eng->stream));
    k_add_f64<<<grid_for(words, 1u << 30), kBlock, 0, eng->stream>>>(c.stage, c.eng.acc_out,
                                                                      words);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(acc, c.stage, words * 8, hipMemcpyDeviceToHost, eng->stream));
    HIP_TRY(hipStreamSynchronize(eng->stream));
    return PG_OK;
}
