// physim oracle: a CPU restatement of the reference's astro / astro2 /
// simple_astro transforms, its Stormer `verlet` integrator and the
// `cube` / `star` / `plummer` initialisers.
//
// TEST INFRASTRUCTURE ONLY. Nothing in the product (physim_amd/) links,
// loads or calls this file. It is loaded by tests/, by the parity check in
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg, always as the
// checker or the timed CPU baseline, never as the thing measured.
//
// Built with `-O2 -ffp-contract=off` and no -march flag so that every FP64
// expression rounds exactly as the reference's Rust (no FMA contraction,
// baseline x86-64 SSE2 code).
//
// Every function cites the reference file:line it restates. Pins (see
// tests/test_oracle_*.py): the reference's 30 tree tests, the ChaCha20
// RFC 8439 / OpenSSL keystream fixture, closed forms (two-body force,
// constant-acceleration Stormer), theta <= 0 == direct sum, brute-force
// opening-test enumeration. The rand-0.9 seeding / float layer
// (rng_seed_from_u64, rng_f64_*) is restated from the crates' published
// algorithms and is "parity unpinned" (no reference output is available in
// this environment to check it against).

#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

// physim-core/src/lib.rs:16-29 (repr(C) Entity, 80 bytes) and :31-37.
struct Entity {
    double x, y, z, vx, vy, vz, radius, mass;
    uintptr_t id;
    bool fixed;
};
struct Acceleration {
    double x, y, z;
};
}
static_assert(sizeof(Entity) == 80, "Entity must be 80 bytes");
static_assert(offsetof(Entity, id) == 64, "id offset");
static_assert(offsetof(Entity, fixed) == 72, "fixed offset");
static_assert(sizeof(Acceleration) == 24, "Acceleration must be 24 bytes");

// ---------------------------------------------------------------------------
// ChaCha block function (RFC 8439 section 2.3 layout with a 64-bit block
// counter in words 12-13 and a 64-bit stream id in words 14-15, as used by
// rand_chacha 0.9.0 -- Cargo.lock:1295-1296). With stream 0 and a counter
// below 2^32 the layout equals RFC 8439's.
// ---------------------------------------------------------------------------
static inline uint32_t rotl32(uint32_t v, int n) { return (v << n) | (v >> (32 - n)); }

static inline void quarter(uint32_t* s, int a, int b, int c, int d) {
    s[a] += s[b]; s[d] ^= s[a]; s[d] = rotl32(s[d], 16);
    s[c] += s[d]; s[b] ^= s[c]; s[b] = rotl32(s[b], 12);
    s[a] += s[b]; s[d] ^= s[a]; s[d] = rotl32(s[d], 8);
    s[c] += s[d]; s[b] ^= s[c]; s[b] = rotl32(s[b], 7);
}

extern "C" void or_chacha_block(const uint32_t key[8], uint64_t counter, uint64_t stream,
                                int rounds, uint32_t out[16]) {
    uint32_t init[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u,
                         key[0], key[1], key[2], key[3], key[4], key[5], key[6], key[7],
                         (uint32_t)counter, (uint32_t)(counter >> 32),
                         (uint32_t)stream, (uint32_t)(stream >> 32)};
    uint32_t s[16];
    memcpy(s, init, sizeof s);
    for (int r = 0; r < rounds; r += 2) {
        quarter(s, 0, 4, 8, 12); quarter(s, 1, 5, 9, 13);
        quarter(s, 2, 6, 10, 14); quarter(s, 3, 7, 11, 15);
        quarter(s, 0, 5, 10, 15); quarter(s, 1, 6, 11, 12);
        quarter(s, 2, 7, 8, 13); quarter(s, 3, 4, 9, 14);
    }
    for (int i = 0; i < 16; ++i) out[i] = s[i] + init[i];
}

// ChaCha8Rng = BlockRng<ChaCha8Core>: a 64-word buffer holding 4 consecutive
// blocks, consumed lowest word first; next_u64 joins two words, low first.
struct OrRng {
    uint32_t key[8];
    uint64_t counter;
    uint32_t buf[64];
    int index;
};

static void rng_refill(OrRng* r) {
    for (int b = 0; b < 4; ++b) or_chacha_block(r->key, r->counter + b, 0, 8, r->buf + 16 * b);
    r->counter += 4;
    r->index = 0;
}

// rand_core 0.9.0 SeedableRng::seed_from_u64: PCG32 (MUL 6364136223846793005,
// INC 11634580027462260723) fills the 32-byte seed, 4 little-endian bytes per
// output. Call sites: astro/src/initialisers.rs:85,296,
// astro/src/octree.rs:365,387,436,499-500. Parity unpinned.
extern "C" void or_rng_seed_from_u64(OrRng* r, uint64_t state) {
    const uint64_t MUL = 6364136223846793005ull, INC = 11634580027462260723ull;
    for (int i = 0; i < 8; ++i) {
        state = state * MUL + INC;
        uint32_t xorshifted = (uint32_t)(((state >> 18) ^ state) >> 27);
        uint32_t rot = (uint32_t)(state >> 59);
        r->key[i] = (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
    }
    r->counter = 0;
    r->index = 64;
}

extern "C" uint64_t or_rng_next_u64(OrRng* r) {
    if (r->index >= 64) rng_refill(r);
    uint64_t lo = r->buf[r->index], hi = r->buf[r->index + 1];
    r->index += 2;
    return lo | (hi << 32);
}

// rand 0.9 StandardUniform for f64: 53 high bits times 2^-53, in [0, 1).
extern "C" double or_rng_f64_standard(OrRng* r) {
    uint64_t u = or_rng_next_u64(r) >> 11;
    return (double)u * (1.0 / 9007199254740992.0);
}

// rand 0.9 UniformFloat::sample_single(low, high): 52 high bits as the
// mantissa of a value in [1,2), minus one, times (high-low), plus low, with
// no fused multiply-add (Rust never contracts). For the ranges used by the
// reference (-1..1 and 0..1) the result is exact and never reaches high.
extern "C" double or_rng_f64_range(OrRng* r, double lo, double hi) {
    const double scale = hi - lo;
    for (;;) {
        uint64_t bits = (or_rng_next_u64(r) >> 12) | 0x3FF0000000000000ull;
        double v12;
        memcpy(&v12, &bits, 8);
        double v01 = v12 - 1.0;
        double res = v01 * scale + lo;
        if (res < hi) return res;
    }
}

extern "C" size_t or_rng_sizeof(void) { return sizeof(OrRng); }

// ---------------------------------------------------------------------------
// Initialisers
// ---------------------------------------------------------------------------

// physim-core/src/lib.rs:94-103 (Entity::new: radius = mass^0.33333).
extern "C" void or_entity_new(double x, double y, double z, double mass, Entity* out) {
    Entity e{};
    e.x = x; e.y = y; e.z = z; e.mass = mass;
    e.radius = pow(mass, 0.33333);
    *out = e;
}

// physim-core/src/lib.rs:115-128 (Entity::random).
extern "C" void or_entity_random(OrRng* r, Entity* out) {
    Entity e{};
    e.x = or_rng_f64_range(r, -1.0, 1.0);
    e.y = or_rng_f64_range(r, -1.0, 1.0);
    e.z = or_rng_f64_range(r, 0.0, 1.0);
    e.mass = 0.005;
    e.radius = 0.02;
    *out = e;
}

// astro/src/initialisers.rs:81-107 (RandomCube::create_entities).
extern "C" void or_gen_cube(uint64_t n, uint64_t seed, double spin, double size,
                            const double centre[3], double mass, uint64_t id, Entity* out) {
    OrRng r;
    or_rng_seed_from_u64(&r, seed);
    const double entity_mass = mass / (double)n;
    for (uint64_t i = 0; i < n; ++i) {
        Entity e;
        or_entity_random(&r, &e);
        e.x *= size; e.y *= size; e.z *= size;
        e.vx = e.y * spin;
        e.vy = -e.x * spin;
        e.x += centre[0]; e.y += centre[1]; e.z += centre[2];
        e.mass = entity_mass;
        e.id = (uintptr_t)id;
        out[i] = e;
    }
}

// astro/src/initialisers.rs:145-191 (SingleStar): defaults radius 0.1, all
// other numbers 0, fixed false.
extern "C" void or_gen_star(double x, double y, double z, double vx, double vy, double vz,
                            double radius, double mass, uint64_t id, int fixed, Entity* out) {
    Entity e{};
    e.x = x; e.y = y; e.z = z; e.vx = vx; e.vy = vy; e.vz = vz;
    e.radius = radius; e.mass = mass; e.id = (uintptr_t)id; e.fixed = fixed != 0;
    *out = e;
}

// astro/src/initialisers.rs:293-330 (Plummer::create_entities, the
// reference's own variant: inverted radius CDF, theta = pi*u, rotation-only
// velocities).
extern "C" void or_gen_plummer_ref(uint64_t n, uint64_t seed, double mass, double a,
                                   const double centre[3], const double v0[3], double spin,
                                   uint64_t id, Entity* out) {
    OrRng r;
    or_rng_seed_from_u64(&r, seed);
    const double m = mass / (double)n;
    for (uint64_t i = 0; i < n; ++i) {
        double c0 = or_rng_f64_standard(&r);
        double c1 = or_rng_f64_standard(&r);
        double c2 = or_rng_f64_standard(&r);
        // powf(0.5) on a non-negative base is lowered to IEEE sqrt (see gravity).
        double rad = a * sqrt(pow(c0, -2.0 / 3.0) - 1.0);
        double theta = M_PI * c2;
        double phi = 2.0 * M_PI * c1;
        Entity e{};
        e.x = rad * sin(theta) * cos(phi) + centre[0];
        e.y = rad * sin(theta) * sin(phi) + centre[1];
        e.z = rad * cos(theta) + centre[2];
        e.mass = m;
        e.radius = 0.01;
        double r2 = rad * rad;
        double a2 = a * a;
        double v_phi = spin * sqrt(mass * r2 / pow(r2 + a2, 1.5));
        e.vx = -v_phi * sin(phi) + v0[0];
        e.vy = v_phi * cos(phi) + v0[1];
        e.vz = 0.0 + v0[2];
        e.id = (uintptr_t)id;
        out[i] = e;
    }
}

// ---------------------------------------------------------------------------
// Star trait for Entity: astro/src/lib.rs:27-113.
// ---------------------------------------------------------------------------
static const double G = 1.0;  // astro/src/lib.rs:27

static int g_fault = 0;  // set where the reference would panic

// astro/src/lib.rs:40-50
static inline void centre_of_mass(const Entity& a, const Entity& b, double out[3]) {
    double total = a.mass + b.mass;
    double inv = 1.0 / total;
    out[0] = (a.mass * a.x + b.mass * b.x) * inv;
    out[1] = (a.mass * a.y + b.mass * b.y) * inv;
    out[2] = (a.mass * a.z + b.mass * b.z) * inv;
}

// astro/src/lib.rs:64-76
static inline Entity fake(const double c[3], double mass) {
    if (std::isnan(c[0])) g_fault |= 2;
    Entity e{};
    e.x = c[0]; e.y = c[1]; e.z = c[2];
    e.radius = 0.0;
    e.mass = mass;
    return e;
}

// astro/src/lib.rs:84-113. powi(2) is x*x; powf(0.5) is lowered by LLVM to
// fabs(sqrt(x)) for the intrinsic (x >= 0 here), i.e. IEEE sqrt.
static inline void gravity(const Entity& a, const Entity& b, double e, double f[3]) {
    double d0 = a.x - b.x, d1 = a.y - b.y, d2 = a.z - b.z;
    double r_norm = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    double r_easing = d0 * d0 + d1 * d1 + d2 * d2 + e;
    double r0 = (b.x - a.x) / r_norm;
    double r1 = (b.y - a.y) / r_norm;
    double r2 = (b.z - a.z) / r_norm;
    double am = a.mass, bm = b.mass;
    f[0] = r0 * G * am * bm / r_easing;
    f[1] = r1 * G * am * bm / r_easing;
    f[2] = r2 * G * am * bm / r_easing;
}

extern "C" void or_gravity(const Entity* a, const Entity* b, double e, double f[3]) {
    gravity(*a, *b, e, f);
}

// ---------------------------------------------------------------------------
// Insertion tree: astro/src/octree.rs:27-216 (D = 3, 8 children) and
// astro/src/quadtree.rs:27-195 (D = 2, 4 children, z centre inherited).
// The bumpalo arena is a vector of nodes linked by index.
// ---------------------------------------------------------------------------
struct TNode {
    double c[3];
    double ext;
    bool has;
    Entity e;
    int64_t body;  // index of the real body held by a leaf, -1 for a fake
    int32_t child[8];
};

struct OrTree {
    int dim;
    int nchild;
    std::vector<TNode> arena;
    uint64_t merges;
    int fault;
};

static int32_t tree_new_node(OrTree* t, const double c[3], double ext) {
    TNode n;
    n.c[0] = c[0]; n.c[1] = c[1]; n.c[2] = c[2];
    n.ext = ext;
    n.has = false;
    n.e = Entity{};
    n.body = -1;
    for (int i = 0; i < 8; ++i) n.child[i] = -1;
    t->arena.push_back(n);
    return (int32_t)t->arena.size() - 1;
}

// octree.rs:160-165 / quadtree.rs:161-165
static inline int octant_id(const OrTree* t, const TNode& n, const double p[3]) {
    int id = (p[0] > n.c[0]) | ((p[1] > n.c[1]) << 1);
    if (t->dim == 3) id |= (p[2] > n.c[2]) << 2;
    return id;
}

// octree.rs:167-215 / quadtree.rs:167-194
static inline void octant_centre(const OrTree* t, const TNode& n, const double p[3], double out[3]) {
    int id = octant_id(t, n, p);
    double h = n.ext / 2.0;
    out[0] = (id & 1) ? n.c[0] + h : n.c[0] - h;
    out[1] = (id & 2) ? n.c[1] + h : n.c[1] - h;
    if (t->dim == 3)
        out[2] = (id & 4) ? n.c[2] + h : n.c[2] - h;
    else
        out[2] = n.c[2];
}

static inline bool no_children(const OrTree* t, const TNode& n) {
    for (int i = 0; i < t->nchild; ++i)
        if (n.child[i] >= 0) return false;
    return true;
}

// octree.rs:59-128 (OctreeNode::push), iterative form of the tail recursion,
// from node `ni` at recursion depth `count0` (the root and 0 for a push).
static void tree_push_at(OrTree* t, int32_t ni, int count0, const Entity& item, int64_t body) {
    for (int count = count0;; ++count) {
        if (count > 64) {  // octree.rs:60-62 "Recursion too deep"
            t->fault |= 1;
            return;
        }
        TNode* n = &t->arena[ni];
        if (!n->has) {
            n->e = item;
            n->body = body;
            n->has = true;
            return;
        }
        const Entity cur = n->e;
        const int64_t cur_body = n->body;
        const bool leaf = no_children(t, *n);
        if (leaf && fabs(cur.x - item.x) < 1e-9 && fabs(cur.y - item.y) < 1e-9 &&
            fabs(cur.z - item.z) < 1e-9) {
            const double c[3] = {item.x, item.y, item.z};
            n->e = fake(c, cur.mass + item.mass);
            n->body = -1;
            t->merges += 1;
            return;
        }
        double com[3];
        centre_of_mass(cur, item, com);
        n->e = fake(com, cur.mass + item.mass);
        n->body = -1;
        if (leaf) {
            const double p[3] = {cur.x, cur.y, cur.z};
            int idx = octant_id(t, *n, p);
            double cc[3];
            octant_centre(t, *n, p, cc);
            double ext = n->ext / 2.0;
            int32_t k = tree_new_node(t, cc, ext);
            n = &t->arena[ni];  // arena may have moved
            t->arena[k].e = cur;
            t->arena[k].body = cur_body;
            t->arena[k].has = true;
            n->child[idx] = k;
        }
        const double p[3] = {item.x, item.y, item.z};
        int idx = octant_id(t, *n, p);
        if (n->child[idx] >= 0) {
            ni = n->child[idx];
            continue;
        }
        double cc[3];
        octant_centre(t, *n, p, cc);
        double ext = n->ext / 2.0;
        int32_t k = tree_new_node(t, cc, ext);
        n = &t->arena[ni];
        t->arena[k].e = item;
        t->arena[k].body = body;
        t->arena[k].has = true;
        n->child[idx] = k;
        return;
    }
}

static void tree_push(OrTree* t, const Entity& item, int64_t body) { tree_push_at(t, 0, 0, item, body); }

extern "C" OrTree* or_tree_new(int dim, const double centre[3], double extent, size_t reserve) {
    OrTree* t = new OrTree();
    t->dim = dim;
    t->nchild = dim == 3 ? 8 : 4;
    t->merges = 0;
    t->fault = 0;
    t->arena.reserve(reserve + 1);
    tree_new_node(t, centre, extent);
    return t;
}

extern "C" void or_tree_free(OrTree* t) { delete t; }

extern "C" void or_tree_push(OrTree* t, const Entity* e, int64_t body) {
    g_fault = 0;
    tree_push(t, *e, body);
    t->fault |= g_fault;
}

extern "C" uint64_t or_tree_merges(const OrTree* t) { return t->merges; }
extern "C" int or_tree_fault(const OrTree* t) { return t->fault; }
extern "C" size_t or_tree_node_count(const OrTree* t) { return t->arena[0].has ? t->arena.size() : 0; }

// octree.rs:130-158 (get_leaves_with_resolution). Appends the interaction
// list (positions and masses of the entities pushed, in the reference's
// order) and optionally the arena index of each contributing node.
//
// margin > 0 adds the build's near field (no reference counterpart): a cell
// at the scale of the body's own leaf (half-width leaf_ext, leaf_ext / 2 or
// leaf_ext / 4: the leaf's level and the two below) is not taken whole while
// loc lies within `margin` of its box (half-width ext + margin on every axis
// the tree splits); it is opened instead.
static void tree_walk(const OrTree* t, const double loc[3], double theta,
                      std::vector<const TNode*>& out, double margin = 0.0,
                      double leaf_ext = 0.0) {
    std::vector<const TNode*> stack;
    stack.reserve(100);
    stack.push_back(&t->arena[0]);
    while (!stack.empty()) {
        const TNode* n = stack.back();
        stack.pop_back();
        if (n->has) {
            double d0 = loc[0] - n->c[0], d1 = loc[1] - n->c[1], d2 = loc[2] - n->c[2];
            double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
            bool take = n->ext / r < theta;
            if (take && margin > 0.0 && n->ext <= leaf_ext && n->ext >= 0.25 * leaf_ext) {
                const double reach = n->ext + margin;
                take = !(fabs(d0) < reach && fabs(d1) < reach && (t->dim == 2 || fabs(d2) < reach));
            }
            if (take) {
                out.push_back(n);
                continue;
            }
        }
        if (!n->has) continue;
        if (no_children(t, *n)) {
            out.push_back(n);
        } else {
            for (int i = 0; i < t->nchild; ++i)
                if (n->child[i] >= 0) stack.push_back(&t->arena[n->child[i]]);
        }
    }
}

// Returns the length of the interaction list; writes up to `cap` entries of
// (x, y, z, mass) into `out4` and the arena node index into `node_ids`.
extern "C" size_t or_tree_walk(const OrTree* t, const double loc[3], double theta, double* out4,
                               int64_t* node_ids, size_t cap) {
    std::vector<const TNode*> res;
    res.reserve(100);
    tree_walk(t, loc, theta, res);
    for (size_t i = 0; i < res.size() && i < cap; ++i) {
        if (out4) {
            out4[4 * i + 0] = res[i]->e.x;
            out4[4 * i + 1] = res[i]->e.y;
            out4[4 * i + 2] = res[i]->e.z;
            out4[4 * i + 3] = res[i]->e.mass;
        }
        if (node_ids) node_ids[i] = res[i] - &t->arena[0];
    }
    return res.size();
}

// Canonical serialisation: preorder, children in octant order 0..n-1.
// rec[k] = {level, octant within parent (-1 for root), leaf body (-1 for an
// internal node, -2 for a merged fake leaf), arena index}; fl[k] =
// {mass, com x, y, z, centre x, y, z, half-width}.
extern "C" size_t or_tree_serialize(const OrTree* t, int64_t* rec, double* fl, size_t cap) {
    if (!t->arena[0].has) return 0;
    struct Item { int32_t node; int level; int oct; };
    std::vector<Item> stack;
    stack.push_back({0, 0, -1});
    size_t k = 0;
    while (!stack.empty()) {
        Item it = stack.back();
        stack.pop_back();
        const TNode& n = t->arena[it.node];
        if (k < cap) {
            bool leaf = no_children(t, n);
            rec[4 * k + 0] = it.level;
            rec[4 * k + 1] = it.oct;
            rec[4 * k + 2] = leaf ? (n.body >= 0 ? n.body : -2) : -1;
            rec[4 * k + 3] = it.node;
            fl[8 * k + 0] = n.e.mass;
            fl[8 * k + 1] = n.e.x;
            fl[8 * k + 2] = n.e.y;
            fl[8 * k + 3] = n.e.z;
            fl[8 * k + 4] = n.c[0];
            fl[8 * k + 5] = n.c[1];
            fl[8 * k + 6] = n.c[2];
            fl[8 * k + 7] = n.ext;
        }
        ++k;
        for (int i = t->nchild - 1; i >= 0; --i)
            if (n.child[i] >= 0) stack.push_back({n.child[i], it.level + 1, i});
    }
    return k;
}

// ---------------------------------------------------------------------------
// Transforms: astro/src/transformers.rs
// ---------------------------------------------------------------------------

// transformers.rs:126-131 / :35-40: max |coord| over every body and axis,
// f64::max fold (NaN-ignoring, as fmax), 1.0 for an empty state.
extern "C" double or_extent(const Entity* s, size_t n) {
    if (n == 0) return 1.0;
    double m = fabs(s[0].x);
    for (size_t i = 0; i < n; ++i) {
        m = fmax(m, fabs(s[i].x));
        m = fmax(m, fabs(s[i].y));
        m = fmax(m, fabs(s[i].z));
    }
    return m;
}

extern "C" OrTree* or_build_tree(int dim, const Entity* s, size_t n) {
    const double c0[3] = {0.0, 0.0, 0.0};
    OrTree* t = or_tree_new(dim, c0, 1.0 * or_extent(s, n), 2 * n);
    g_fault = 0;
    for (size_t i = 0; i < n; ++i) tree_push(t, s[i], (int64_t)i);
    t->fault |= g_fault;
    return t;
}

// The same tree as or_build_tree, built on `threads` host threads. Each
// node's entity is the fold of COM/mass over its bodies in insertion order
// (octree.rs:81-83 at every push through it; the merge rule of :69-80 only
// touches leaves), and a node's subtree depends only on its own bodies in
// insertion order. So the subtrees of the cells at depth kMtDepth are built
// independently (each from an empty node at that depth, its bodies pushed
// in order) and the entities of the nodes above are folds over their bodies
// in order. Where that equivalence could fail -- a node above that depth
// with fewer than two bodies, or whose first two bodies lie within the
// merge distance -- the serial build runs instead.
static bool close_pair(const Entity& a, const Entity& b) {
    return fabs(a.x - b.x) < 1e-9 && fabs(a.y - b.y) < 1e-9 && fabs(a.z - b.z) < 1e-9;
}

static Entity fold(const Entity* s, const uint32_t* idx, size_t cnt) {
    Entity cur = s[idx[0]];
    for (size_t k = 1; k < cnt; ++k) {
        double com[3];
        centre_of_mass(cur, s[idx[k]], com);
        cur = fake(com, cur.mass + s[idx[k]].mass);
    }
    return cur;
}

extern "C" OrTree* or_build_tree_mt(int dim, const Entity* s, size_t n, int threads) {
    constexpr int kMtDepth = 3;
    if (n < 4096 || threads <= 1) return or_build_tree(dim, s, n);
    const double c0[3] = {0.0, 0.0, 0.0};
    OrTree* t = or_tree_new(dim, c0, 1.0 * or_extent(s, n), 0);
    const int bits = dim;  // octant bits per level
    const uint32_t ncells[kMtDepth + 1] = {1u, 1u << bits, 1u << (2 * bits), 1u << (3 * bits)};
    // each body's cell path to depth kMtDepth, by the reference's descent
    std::vector<uint32_t> cell(n);
    const TNode root = t->arena[0];
#pragma omp parallel for schedule(static) num_threads(threads)
    for (size_t i = 0; i < n; ++i) {
        const double p[3] = {s[i].x, s[i].y, s[i].z};
        TNode nd = root;
        uint32_t id = 0;
        for (int d = 0; d < kMtDepth; ++d) {
            const int o = octant_id(t, nd, p);
            TNode ch;
            octant_centre(t, nd, p, ch.c);
            ch.ext = nd.ext / 2.0;
            nd = ch;
            id = (id << bits) | (uint32_t)o;
        }
        cell[i] = id;
    }
    // the bodies of every node at depth d <= kMtDepth in insertion order: a
    // stable counting sort by the node's prefix of the cell path
    std::vector<size_t> start[kMtDepth + 1];
    std::vector<uint32_t> order[kMtDepth + 1];
#pragma omp parallel for schedule(static, 1) num_threads(threads)
    for (int d = 0; d <= kMtDepth; ++d) {
        const int sh = bits * (kMtDepth - d);
        start[d].assign(ncells[d] + 1, 0);
        for (size_t i = 0; i < n; ++i) ++start[d][(cell[i] >> sh) + 1];
        for (uint32_t c = 0; c < ncells[d]; ++c) start[d][c + 1] += start[d][c];
        order[d].resize(n);
        std::vector<size_t> fill(start[d].begin(), start[d].end() - 1);
        for (size_t i = 0; i < n; ++i) order[d][fill[cell[i] >> sh]++] = (uint32_t)i;
    }
    auto count_of = [&](int d, uint32_t q) { return start[d][q + 1] - start[d][q]; };
    // checks and entities of the nodes above the subtrees; the subtrees
    std::vector<Entity> ent[kMtDepth];
    for (int d = 0; d < kMtDepth; ++d) ent[d].assign(ncells[d], Entity{});
    const uint32_t nleaf = ncells[kMtDepth];
    std::vector<OrTree*> sub(nleaf, nullptr);
    int bad = 0;
    const uint32_t nupper = ncells[0] + ncells[1] + ncells[2];
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads) reduction(| : bad)
    for (uint32_t task = 0; task < nupper + nleaf; ++task) {
        if (task < nupper) {
            const int d = task < ncells[0] ? 0 : task < ncells[0] + ncells[1] ? 1 : 2;
            const uint32_t q = task - (d == 0 ? 0u : d == 1 ? ncells[0] : ncells[0] + ncells[1]);
            const size_t cnt = count_of(d, q);
            if (cnt == 0) continue;
            const uint32_t* ids = order[d].data() + start[d][q];
            if (cnt < 2 || close_pair(s[ids[0]], s[ids[1]])) {
                bad |= 1;
                continue;
            }
            ent[d][q] = fold(s, ids, cnt);
        } else {
            const uint32_t c = task - nupper;
            const size_t cnt = count_of(kMtDepth, c);
            if (cnt == 0) continue;
            const uint32_t* ids = order[kMtDepth].data() + start[kMtDepth][c];
            const double p[3] = {s[ids[0]].x, s[ids[0]].y, s[ids[0]].z};
            TNode nd = root;
            for (int d = 0; d < kMtDepth; ++d) {
                TNode ch;
                octant_centre(t, nd, p, ch.c);
                ch.ext = nd.ext / 2.0;
                nd = ch;
            }
            OrTree* st = or_tree_new(dim, nd.c, nd.ext, 2 * cnt);
            for (size_t k = 0; k < cnt; ++k) tree_push_at(st, 0, kMtDepth, s[ids[k]], (int64_t)ids[k]);
            sub[c] = st;
        }
    }
    if (bad) {
        for (OrTree* st : sub)
            if (st) or_tree_free(st);
        or_tree_free(t);
        return or_build_tree(dim, s, n);
    }
    // one arena: the nodes above depth kMtDepth, then each subtree (links
    // offset), copied in parallel into place
    std::vector<int32_t> idx_prev(1, 0), idx_cur;
    t->arena[0].e = ent[0][0];
    t->arena[0].has = true;
    t->arena[0].body = -1;
    std::vector<size_t> sub_base(nleaf, 0);
    {
        size_t total = t->arena.size();
        for (int d = 1; d < kMtDepth; ++d)
            for (uint32_t q = 0; q < ncells[d]; ++q) total += count_of(d, q) ? 1 : 0;
        for (uint32_t c = 0; c < nleaf; ++c) total += sub[c] ? sub[c]->arena.size() : 0;
        t->arena.reserve(total);
    }
    for (int d = 1; d <= kMtDepth; ++d) {
        idx_cur.assign(ncells[d], -1);
        for (uint32_t q = 0; q < ncells[d]; ++q) {
            const uint32_t parent = q >> bits, o = q & ((1u << bits) - 1u);
            if (idx_prev[parent] < 0 || count_of(d, q) == 0) continue;
            int32_t k;
            if (d < kMtDepth) {
                // the centre by the reference's descent along any body's path
                const uint32_t b0 = order[d][start[d][q]];
                const double p[3] = {s[b0].x, s[b0].y, s[b0].z};
                double cc[3];
                octant_centre(t, t->arena[idx_prev[parent]], p, cc);
                k = tree_new_node(t, cc, t->arena[idx_prev[parent]].ext / 2.0);
                t->arena[k].e = ent[d][q];
                t->arena[k].has = true;
                t->arena[k].body = -1;
            } else {
                k = (int32_t)t->arena.size();
                sub_base[q] = (size_t)k;
                t->arena.insert(t->arena.end(), sub[q]->arena.size(), TNode{});
                t->merges += sub[q]->merges;
                t->fault |= sub[q]->fault;
            }
            t->arena[idx_prev[parent]].child[o] = k;
            idx_cur[q] = k;
        }
        idx_prev.swap(idx_cur);
    }
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
    for (uint32_t c = 0; c < nleaf; ++c) {
        OrTree* st = sub[c];
        if (!st) continue;
        const int32_t base = (int32_t)sub_base[c];
        TNode* dst = t->arena.data() + base;
        for (size_t k = 0; k < st->arena.size(); ++k) {
            TNode nd = st->arena[k];
            for (int r = 0; r < 8; ++r)
                if (nd.child[r] >= 0) nd.child[r] += base;
            dst[k] = nd;
        }
        or_tree_free(st);
    }
    return t;
}

// transformers.rs:123-160 (astro2, dim 3) and :32-69 (astro, dim 2).
// Accumulates into acc (+=). Returns the tree fault flags (0 = ok).
// `counts`, when given, receives each body's interaction-list length.
extern "C" int or_bh_transform(int dim, const Entity* s, size_t n, double theta, double e,
                               Acceleration* acc, int64_t* counts) {
    OrTree* t = or_build_tree(dim, s, n);
    int fault = t->fault;
    std::vector<const TNode*> list;
    for (size_t i = 0; i < n; ++i) {
        if (counts) counts[i] = 0;
        const Entity& a = s[i];
        if (a.fixed) continue;
        double f[3] = {0.0, 0.0, 0.0};
        const double loc[3] = {a.x, a.y, a.z};
        list.clear();
        tree_walk(t, loc, theta, list);
        if (counts) counts[i] = (int64_t)list.size();
        for (const TNode* nb : list) {
            const Entity& b = nb->e;
            if (a.x == b.x && a.y == b.y && a.z == b.z) continue;
            double fij[3];
            gravity(a, b, e, fij);
            f[0] += fij[0]; f[1] += fij[1]; f[2] += fij[2];
        }
        acc[i].x += f[0] / a.mass;
        acc[i].y += f[1] / a.mass;
        acc[i].z += f[2] / a.mass;
    }
    or_tree_free(t);
    return fault;
}

// The same transform with the reference's per-body allocation pattern
// (fresh result and stack vectors per walk, copies of the 80-byte entities)
// so that its wall time is representative of the reference. Used only as the
// timed CPU baseline.
extern "C" int or_bh_transform_timed(int dim, const Entity* s, size_t n, double theta, double e,
                                     Acceleration* acc) {
    OrTree* t = or_build_tree(dim, s, n);
    int fault = t->fault;
    for (size_t i = 0; i < n; ++i) {
        const Entity& a = s[i];
        if (a.fixed) continue;
        double f[3] = {0.0, 0.0, 0.0};
        const double loc[3] = {a.x, a.y, a.z};
        std::vector<const TNode*> list;
        list.reserve(100);
        tree_walk(t, loc, theta, list);
        std::vector<Entity> bs;
        bs.reserve(list.size());
        for (const TNode* nb : list) bs.push_back(nb->e);
        for (const Entity& b : bs) {
            if (a.x == b.x && a.y == b.y && a.z == b.z) continue;
            double fij[3];
            gravity(a, b, e, fij);
            f[0] += fij[0]; f[1] += fij[1]; f[2] += fij[2];
        }
        acc[i].x += f[0] / a.mass;
        acc[i].y += f[1] / a.mass;
        acc[i].z += f[2] / a.mass;
    }
    or_tree_free(t);
    return fault;
}

// transformers.rs:219-244 (simple_astro).
extern "C" void or_simple_transform(const Entity* s, size_t n, double e, Acceleration* acc) {
    for (size_t i = 0; i < n; ++i) {
        const Entity& a = s[i];
        if (a.fixed) continue;
        double f[3] = {0.0, 0.0, 0.0};
        for (size_t j = 0; j < n; ++j) {
            const Entity& b = s[j];
            if (a.x == b.x && a.y == b.y && a.z == b.z) continue;
            double fij[3];
            gravity(a, b, e, fij);
            f[0] += fij[0]; f[1] += fij[1]; f[2] += fij[2];
        }
        acc[i].x += f[0] / a.mass;
        acc[i].y += f[1] / a.mass;
        acc[i].z += f[2] / a.mass;
    }
}

// Direct sum for a subset of target bodies (rows) against all sources, the
// same arithmetic as or_simple_transform, so that large-N checks can sample.
extern "C" void or_simple_transform_rows(const Entity* s, size_t n, double e, const int64_t* rows,
                                         size_t nrows, Acceleration* acc) {
    for (size_t r = 0; r < nrows; ++r) {
        const Entity& a = s[rows[r]];
        if (a.fixed) continue;
        double f[3] = {0.0, 0.0, 0.0};
        for (size_t j = 0; j < n; ++j) {
            const Entity& b = s[j];
            if (a.x == b.x && a.y == b.y && a.z == b.z) continue;
            double fij[3];
            gravity(a, b, e, fij);
            f[0] += fij[0]; f[1] += fij[1]; f[2] += fij[2];
        }
        acc[r].x += f[0] / a.mass;
        acc[r].y += f[1] / a.mass;
        acc[r].z += f[2] / a.mass;
    }
}

// The astro2 / astro transform (or_bh_transform's arithmetic) for a subset of
// target bodies (rows) against the whole tree, so that large-N checks can
// sample. acc[r] += the acceleration of body rows[r]; counts[r] (optional)
// receives its interaction-list length. Returns the tree fault flags.
extern "C" int or_bh_transform_rows(int dim, const Entity* s, size_t n, double theta, double e,
                                    const int64_t* rows, size_t nrows, Acceleration* acc,
                                    int64_t* counts) {
    OrTree* t = or_build_tree(dim, s, n);
    int fault = t->fault;
    std::vector<const TNode*> list;
    for (size_t r = 0; r < nrows; ++r) {
        if (counts) counts[r] = 0;
        const Entity& a = s[rows[r]];
        if (a.fixed) continue;
        double f[3] = {0.0, 0.0, 0.0};
        const double loc[3] = {a.x, a.y, a.z};
        list.clear();
        tree_walk(t, loc, theta, list);
        if (counts) counts[r] = (int64_t)list.size();
        for (const TNode* nb : list) {
            const Entity& b = nb->e;
            if (a.x == b.x && a.y == b.y && a.z == b.z) continue;
            double fij[3];
            gravity(a, b, e, fij);
            f[0] += fij[0]; f[1] += fij[1]; f[2] += fij[2];
        }
        acc[r].x += f[0] / a.mass;
        acc[r].y += f[1] / a.mass;
        acc[r].z += f[2] / a.mass;
    }
    or_tree_free(t);
    return fault;
}

// Half-width of the leaf cell that holds loc (the descent of octree.rs:
// 160-165 from the root until a node without children).
static double leaf_ext(const OrTree* t, const double loc[3]) {
    const TNode* n = &t->arena[0];
    for (;;) {
        if (no_children(t, *n)) return n->ext;
        const int32_t c = n->child[octant_id(t, *n, loc)];
        if (c < 0) return n->ext;
        n = &t->arena[c];
    }
}

// or_bh_transform_rows with the near field: body a's margin is `near` widths
// of its own leaf cell (2 * its half-width). rows = NULL: every body.
extern "C" int or_bh_transform_near(int dim, const Entity* s, size_t n, double theta, double e,
                                    double near, const int64_t* rows, size_t nrows,
                                    Acceleration* acc, int64_t* counts) {
    OrTree* t = or_build_tree(dim, s, n);
    int fault = t->fault;
    std::vector<const TNode*> list;
    const size_t m = rows ? nrows : n;
    for (size_t r = 0; r < m; ++r) {
        if (counts) counts[r] = 0;
        const Entity& a = s[rows ? rows[r] : r];
        if (a.fixed) continue;
        double f[3] = {0.0, 0.0, 0.0};
        const double loc[3] = {a.x, a.y, a.z};
        list.clear();
        const double lx = leaf_ext(t, loc);
        tree_walk(t, loc, theta, list, near * (2.0 * lx), lx);
        if (counts) counts[r] = (int64_t)list.size();
        for (const TNode* nb : list) {
            const Entity& b = nb->e;
            if (a.x == b.x && a.y == b.y && a.z == b.z) continue;
            double fij[3];
            gravity(a, b, e, fij);
            f[0] += fij[0]; f[1] += fij[1]; f[2] += fij[2];
        }
        acc[r].x += f[0] / a.mass;
        acc[r].y += f[1] / a.mass;
        acc[r].z += f[2] / a.mass;
    }
    or_tree_free(t);
    return fault;
}

// ---------------------------------------------------------------------------
// verlet: integrators/src/verlet.rs:19-108
// ---------------------------------------------------------------------------
struct OrVerlet {
    std::vector<Entity> previous_state;
};

extern "C" OrVerlet* or_verlet_new(void) { return new OrVerlet(); }
extern "C" void or_verlet_free(OrVerlet* v) { delete v; }

// verlet.rs:85-108 after acc_fn has filled `acc`.
extern "C" void or_verlet_integrate(OrVerlet* v, const Entity* ent, Entity* new_state, size_t n,
                                    const Acceleration* acc, double dt) {
    const double dt2 = dt * dt;  // dt.powi(2)
    if (v->previous_state.size() != n) {
        // verlet.rs:24-50 (initial_integration)
        v->previous_state.assign(ent, ent + n);
        for (size_t i = 0; i < n; ++i) {
            const Entity& e = ent[i];
            const Acceleration& a = acc[i];
            Entity ne = e;
            ne.x = e.x + e.vx * dt + 0.5 * a.x * dt2;
            ne.y = e.y + e.vy * dt + 0.5 * a.y * dt2;
            ne.z = e.z + e.vz * dt + 0.5 * a.z * dt2;
            ne.vx = e.vx + a.x * dt;
            ne.vy = e.vy + a.y * dt;
            ne.vz = e.vz + a.z * dt;
            new_state[i] = ne;
        }
    } else {
        // verlet.rs:52-82 (integration)
        for (size_t i = 0; i < n; ++i) {
            const Entity& e = ent[i];
            const Entity& p = v->previous_state[i];
            const Acceleration& a = acc[i];
            double x = 2.0 * e.x - p.x + a.x * dt2;
            double y = 2.0 * e.y - p.y + a.y * dt2;
            double z = 2.0 * e.z - p.z + a.z * dt2;
            Entity ne = e;
            ne.x = x; ne.y = y; ne.z = z;
            ne.vx = (x - e.x) / dt;
            ne.vy = (y - e.y) / dt;
            ne.vz = (z - e.z) / dt;
            new_state[i] = ne;
        }
        v->previous_state.assign(ent, ent + n);
    }
}

// One pipeline step in the reference's order (pipeline.rs:143-182 with a
// single gravity transform and verlet): zeroed accelerations (verlet.rs:93),
// acc_fn (verlet.rs:94), integrate, state = new_state (pipeline.rs:173).
// kind: 0 = simple_astro, 2 = astro (quadtree), 3 = astro2 (octree).
extern "C" int or_pipeline_steps(OrVerlet* v, Entity* state, size_t n, int kind, double theta,
                                 double e, double dt, int steps) {
    std::vector<Acceleration> acc(n);
    std::vector<Entity> new_state(state, state + n);
    int fault = 0;
    for (int s = 0; s < steps; ++s) {
        std::fill(acc.begin(), acc.end(), Acceleration{0.0, 0.0, 0.0});
        if (kind == 0)
            or_simple_transform(state, n, e, acc.data());
        else
            fault |= or_bh_transform(kind, state, n, theta, e, acc.data(), nullptr);
        or_verlet_integrate(v, state, new_state.data(), n, acc.data(), dt);
        std::copy(new_state.begin(), new_state.end(), state);
    }
    return fault;
}

// Timed CPU baseline of one pipeline step (transform + integrate only).
extern "C" int or_pipeline_steps_timed(OrVerlet* v, Entity* state, size_t n, double theta,
                                       double e, double dt, int steps) {
    std::vector<Acceleration> acc(n);
    std::vector<Entity> new_state(state, state + n);
    int fault = 0;
    for (int s = 0; s < steps; ++s) {
        std::fill(acc.begin(), acc.end(), Acceleration{0.0, 0.0, 0.0});
        fault |= or_bh_transform_timed(3, state, n, theta, e, acc.data());
        or_verlet_integrate(v, state, new_state.data(), n, acc.data(), dt);
        std::copy(new_state.begin(), new_state.end(), state);
    }
    return fault;
}

// The timed baseline on `threads` host threads (OpenMP): the insertion
// tree of transformers.rs:132-135 built in parallel (or_build_tree_mt, the
// same tree bit for bit), then the per-body walks and pair sums of
// transformers.rs:137-159 split over the threads (each body's own sum keeps
// the reference's order, so the accelerations equal the serial ones bit for
// bit), then the Stormer update (verlet.rs:52-82) split the same way.
extern "C" int or_pipeline_steps_timed_mt(OrVerlet* v, Entity* state, size_t n, double theta,
                                          double e, double dt, int steps, int threads) {
    std::vector<Acceleration> acc(n);
    std::vector<Entity> new_state(state, state + n);
    int fault = 0;
    const double dt2 = dt * dt;
    for (int st = 0; st < steps; ++st) {
        std::fill(acc.begin(), acc.end(), Acceleration{0.0, 0.0, 0.0});
        OrTree* t = or_build_tree_mt(3, state, n, threads);
        fault |= t->fault;
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads)
        for (size_t i = 0; i < n; ++i) {
            const Entity& a = state[i];
            if (a.fixed) continue;
            double f[3] = {0.0, 0.0, 0.0};
            const double loc[3] = {a.x, a.y, a.z};
            std::vector<const TNode*> list;
            list.reserve(100);
            tree_walk(t, loc, theta, list);
            std::vector<Entity> bs;
            bs.reserve(list.size());
            for (const TNode* nb : list) bs.push_back(nb->e);
            for (const Entity& b : bs) {
                if (a.x == b.x && a.y == b.y && a.z == b.z) continue;
                double fij[3];
                gravity(a, b, e, fij);
                f[0] += fij[0]; f[1] += fij[1]; f[2] += fij[2];
            }
            acc[i].x += f[0] / a.mass;
            acc[i].y += f[1] / a.mass;
            acc[i].z += f[2] / a.mass;
        }
        or_tree_free(t);
        if (v->previous_state.size() != n) {
            or_verlet_integrate(v, state, new_state.data(), n, acc.data(), dt);
        } else {
            const Entity* p = v->previous_state.data();
#pragma omp parallel for schedule(static) num_threads(threads)
            for (size_t i = 0; i < n; ++i) {
                const Entity& en = state[i];
                const Acceleration& a = acc[i];
                double x = 2.0 * en.x - p[i].x + a.x * dt2;
                double y = 2.0 * en.y - p[i].y + a.y * dt2;
                double z = 2.0 * en.z - p[i].z + a.z * dt2;
                Entity ne = en;
                ne.x = x; ne.y = y; ne.z = z;
                ne.vx = (x - en.x) / dt;
                ne.vy = (y - en.y) / dt;
                ne.vz = (z - en.z) / dt;
                new_state[i] = ne;
            }
            v->previous_state.assign(state, state + n);
        }
        std::copy(new_state.begin(), new_state.end(), state);
    }
    return fault;
}

// The compiler that built this file (reported beside the CPU baseline).
extern "C" const char* or_compiler(void) { return __VERSION__; }

extern "C" int or_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// acc_fn of the pipeline (pipeline.rs:137-141) with one gravity transform:
// accelerations zeroed by the integrator, then the transform adds into them.
static int acc_fn(int kind, const Entity* s, size_t n, double theta, double e,
                  Acceleration* acc) {
    std::fill(acc, acc + n, Acceleration{0.0, 0.0, 0.0});
    if (kind == 0) {
        or_simple_transform(s, n, e, acc);
        return 0;
    }
    return or_bh_transform(kind, s, n, theta, e, acc, nullptr);
}

// integrators/src/euler.rs:17-49: x' = x + v dt + 0.5 a dt^2 (dt.powi(2) =
// dt * dt), v' = v + a dt, every body (no case for `fixed`: a fixed body has
// a = 0 and moves with its velocity); other fields copied.
extern "C" int or_euler_step(Entity* state, size_t n, int kind, double theta, double e,
                             double dt) {
    std::vector<Acceleration> acc(n);
    int fault = acc_fn(kind, state, n, theta, e, acc.data());
    const double dt2 = dt * dt;
    for (size_t i = 0; i < n; ++i) {
        Entity& en = state[i];
        const Acceleration& a = acc[i];
        const double x = en.x + en.vx * dt + 0.5 * a.x * dt2;
        const double y = en.y + en.vy * dt + 0.5 * a.y * dt2;
        const double z = en.z + en.vz * dt + 0.5 * a.z * dt2;
        const double vx = en.vx + a.x * dt;
        const double vy = en.vy + a.y * dt;
        const double vz = en.vz + a.z * dt;
        en.x = x; en.y = y; en.z = z;
        en.vx = vx; en.vy = vy; en.vz = vz;
    }
    return fault;
}

// integrators/src/rk4.rs:24-182. Four acc_fn calls (:36, :70, :104, :138) on
// the base state and on base + 0.5 k1, base + 0.5 k2, base + k3 (positions
// and velocities of every body, fixed ones included, other fields copied);
// k = (dt * v_stage, dt * a_stage); the result base + (k1 + 2 k2 + 2 k3 + k4)
// / 6.0 summed left to right, and for a fixed body its position with zero
// velocity (:160-174).
extern "C" int or_rk4_step(Entity* state, size_t n, int kind, double theta, double e, double dt) {
    struct K6 { double x, y, z, vx, vy, vz; };
    std::vector<Acceleration> f(n);
    std::vector<Entity> tmp(state, state + n);
    std::vector<K6> k[4];
    int fault = 0;
    for (int s = 0; s < 4; ++s) {
        const Entity* cur = s == 0 ? state : tmp.data();
        fault |= acc_fn(kind, cur, n, theta, e, f.data());
        k[s].resize(n);
        for (size_t i = 0; i < n; ++i) {
            const Entity& c = cur[i];
            k[s][i] = K6{dt * c.vx, dt * c.vy, dt * c.vz, dt * f[i].x, dt * f[i].y, dt * f[i].z};
        }
        if (s == 3) break;
        const double w = s < 2 ? 0.5 : 1.0;
        for (size_t i = 0; i < n; ++i) {
            const Entity& b = state[i];
            const K6& q = k[s][i];
            Entity t = b;
            t.x = b.x + w * q.x;
            t.y = b.y + w * q.y;
            t.z = b.z + w * q.z;
            t.vx = b.vx + w * q.vx;
            t.vy = b.vy + w * q.vy;
            t.vz = b.vz + w * q.vz;
            tmp[i] = t;
        }
    }
    for (size_t i = 0; i < n; ++i) {
        Entity& b = state[i];
        const K6 &a = k[0][i], &c = k[1][i], &d = k[2][i], &g = k[3][i];
        if (!b.fixed) {
            const double x = b.x + (a.x + 2.0 * c.x + 2.0 * d.x + g.x) / 6.0;
            const double y = b.y + (a.y + 2.0 * c.y + 2.0 * d.y + g.y) / 6.0;
            const double z = b.z + (a.z + 2.0 * c.z + 2.0 * d.z + g.z) / 6.0;
            const double vx = b.vx + (a.vx + 2.0 * c.vx + 2.0 * d.vx + g.vx) / 6.0;
            const double vy = b.vy + (a.vy + 2.0 * c.vy + 2.0 * d.vy + g.vy) / 6.0;
            const double vz = b.vz + (a.vz + 2.0 * c.vz + 2.0 * d.vz + g.vz) / 6.0;
            b.x = x; b.y = y; b.z = z;
            b.vx = vx; b.vy = vy; b.vz = vz;
        } else {
            b.vx = 0.0; b.vy = 0.0; b.vz = 0.0;
        }
    }
    return fault;
}

// ---------------------------------------------------------------------------
// Energy and momentum, consistent with the force law of astro/src/lib.rs:84-113
// (|F| = m_a m_b / (r^2 + e)): U(r) = -(m_a m_b / sqrt(e)) atan(sqrt(e) / r),
// U = -m_a m_b / r for e = 0. The reference's own energysink
// (debug/src/energysink.rs:78-99) overwrites instead of accumulating, so it is
// not restated. out = {K, V, E, Px, Py, Pz}. Pairs at identical positions are
// skipped, as the force skips them.
// ---------------------------------------------------------------------------
extern "C" void or_energy(const Entity* s, size_t n, double e, double out[6]) {
    long double K = 0, V = 0, P[3] = {0, 0, 0};
    const double se = sqrt(e);
    for (size_t i = 0; i < n; ++i) {
        const Entity& a = s[i];
        K += 0.5L * a.mass * ((long double)a.vx * a.vx + (long double)a.vy * a.vy +
                              (long double)a.vz * a.vz);
        P[0] += (long double)a.mass * a.vx;
        P[1] += (long double)a.mass * a.vy;
        P[2] += (long double)a.mass * a.vz;
        for (size_t j = i + 1; j < n; ++j) {
            const Entity& b = s[j];
            if (a.x == b.x && a.y == b.y && a.z == b.z) continue;
            double d0 = a.x - b.x, d1 = a.y - b.y, d2 = a.z - b.z;
            double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
            double mm = a.mass * b.mass;
            V += e > 0.0 ? -(mm / se) * atan(se / r) : -mm / r;
        }
    }
    out[0] = (double)K;
    out[1] = (double)V;
    out[2] = (double)(K + V);
    out[3] = (double)P[0];
    out[4] = (double)P[1];
    out[5] = (double)P[2];
}
