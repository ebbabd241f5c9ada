// Initial conditions: the reference's `cube`, `star` and `plummer`
// initialisers (astro/src/initialisers.rs) over a ChaCha8 stream laid out as
// rand_chacha 0.9.0's ChaCha8Rng::seed_from_u64, plus a virial-equilibrium
// Plummer sphere that the reference does not have.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "pg_engine.h"
#include "physim_gpu.h"

namespace {

class ChaCha8Stream {
   public:
    // rand_core 0.9 seed_from_u64: a PCG32 sequence supplies the eight key words.
    explicit ChaCha8Stream(uint64_t seed) {
        uint64_t state = seed;
        for (uint32_t& w : key_) {
            state = state * 6364136223846793005ull + 11634580027462260723ull;
            const uint32_t xs = (uint32_t)(((state >> 18) ^ state) >> 27);
            const uint32_t rot = (uint32_t)(state >> 59);
            w = (xs >> rot) | (xs << ((32u - rot) & 31u));
        }
    }

    // Two consecutive keystream words, low word first.
    uint64_t next_u64() {
        if (pos_ >= 64) refill();
        const uint64_t v = (uint64_t)words_[pos_] | ((uint64_t)words_[pos_ + 1] << 32);
        pos_ += 2;
        return v;
    }

    // [0, 1) with 53 random bits (StandardUniform for f64).
    double unit53() { return (double)(next_u64() >> 11) * 0x1.0p-53; }

    // lo + [0,1) * (hi - lo) from 52 random bits (UniformFloat::sample_single).
    double range(double lo, double hi) {
        const double scale = hi - lo;
        for (;;) {
            uint64_t bits = (next_u64() >> 12) | 0x3FF0000000000000ull;
            double one_two;
            memcpy(&one_two, &bits, sizeof one_two);
            const double r = (one_two - 1.0) * scale + lo;
            if (r < hi) return r;
        }
    }

   private:
    static uint32_t rotl(uint32_t v, int c) { return (v << c) | (v >> (32 - c)); }

    // Four 8-round ChaCha blocks at consecutive 64-bit counters, stream 0.
    void refill() {
        for (int blk = 0; blk < 4; ++blk) {
            const uint64_t ctr = counter_ + (uint64_t)blk;
            uint32_t in[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u,
                               key_[0], key_[1], key_[2], key_[3],
                               key_[4], key_[5], key_[6], key_[7],
                               (uint32_t)ctr, (uint32_t)(ctr >> 32), 0u, 0u};
            uint32_t x[16];
            memcpy(x, in, sizeof x);
            static const int qr[8][4] = {{0, 4, 8, 12}, {1, 5, 9, 13}, {2, 6, 10, 14},
                                         {3, 7, 11, 15}, {0, 5, 10, 15}, {1, 6, 11, 12},
                                         {2, 7, 8, 13}, {3, 4, 9, 14}};
            for (int round = 0; round < 8; round += 2) {
                for (const auto& q : qr) {
                    uint32_t &a = x[q[0]], &b = x[q[1]], &c = x[q[2]], &d = x[q[3]];
                    a += b; d = rotl(d ^ a, 16);
                    c += d; b = rotl(b ^ c, 12);
                    a += b; d = rotl(d ^ a, 8);
                    c += d; b = rotl(b ^ c, 7);
                }
            }
            for (int k = 0; k < 16; ++k) words_[16 * blk + k] = x[k] + in[k];
        }
        counter_ += 4;
        pos_ = 0;
    }

    uint32_t key_[8];
    uint64_t counter_ = 0;
    uint32_t words_[64];
    int pos_ = 64;
};

Entity blank() {
    Entity e;
    memset(&e, 0, sizeof e);
    return e;
}

}  // namespace

extern "C" int pg_gen_cube(uint64_t n, uint64_t seed, double spin, double size,
                           const double centre[3], double mass, uint64_t id, Entity* out) {
    if (n && !out) return pg_set_error(PG_ERR_ARG, "out is NULL");
    const double c0 = centre ? centre[0] : 0.0, c1 = centre ? centre[1] : 0.0,
                 c2 = centre ? centre[2] : 0.0;
    ChaCha8Stream rng(seed);
    const double each = mass / (double)n;
    for (uint64_t i = 0; i < n; ++i) {
        // Entity::random (physim-core/src/lib.rs:115-128)
        Entity e = blank();
        e.x = rng.range(-1.0, 1.0);
        e.y = rng.range(-1.0, 1.0);
        e.z = rng.range(0.0, 1.0);
        e.radius = 0.02;
        // RandomCube::create_entities (initialisers.rs:89-104)
        e.x *= size;
        e.y *= size;
        e.z *= size;
        e.vx = e.y * spin;
        e.vy = -e.x * spin;
        e.x += c0;
        e.y += c1;
        e.z += c2;
        e.mass = each;
        e.id = (uintptr_t)id;
        out[i] = e;
    }
    return PG_OK;
}

extern "C" int pg_gen_star(double x, double y, double z, double vx, double vy, double vz,
                           double radius, double mass, uint64_t id, int fixed, Entity* out) {
    if (!out) return pg_set_error(PG_ERR_ARG, "out is NULL");
    Entity e = blank();
    e.x = x; e.y = y; e.z = z;
    e.vx = vx; e.vy = vy; e.vz = vz;
    e.radius = radius;
    e.mass = mass;
    e.id = (uintptr_t)id;
    e.fixed = fixed != 0;
    *out = e;
    return PG_OK;
}

extern "C" int pg_gen_plummer_ref(uint64_t n, uint64_t seed, double mass, double a,
                                  const double centre[3], const double v[3], double spin,
                                  uint64_t id, Entity* out) {
    if (n && !out) return pg_set_error(PG_ERR_ARG, "out is NULL");
    const double c[3] = {centre ? centre[0] : 0.0, centre ? centre[1] : 0.0,
                         centre ? centre[2] : 0.0};
    const double v0[3] = {v ? v[0] : 0.0, v ? v[1] : 0.0, v ? v[2] : 0.0};
    ChaCha8Stream rng(seed);
    const double each = mass / (double)n;
    const double pi = 3.141592653589793;
    for (uint64_t i = 0; i < n; ++i) {
        const double u0 = rng.unit53(), u1 = rng.unit53(), u2 = rng.unit53();
        const double r = a * std::sqrt(std::pow(u0, -2.0 / 3.0) - 1.0);
        const double th = pi * u2;
        const double ph = 2.0 * pi * u1;
        Entity e = blank();
        e.x = r * std::sin(th) * std::cos(ph) + c[0];
        e.y = r * std::sin(th) * std::sin(ph) + c[1];
        e.z = r * std::cos(th) + c[2];
        e.mass = each;
        e.radius = 0.01;
        const double r2 = r * r, a2 = a * a;
        const double vphi = spin * std::sqrt(mass * r2 / std::pow(r2 + a2, 1.5));
        e.vx = -vphi * std::sin(ph) + v0[0];
        e.vy = vphi * std::cos(ph) + v0[1];
        e.vz = 0.0 + v0[2];
        e.id = (uintptr_t)id;
        out[i] = e;
    }
    return PG_OK;
}

extern "C" int pg_gen_plummer_virial(uint64_t n, uint64_t seed, double mass, double a,
                                     const double centre[3], const double v[3], uint64_t id,
                                     Entity* out) {
    if (n && !out) return pg_set_error(PG_ERR_ARG, "out is NULL");
    if (!(a > 0.0)) return pg_set_error(PG_ERR_ARG, "scale radius must be positive");
    ChaCha8Stream rng(seed ^ 0x9E3779B97F4A7C15ull);
    const double each = mass / (double)n;
    const double two_pi = 6.283185307179586;
    double cm[6] = {0, 0, 0, 0, 0, 0};
    for (uint64_t i = 0; i < n; ++i) {
        // radius from the inverted cumulative mass M(r)/M = r^3 / (r^2 + a^2)^{3/2}
        double u;
        do u = rng.unit53(); while (u <= 0.0);
        const double r = a / std::sqrt(std::pow(u, -2.0 / 3.0) - 1.0);
        // isotropic direction
        const double cz = 1.0 - 2.0 * rng.unit53();
        const double ph = two_pi * rng.unit53();
        const double sz = std::sqrt(std::fmax(0.0, 1.0 - cz * cz));
        // speed: q = v / v_esc with density q^2 (1 - q^2)^{7/2} (von Neumann rejection)
        double q, g;
        do {
            q = rng.unit53();
            g = 0.1 * rng.unit53();
        } while (g > q * q * std::pow(1.0 - q * q, 3.5));
        const double vesc = std::sqrt(2.0 * mass / std::sqrt(r * r + a * a));
        const double sp = q * vesc;
        const double vcz = 1.0 - 2.0 * rng.unit53();
        const double vph = two_pi * rng.unit53();
        const double vsz = std::sqrt(std::fmax(0.0, 1.0 - vcz * vcz));
        Entity e = blank();
        e.x = r * sz * std::cos(ph);
        e.y = r * sz * std::sin(ph);
        e.z = r * cz;
        e.vx = sp * vsz * std::cos(vph);
        e.vy = sp * vsz * std::sin(vph);
        e.vz = sp * vcz;
        e.mass = each;
        e.radius = 0.01;
        e.id = (uintptr_t)id;
        out[i] = e;
        cm[0] += e.x; cm[1] += e.y; cm[2] += e.z;
        cm[3] += e.vx; cm[4] += e.vy; cm[5] += e.vz;
    }
    const double shift[6] = {
        (centre ? centre[0] : 0.0) - cm[0] / (double)n, (centre ? centre[1] : 0.0) - cm[1] / (double)n,
        (centre ? centre[2] : 0.0) - cm[2] / (double)n, (v ? v[0] : 0.0) - cm[3] / (double)n,
        (v ? v[1] : 0.0) - cm[4] / (double)n, (v ? v[2] : 0.0) - cm[5] / (double)n};
    for (uint64_t i = 0; i < n; ++i) {
        out[i].x += shift[0]; out[i].y += shift[1]; out[i].z += shift[2];
        out[i].vx += shift[3]; out[i].vy += shift[4]; out[i].vz += shift[5];
    }
    return PG_OK;
}
