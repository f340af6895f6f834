// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11), and the per-history
// random stream the transport draws from.
//
// One history owns one stream: key = 64-bit master seed, counter =
// {block, history, spot, (scenario << 8) | purpose}.  Each block yields four
// 32-bit words; `block` advances every four draws.  Nothing is stored per
// history on the device, so any history can be replayed on the host from its
// coordinates alone.  This replaces the reference's per-history
// curand_init(transport_seed[i]) (mqi_transport.hpp:147-149) and its host
// side seed array (mqi_tps_env.hpp:1537-1550).
#ifndef MQI_PHILOX_HPP
#define MQI_PHILOX_HPP

#include "mqi_pmath.hpp"

namespace mqi_rng
{

struct u32x4 {
    uint32_t v[4];
};

MQI_HD inline void
mulhilo32(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
    const uint64_t p = (uint64_t) a * (uint64_t) b;
    *hi              = (uint32_t) (p >> 32);
    *lo              = (uint32_t) p;
}

MQI_HD inline u32x4
philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#if defined(__HIP__)
#pragma unroll
#endif
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo32(M0, c.v[0], &hi0, &lo0);
        mulhilo32(M1, c.v[2], &hi1, &lo1);
        u32x4 n;
        n.v[0] = hi1 ^ c.v[1] ^ k0;
        n.v[1] = lo1;
        n.v[2] = hi0 ^ c.v[3] ^ k1;
        n.v[3] = lo0;
        c      = n;
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// STREAM_STEP and STREAM_INTERACTION: the production transport's
// block-per-step and block-per-interaction draws (mqi_device.hpp rng_t)
enum purpose_t : uint32_t { STREAM_TRANSPORT = 0, STREAM_VERTEX = 1, STREAM_STEP = 2, STREAM_INTERACTION = 3 };

// Per-history stream.  uniform() returns k * 2^-24, k in [1, 2^24], i.e. a
// float in (0, 1] like curand_uniform; normal() is Box-Muller on a pair of
// uniforms with the second variate cached, evaluated in double with the
// portable functions and rounded once to float.
struct stream {
    uint32_t key0, key1;
    uint32_t c1, c2, c3;
    uint32_t block;
    uint32_t q0, q1, q2, q3;   // unread words of the current block, next first
    int      nbuf;
    int      has_spare;
    double   spare;

    MQI_HD void
    init(uint64_t seed, uint32_t history, uint32_t spot, uint32_t scenario, uint32_t purpose) {
        key0      = (uint32_t) seed;
        key1      = (uint32_t) (seed >> 32);
        c1        = history;
        c2        = spot;
        c3        = (scenario << 8) | purpose;
        block     = 0;
        nbuf      = 0;
        has_spare = 0;
        spare     = 0.0;
    }

    // Words of a block are consumed in the order o[0], o[1], o[2], o[3].  The
    // queue is shifted rather than indexed: a runtime index into an array
    // member would put the stream in private memory on the GPU.
    MQI_HD uint32_t
    next_u32() {
        if (nbuf == 0) {
            u32x4 c;
            c.v[0] = block++;
            c.v[1] = c1;
            c.v[2] = c2;
            c.v[3] = c3;
            u32x4 o = philox4x32_10(c, key0, key1);
            q0      = o.v[0];
            q1      = o.v[1];
            q2      = o.v[2];
            q3      = o.v[3];
            nbuf    = 4;
        }
        const uint32_t r = q0;
        q0               = q1;
        q1               = q2;
        q2               = q3;
        --nbuf;
        return r;
    }

    MQI_HD float
    uniform() {
        return to_uniform(next_u32());
    }

    // k * 2^-24, k in [1, 2^24]: a float in (0, 1] like curand_uniform
    MQI_HD static float
    to_uniform(uint32_t w) {
        return (float) ((w >> 8) + 1u) * (1.0f / 16777216.0f);
    }

    // block `b` of this history's stream for another purpose (same key and
    // history, spot, scenario)
    MQI_HD u32x4
    block_of(uint32_t purpose, uint32_t b) const {
        u32x4 c;
        c.v[0] = b;
        c.v[1] = c1;
        c.v[2] = c2;
        c.v[3] = (c3 & ~0xffu) | purpose;
        return philox4x32_10(c, key0, key1);
    }

    // Standard normal variate in double (Box-Muller pair, second one cached).
    MQI_HD double
    gauss() {
        if (has_spare) {
            has_spare = 0;
            return spare;
        }
        const double u1 = (double) uniform();
        const double u2 = (double) uniform();
        const double r  = mqi_pm::pm_sqrt(-2.0 * mqi_pm::pm_log(u1));
        double       s, c;
        mqi_pm::pm_sincos(6.283185307179586 * u2, &s, &c);
        spare     = r * s;
        has_spare = 1;
        return r * c;
    }
};

}   // namespace mqi_rng

#endif
