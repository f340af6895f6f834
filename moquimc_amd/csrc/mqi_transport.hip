// Transport kernels for gfx950: fippel stepping, ionization with straggling,
// multiple Coulomb scattering, p-p elastic, p-O elastic/inelastic, the
// dose / LETd / statistics / Dij scorers, and the two launch forms:
//
//   batched  one persistent launch over every (scenario, spot) history of a
//            plan; lanes pull histories from a device counter with one
//            atomic per wavefront refill (ballot + popcount), so a lane whose
//            track ends takes the next secondary or history on the next
//            iteration while its neighbours keep stepping.
//   plain    one launch per spot, histories split statically over threads
//            (the reference's mc::transport_particles_patient launch,
//            mqi_transport.hpp:113-250, mqi_tps_env.hpp:1058-1129).
//
// Compiled twice: MQI_EXACT=1 (namespace mqi_exact, -ffp-contract=off,
// portable math: bit-identical to the CPU restatement) and MQI_EXACT=0
// (namespace mqi_fast, device math library, FMA contraction).
#include "mqi_device.hpp"
#include "mqi_kernels.h"

#define MQI_TABLE_QUALIFIER static __constant__
#include "mqi_tables.inc"

#if MQI_EXACT
#define MQI_NS mqi_exact
#else
#define MQI_NS mqi_fast
#endif

using namespace mqi_dev;

namespace MQI_NS
{

// ------------------------------------------------------------------ physics
// Per-lane context: LDS tables, grid view, RNG, the lane's secondary stack.
struct ctx {
    const float* T;   // LDS tables
    grid_view    g;
    rng_t        rng;
    float*       stack;   // HBM, stack_cap * MQI_STACK_FLOATS floats for this lane
    int          sp;
    int          cap;
    float        fwd[9];   // held by value: a pointer into the kernel
    float        trans[3]; // arguments would force them into private memory
    unsigned     drops;
    int          child;    // world child the track is in: 0 .. nbl - 1 beamline, nbl the patient
    int          nbl;      // beamline children
    const float* blxf;     // their transforms (fwd[9], inv[9], translation[3] each), HBM
};

// p_ionization_tabulated::cross_section (mqi_p_ionization.hpp:254-268): Ei=0.1, Ef=299.6, step 0.5
MQI_DI float cs_ion(const ctx& c, float Ek, float rho) {
    float cs = 0;
    if (Ek >= 0.1f && Ek <= 299.6f) {
        uint16_t idx0 = uint16_t((Ek - 0.1f) / 0.5f);
        uint16_t idx1 = idx0 + 1;
        float    x0   = 0.1f + idx0 * 0.5f;
        float    x1   = x0 + 0.5f;
        cs            = intpl(Ek, x0, x1, c.T[T_CS_ION + idx0], c.T[T_CS_ION + idx1]);
    }
    cs *= rho;
    return cs;
}
// pp / p-O tables: Ek_min 0.5, Ek_max 300, dEk 0.5 (mqi_pp_elastic.hpp:221-235 and twins)
MQI_DI float cs_tab(const ctx& c, int off, float Ek, float rho) {
    float cs = 0;
    if (Ek >= 0.5f && Ek <= 300.0f) {
        uint16_t idx0 = uint16_t((Ek - 0.5f) / 0.5f);
        uint16_t idx1 = idx0 + 1;
        float    x0   = 0.5f + idx0 * 0.5f;
        float    x1   = x0 + 0.5;
        cs            = intpl(Ek, x0, x1, c.T[off + idx0], c.T[off + idx1]);
    }
    cs *= rho;
    return cs;
}
// dEdx (mqi_p_ionization.hpp:271-286)
MQI_DI float dedx(const ctx& c, float Ek) {
    float pw = 0;
    if (Ek >= 0.1f && Ek <= 299.6f) {
        uint16_t idx0 = uint16_t((Ek - 0.1f) / 0.5f);
        uint16_t idx1 = idx0 + 1;
        float    x0   = 0.1f + idx0 * 0.5f;
        float    x1   = x0 + 0.5f;
        pw            = intpl(Ek, x0, x1, c.T[T_DEDX + idx0], c.T[T_DEDX + idx1]);
    } else if (Ek < 0.1f && Ek > 0) {
        pw = c.T[T_DEDX];
    }
    return -1.0 * pw;
}

// energy_loss + energy_straggling (mqi_p_ionization.hpp:297-345); the
// residual-range walk stops at index 0 where the reference's uint16_t
// counter would wrap past zero.
template<bool DBG>
MQI_DI float energy_loss(ctx& c, const rel_q& rel, float rho, float step_length, float sp) {
    const float* R   = c.T + T_RANGE;
    float        L_w = step_length * sp * rho / units::water_density;
    uint16_t     n   = uint16_t((rel.Ek - 0.1f) / 0.5f);
    float        x0  = 0.1f + n * 0.5f;
    float        x1  = x0 + 0.5f;
    if (x0 > rel.Ek) n -= 1;
    if (x1 < rel.Ek) n += 1;
    float r = intpl(rel.Ek, x0, x1, R[n], R[n + 1]);
    if (r < L_w) return rel.Ek;
    r -= L_w;
#if MQI_EXACT
    for (;;) {
        if (r >= R[n]) break;
        if (n == 0) break;
        --n;
        if (n == 0) break;
    }
#else
    // The walk above ends at the largest n in [1, n0] with R[n] <= r (0 if
    // none), R being non-decreasing.  Walking from the current bin takes one
    // LDS round trip per 0.5 MeV lost, the longest lane setting the wave's
    // count (8.3 per step on the 2000-spot CT plan); here the walk starts at
    // the bin the range slope of the current bin predicts, and goes up or
    // down from there to the same n.
    {
        const int n0 = n;
        const float slope = R[n0 + 1] - R[n0];   // mm per 0.5 MeV
        int         k     = slope > 0.0f ? (int) (L_w / slope) : 0;
        int         m     = n0 - (k < n0 ? k : n0);
        if (m >= 1 && R[m] > r) {
            do {
                --m;
            } while (m >= 1 && R[m] > r);
            if (m < 1) m = 0;
        } else {
            while (m < n0 && R[m + 1] <= r) ++m;
        }
        n = (uint16_t) m;
    }
#endif
    x0            = 0.1f + n * 0.5f;
    x1            = x0 + 0.5f;
    float dE_mean = rel.Ek - intpl(r, R[n], R[n + 1], x0, x1);
    float Te      = (rel.Te_max >= 0.08511) ? 0.08511 : rel.Te_max;
    float O_sq    = units::two_pi_re2_mc2_h2o * rho / units::water_density * L_w;
    O_sq *= Te / rel.beta_sq * (1.0 - 0.5 * rel.beta_sq);
    float ret = c.rng.step_normal(0, dE_mean, MQ_SQRTF(O_sq));
    if (ret < 0) ret *= -1.0;
    return ret;
}

// along_step (mqi_p_ionization.hpp:348-420)
template<bool DBG>
MQI_DI void along_step(ctx& c, trk& t, float len, float rho, float sp) {
    rel_q rel = relq(t.ke0);
    float dE  = energy_loss<DBG>(c, rel, rho, len, sp);
    float r   = 1.0;
    if (dE >= t.ke0) {
        r         = t.ke0 / dE;
        t.stopped = true;
    }
    float P     = momentum(rel);
    float X0    = radiation_length<DBG>(rho);
    float th_sq = ((13.9f / P) * (13.9f / P) / rel.beta_sq) * len / X0;
    float th    = MQ_SQRTF(th_sq);
    th          = c.rng.step_normal(1, 0, MQ_SQRTF(2.0f) * th);
    if (th < 0) th *= -1.0;
    float phi = 2.0 * 3.14159265358979323846 * c.rng.step_uniform(1);
    rotate_direction(t.d1, th, phi);
    t.dE += dE * r;
    update_pos(t, r * len);
    t.ke1 -= dE * r;
}

// secondary placement: pos -> fwd (pos - T) + T, dir -> fwd dir (e.g.
// mqi_pp_elastic.hpp:190-196), with the transform of the child the track is
// in.  In the patient (c.fwd, c.trans) at the push; a secondary made in a
// beamline child is stored unplaced with the child's index and placed with
// that child's transform when it is popped (the same operations, so the same
// bits), which keeps the beamline transforms out of the physics.
MQI_DI v3 place(const ctx& c, v3 p) {
    MQI_NO_CONTRACT
    v3 q = mk(p.x - c.trans[0], p.y - c.trans[1], p.z - c.trans[2]);
    q    = mv(c.fwd, q);
    return mk(q.x + c.trans[0], q.y + c.trans[1], q.z + c.trans[2]);
}
MQI_DI v3 turn(const ctx& c, v3 d) { return mv(c.fwd, d); }
// the same with a child's transform (fwd[9], inv[9], translation[3]) in HBM
MQI_DI v3 place_xf(const float* f, v3 p) {
    MQI_NO_CONTRACT
    v3 q = mv(f, mk(p.x - f[18], p.y - f[19], p.z - f[20]));
    return mk(q.x + f[18], q.y + f[19], q.z + f[20]);
}

// track_stack_t::push_secondary: dropped when the stack is full (mqi_track_stack.hpp:36-41).
// A pushed track keeps vtx0, vtx1.ke and dE; vtx1 position/direction are
// overwritten when it is popped, local_dE is always zero, status is CREATED.
MQI_DI void push(ctx& c, float ke0, v3 p0, v3 d0, float ke1, float dE, float tag) {
    if (c.sp < c.cap) {
        float* s = c.stack + MQI_STACK_FLOATS * c.sp;
        s[0]     = ke0;
        s[1]     = p0.x;
        s[2]     = p0.y;
        s[3]     = p0.z;
        s[4]     = d0.x;
        s[5]     = d0.y;
        s[6]     = d0.z;
        s[7]     = ke1;
        s[8]     = dE;
        s[9]     = tag;
        ++c.sp;
    } else {
        ++c.drops;
    }
}

// push a daughter made at (p, d) in the current child
MQI_DI void push_placed(ctx& c, float ke0, v3 p, v3 d, float ke1, float dE) {
    if (c.child < c.nbl)
        push(c, ke0, p, d, ke1, dE, (float) c.child);
    else
        push(c, ke0, place(c, p), turn(c, d), ke1, dE, -1.0f);
}

// p_ionization_tabulated::post_step: delta-ray (mqi_p_ionization.hpp:424-477)
template<bool DBG>
MQI_DI void post_ion(ctx& c, trk& t) {
    rel_q       rel   = relq(t.ke1);
    const float T_cut = DBG ? 0.08511 * 1.0 : 0.0815 * 1.0;
    float       Te, n;
    // acceptance probability is above 1/2 for any finite energy; the bound
    // only ends the loop for non-finite input (which the reference would spin on)
    for (int it = 0; it < 4096; ++it) {
        n  = it == 0 ? c.rng.int_uniform(1) : c.rng.uniform();
        Te = T_cut * rel.Te_max;
        Te /= ((1.0 - n) * rel.Te_max + n * T_cut);
        const float acc = it == 0 ? c.rng.int_uniform(2) : c.rng.uniform();
        if (acc < 1.0 - rel.beta_sq * Te / rel.Te_max + Te * Te / (2.0 * rel.Et_sq)) break;
    }
    if (DBG) {
        push_placed(c, 0.0f, t.p0, t.d0, 0.0f, Te);
    } else {
        t.dE += Te;
    }
    t.ke1 -= Te;
}

// pp_elastic_tabulated::post_step (mqi_pp_elastic.hpp:119-219)
MQI_DI void post_pp(ctx& c, trk& t) {
    rel_q rel       = relq(t.ke1);
    float min_value = 0.5f / rel.Ek;
    float u         = c.rng.uniform() * (1.0 - 2.0 * min_value) + min_value;
    float E1        = rel.Et;
    float dE        = rel.Ek * u;
    float E3        = (rel.Ek - dE) + units::Mp;
    float E4        = dE + units::Mp;
    float P1        = momentum(rel);
    float P3        = MQ_SQRTF(E3 * E3 - units::Mp_sq);
    float P4        = MQ_SQRTF(E4 * E4 - units::Mp_sq);
    float cos_th3   = E1 * E3 - units::Mp_sq - units::Mp * (E1 - E3);
    cos_th3 /= (P1 * P3);
    float cos_th34 = E3 * E4 - E1 * units::Mp;
    cos_th34 /= (P3 * P4);
    if (cos_th3 > 1.0)
        cos_th3 = 1.0;
    else if (cos_th3 < -1.0)
        cos_th3 = -1.0;
    if (cos_th34 > 1.0)
        cos_th34 = 1.0;
    else if (cos_th34 < -1.0)
        cos_th34 = -1.0;
    float th3 = MQ_ACOSF(cos_th3);
    float th4 = th3 - MQ_ACOSF(cos_th34);
    float phi = 2.0 * 3.14159265358979323846 * c.rng.uniform();
    t.ke1 -= dE;
    rotate_direction(t.d1, th3, phi);
    v3 dd = t.d1;
    rotate_direction(dd, th4, phi);
    push_placed(c, dE, t.p1, dd, dE, 0.0f);
}

// po_elastic::post_step (mqi_po_elastic.hpp:96-217); exponential not truncated (CPU semantics)
template<bool DBG>
MQI_DI void post_poe(ctx& c, trk& t) {
    rel_q rel = relq(t.ke1);
    if (rel.Ek <= 5.5f) {
        float dE = rel.Ek;
        if (DBG)
            push_placed(c, dE, t.p1, t.d1, 0.0f, dE);
        else
            t.ldE += dE;
        t.ke1 -= dE;
        t.stopped = true;
    } else {
        float Tp_avg = 0.65 * MQ_EXPD(-0.0013 * rel.Ek);
        Tp_avg -= 0.71 * MQ_EXPD(-0.0177 * rel.Ek);
        float Tp_max = (2.0 * units::Mo * rel.beta_sq * rel.gamma_sq);
        Tp_max /= (1.0 + 2.0 * rel.gamma * units::MoMp + units::MoMp_sq);
        (void) Tp_max;
        float avg     = (float) (1.0 / Tp_avg);
        float uu      = c.rng.uniform();
        float dE      = -MQ_LOGF(uu) / avg;
        float E1      = rel.Ek * (rel.Ek + 2.0 * units::Mp);
        float E3      = (rel.Ek - dE) * (rel.Ek - dE + 2.0 * units::Mp);
        float cos_th3 = (E1 + E3 - dE * (dE + 2.0 * units::Mo)) / 2.0 / MQ_SQRTF(E1 * E3);
        if (cos_th3 > 1.0) cos_th3 = 1.0;
        if (cos_th3 < -1.0) cos_th3 = -1.0;
        float th3 = MQ_ACOSF(cos_th3);
        float phi = 2.0 * 3.14159265358979323846 * c.rng.uniform();
        if (DBG)
            push_placed(c, dE, t.p0, t.d0, 0.0f, dE);
        else
            t.ldE += dE;
        t.ke1 -= dE;
        rotate_direction(t.d1, th3, phi);
    }
}

// po_inelastic_tabulated::post_step (mqi_po_inelastic.hpp:158-288)
template<bool DBG>
MQI_DI void post_poi(ctx& c, trk& t) {
    const float Ek = t.ke1;
    float       Eb = 5.0f;
    float       Er = Ek;
    const float E_mini = 2.0f, E_ratio = 0.65f;
    float       Prob_2nd, Prob_long, power;
    if (Ek <= 215 && Ek > 200) {
        Prob_2nd  = 0.78f;
        Prob_long = Prob_2nd + (1 - Prob_2nd) * 0.9;
        power     = 0.4f;
    } else if (Ek > 215) {
        Prob_2nd  = 0.78f;
        Prob_long = Prob_2nd + (1 - Prob_2nd) * 1.0;
        power     = 0.4f;
    } else if (Ek <= 200 && Ek > 150) {
        Prob_2nd  = 0.72f;
        Prob_long = Prob_2nd + (1 - Prob_2nd) * 0.83;
        power     = 0.45f;
    } else {
        Prob_2nd  = 0.7f;
        Prob_long = Prob_2nd + (1 - Prob_2nd) * 0.83;
        power     = 0.52f;
    }
    while ((Er - Eb) > E_mini) {
        Er -= Eb;
        float u  = c.rng.uniform();
        float dE = MQ_POWF(u, power) * (Er - E_mini) + E_mini;
        if (dE >= Er) dE = Er;
        Er -= dE;
        t.ke1 -= dE + Eb;
        float zeta = c.rng.uniform();
        if (zeta < Prob_2nd) {
            float cos_th = (2.0 * dE / Ek - 1.0) + 2.0 * (1 - dE / Ek) * c.rng.uniform();
            if (cos_th < -1) cos_th = -1;
            if (cos_th > 1) cos_th = 1;
            float th  = MQ_ACOSF(cos_th);
            float phi = 2.0 * 3.14159265358979323846 * c.rng.uniform();
            v3    dd  = t.d1;
            rotate_direction(dd, th, phi);
            push_placed(c, dE, t.p1, dd, dE, 0.0f);
        } else if (zeta < Prob_long) {
            // long-range energy leaves the track
        } else if (DBG) {
            push_placed(c, dE, t.p0, t.d0, 0.0f, dE);
        }
        Eb *= E_ratio;
    }
    t.dE += Er;
    t.ke1 -= Er;
    t.stopped = true;
}

// fippel_physics::stepping (mqi_fippel_physics.hpp:68-216).  The three
// branches' along_step calls are merged into one call with the branch's
// length; the random-number order is unchanged.  The stopping power ratio
// the reference evaluates at several points of a step (stepping,
// energy_loss, last_step, the dose hit) always has the arguments (rho,
// ke0 of the step): it is evaluated once and returned in sp for scoring.
template<bool DBG>
MQI_DI void stepping(ctx& c, trk& t, float rho, float d2b, float& sp) {
    sp = 0.0f;   // vacuum: the dose hit returns before reading it
    if (rho < 1.0e-7) {
        update_pos(t, d2b);
        return;
    } else if (rho > 99.9) {
        t.stopped = true;
        sp        = spr<DBG>(rho, t.ke0);
        return;
    }
    if (t.ke0 <= 0.5f) {
        if (t.ke0 < 0) t.ke0 = 0;
        sp = spr<DBG>(rho, t.ke0);
        t.dE += t.ke0;
        t.ke1 -= t.ke0;
        // last_step (mqi_p_ionization.hpp:481-490)
        rel_q rel = relq(t.ke0);
        float L_w = 0;
        if (t.dE > 0) L_w = -t.dE / dedx(c, rel.Ek);
        float step_length = L_w * units::water_density / (sp * rho);
        update_pos(t, step_length);
        t.stopped = true;
        return;
    }
    rel_q rel = relq(t.ke0);
    float cms = 1.0f;
    sp        = spr<DBG>(rho, rel.Ek);   // rel.Ek == t.ke0
    cms       = cms * sp * rho / units::water_density;
    float max_loss_energy = -1.0 * cms * dedx(c, rel.Ek);
    float cs1[4]  = { cs_ion(c, rel.Ek, rho), cs_tab(c, T_CS_PP, rel.Ek, rho), cs_tab(c, T_CS_POE, rel.Ek, rho),
                      cs_tab(c, T_CS_POI, rel.Ek, rho) };
    float cs1_sum = cs1[0] + cs1[1] + cs1[2] + cs1[3];
    rel_q rel_de  = relq(t.ke0 - max_loss_energy);
    float cs2[4]  = { cs_ion(c, rel_de.Ek, rho), cs_tab(c, T_CS_PP, rel_de.Ek, rho),
                      cs_tab(c, T_CS_POE, rel_de.Ek, rho), cs_tab(c, T_CS_POI, rel_de.Ek, rho) };
    float cs2_sum = cs2[0] + cs2[1] + cs2[2] + cs2[3];
    const bool first = cs1_sum >= cs2_sum;
    float cs_sum  = first ? cs1_sum : cs2_sum;
    float cs0 = first ? cs1[0] : cs2[0], csa = first ? cs1[1] : cs2[1], csb = first ? cs1[2] : cs2[2],
          csc = first ? cs1[3] : cs2[3];
    c.rng.begin_step();
    float prob       = c.rng.step_uniform(0);
    float mfp        = -1.0f * MQ_LOGF(prob) / cs_sum;
    float step_limit = cms * units::water_density / (sp * rho);
    int   br;
    float len;
    if (d2b < mfp && d2b < step_limit) {
        br  = 0;
        len = d2b;
    } else if ((mfp < d2b || fabsf(mfp - d2b) < 1e-3f) &&
               (mfp < step_limit || fabsf(mfp - step_limit) < 1e-3f)) {
        br  = 1;
        len = mfp;
    } else {
        br  = 2;
        len = step_limit;
    }
    along_step<DBG>(c, t, len, rho, sp);
    if (br != 1) return;
    if (t.ke1 <= 0.5f) return;
    c.rng.begin_interaction();
    float u = cs_sum * c.rng.int_uniform(0);
    t.d1    = t.d0;
    if (u < cs0) {
        post_ion<DBG>(c, t);
    } else if (u < (cs0 + csa)) {
        post_pp(c, t);
    } else if (u < (cs0 + csa + csb)) {
        post_poe<DBG>(c, t);
    } else if (u < (cs0 + csa + csb + csc)) {
        post_poi<DBG>(c, t);
    }
}

// ------------------------------------------------------------------ scoring
// hit functions (mqi_scorer_energy_deposit.hpp:43-139)
template<bool DBG>
MQI_DI double dose_hit(const trk& t, float density, float volume, float sp) {
    if (density < 1.0e-7) return 0.0;
#if MQI_EXACT
    return (t.dE + t.ldE) * 1.60218e-10 / (volume * density * sp);
#else
    // production: the per-step value in float (it is accumulated in FP64)
    return (double) ((t.dE + t.ldE) * 1.60218e-10f / (volume * density * sp));
#endif
}
// LETt_weight1 / LETt_weight2 (mqi_scorer_energy_deposit.hpp:141-177)
MQI_DI double lett_hit(const trk& t, float density, bool numer) {
    density *= 1000.0;
    double length = (t.p1.x - t.p0.x) * (t.p1.x - t.p0.x);
    length += (t.p1.y - t.p0.y) * (t.p1.y - t.p0.y);
    length += (t.p1.z - t.p0.z) * (t.p1.z - t.p0.z);
    length = MQ_SQRTD(length);
    if (length <= 0) return 0.0;
    if (!numer) return length;
    double let = t.dE / length / density;
    return length * let;
}
// energy_deposit (:14-20); dose_to_medium of primaries in J/kg (:80-92)
MQI_DI double edep_hit(const trk& t) { return t.dE + t.ldE; }
MQI_DI double dm_hit(const trk& t, float density, float volume) {
    return t.primary ? t.dE * 1.60218e-13 * 1000.0 / (volume * density) : 0.0;
}

MQI_DI double letd_hit(const trk& t, float density, bool numer) {
    density *= 1000.0;
    double length = (t.p1.x - t.p0.x) * (t.p1.x - t.p0.x);
    length += (t.p1.y - t.p0.y) * (t.p1.y - t.p0.y);
    length += (t.p1.z - t.p0.z) * (t.p1.z - t.p0.z);
    length = MQ_SQRTD(length);
    if (length <= 0) return 0.0;
    double let = t.dE / length / density;
    if (let >= 25.0) return 0;
    return numer ? t.dE * let : t.dE * 1.0;
}

// Dij: per-spot open-addressing table of (voxel, dose) in HBM.  Keys are
// claimed with one 32-bit CAS, values added with FP64 atomics.  Linear
// probing stops after MQI_DIJ_MAX_PROBE slots (a table that full is an
// overflow: the run reports it instead of crawling through a full table).
#define MQI_DIJ_MAX_PROBE 4096
MQI_DI void dij_insert(const transport_params& P, uint64_t item, uint32_t voxel, double value) {
    if (value <= 0) return;
    const uint64_t base = P.dij_off[item];
    const uint64_t cap  = P.dij_off[item + 1] - base;
    uint32_t*      keys = P.dij_keys + base;
    double*        vals = P.dij_vals + base;
    uint32_t       h    = voxel * 2654435761u;
    uint64_t       slot = ((uint64_t) h * cap) >> 32;
    const uint64_t n_probe = cap < MQI_DIJ_MAX_PROBE ? cap : MQI_DIJ_MAX_PROBE;
    // A key changes once, from empty to its voxel, so a plain load that
    // shows a voxel is final: the slot is ours, or taken by another voxel.
    // Only a load that shows it empty (possibly a stale cached line) needs
    // the compare-and-swap, which is performed beyond the caches.
    for (uint64_t probe = 0; probe < n_probe; ++probe) {
        uint32_t k = keys[slot];
        if (k == 0xffffffffu) k = atomicCAS(keys + slot, 0xffffffffu, voxel);
        if (k == 0xffffffffu || k == voxel) {
            atomicAdd(vals + slot, value);
            return;
        }
        slot = (slot + 1 == cap) ? 0 : slot + 1;
    }
    atomicOr(P.status, MQI_STATUS_DIJ_FULL);
}

// Deposits of consecutive steps into the same voxel (and scenario, and Dij
// table) are summed per lane and added to the scorer arrays, or inserted
// into the Dij table, once, when the lane moves on (each-scorer-once scoring
// only; the sums differ from per-step atomics only in FP64 summation order).
// Scorer slots beyond MQI_DEFER are scored per step.  The sums live in the lane's LDS slots (v[s * blockDim.x]), not
// in registers: held across the loop they pushed the kernel into scratch,
// and every scratch reload waits on the one counter the atomics in flight
// also hold.
struct pending {
    int64_t  cnb;    // voxel of the pending sums, -1: none
    uint64_t off;    // scenario offset of the pending sums
    uint32_t item;   // Dij table (scenario, spot) of the pending sums
};

MQI_DI void flush(const transport_params& P, pending& A, double* v) {
    if (A.cnb >= 0) {
        // accumulated slots are positive, +inf or NaN (the per-step path adds
        // every value that is not <= 0, NaN included); untouched slots are 0
#pragma unroll
        for (int s = 0; s < MQI_DEFER; ++s) {
            const double x = v[s * blockDim.x];
            if (x == 0.0) continue;
            if (P.kinds[s] == MQI_KIND_DIJ)
                dij_insert(P, A.item, (uint32_t) A.cnb, x);
#if !MQI_EXACT
            else if (P.pair && s < P.group_ns)
                atomicAdd(P.pair_out + (uint64_t) P.pair * (A.off + (uint64_t) A.cnb) + s, x);
#endif
            else
                atomicAdd(P.out[s] + A.off + (uint64_t) A.cnb, x);
        }
    }
    A.cnb = -1;
}

#if !MQI_EXACT
// The flush of the step with paired scoring, run by every lane that steps
// (f: this lane leaves its voxel).  Lanes 2m and 2m+1 exchange their sums:
// the first atomic instruction adds both of lane 2m's sums (scorer 0 from
// lane 2m, scorer 1 from lane 2m+1), the second both of lane 2m+1's, so the
// two adds of a voxel land in one line of the interleaved array in one
// instruction.  A lane whose partner does not step adds both of its own.
MQI_DI void flush_pair(const transport_params& P, pending& A, double* v, bool f) {
    const bool     has = f && A.cnb >= 0;
    const double   a   = has ? v[0] : 0.0;
    const double   b   = has ? v[blockDim.x] : 0.0;
    const uint64_t k   = has ? A.off + (uint64_t) A.cnb : 0;
    const int      me  = __lane_id();
    const bool     pact = (__ballot(1) >> (me ^ 1)) & 1ull;
    const uint64_t kp  = __shfl_xor((unsigned long long) k, 1);
    const double   ap  = __shfl_xor(a, 1);
    const double   bp  = __shfl_xor(b, 1);
    const bool     odd = me & 1;
    uint64_t       eA = 2 * k, eB = 2 * k + 1;   // partner stepping: overwritten
    double         xA = a, xB = b;
    if (pact) {
        eA = odd ? 2 * kp + 1 : 2 * k;
        xA = odd ? bp : a;
        eB = odd ? 2 * k + 1 : 2 * kp;
        xB = odd ? b : ap;
    }
    if (xA != 0.0) atomicAdd(P.pair_out + eA, xA);
    if (xB != 0.0) atomicAdd(P.pair_out + eB, xB);
    if (has) {
#pragma unroll
        for (int s = 2; s < MQI_DEFER; ++s) {
            const double x = v[s * blockDim.x];
            if (x != 0.0) atomicAdd(P.out[s] + A.off + (uint64_t) A.cnb, x);
        }
        A.cnb = -1;
    }
}

// The same for three or four grouped scorers (interleaved four doubles per
// element): in a group of four stepping lanes, instruction j adds the sums of
// lane 4m+j's voxel, lane 4m+q adding that lane's sum of scorer q, so all of
// a voxel's sums go out as one line request.  Lanes of a group with a lane
// that does not step add their own sums.
MQI_DI void flush_quad(const transport_params& P, pending& A, double* v, bool f) {
    const bool     has = f && A.cnb >= 0;
    const double   x0  = has ? v[0] : 0.0;
    const double   x1  = has ? v[blockDim.x] : 0.0;
    const double   x2  = has ? v[2 * blockDim.x] : 0.0;
    const double   x3  = (has && P.group_ns > 3) ? v[3 * blockDim.x] : 0.0;
    const uint64_t k   = has ? A.off + (uint64_t) A.cnb : 0;
    const int      me  = __lane_id();
    const int      q   = me & 3;
    const int      g0  = me & ~3;
    if (((__ballot(1) >> g0) & 0xfull) == 0xfull) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint64_t kj = __shfl((unsigned long long) k, g0 + j);
            const double   y0 = __shfl(x0, g0 + j);
            const double   y1 = __shfl(x1, g0 + j);
            const double   y2 = __shfl(x2, g0 + j);
            const double   y3 = __shfl(x3, g0 + j);
            const double   y  = q == 0 ? y0 : (q == 1 ? y1 : (q == 2 ? y2 : y3));
            if (y != 0.0) atomicAdd(P.pair_out + 4 * kj + q, y);
        }
    } else if (has) {
        if (x0 != 0.0) atomicAdd(P.pair_out + 4 * k, x0);
        if (x1 != 0.0) atomicAdd(P.pair_out + 4 * k + 1, x1);
        if (x2 != 0.0) atomicAdd(P.pair_out + 4 * k + 2, x2);
        if (x3 != 0.0) atomicAdd(P.pair_out + 4 * k + 3, x3);
    }
    if (has) A.cnb = -1;
}
#endif

template<bool DBG>
MQI_DI void score(const transport_params& P, const trk& t, uint64_t cnb, float density, float volume, float sp,
                  uint64_t dij_item, uint64_t scen_off, pending& A, double* pv) {
    const int     n   = P.n_scorers;
    const int32_t roi = (int32_t) (uint32_t) cnb;   // roi_t DIRECT
    // regular scorers score through roi_->idx(cnb) > 0, the statistics pair
    // through get_mask_idx(cnb) >= 0 (mqi_transport.hpp:342-364); a CONTOUR
    // ROI's answer per voxel is in roi_flags
    const uint32_t fl      = P.roi_flags ? (uint32_t) P.roi_flags[cnb] : (roi > 0 ? 3u : 2u);
    const bool     in_reg  = fl & 1u;
    const bool     in_stat = fl & 2u;
    // Both loops run over the fixed maximum with constant indices so that the
    // scorer table stays in kernel arguments / registers (a runtime index
    // into it would move the whole argument block to private memory).
    const double d = dose_hit<DBG>(t, density, volume, sp);
    double       v[MQI_MAX_SCORERS];
#pragma unroll
    for (int s = 0; s < MQI_MAX_SCORERS; ++s) {
        v[s] = 0.0;
        if (s >= n) continue;
        const int k = P.kinds[s];
        if (k == MQI_KIND_LETD_NUM)
            v[s] = letd_hit(t, density, true);
        else if (k == MQI_KIND_LETD_DEN)
            v[s] = letd_hit(t, density, false);
        else if (k == MQI_KIND_LETT_NUM)
            v[s] = lett_hit(t, density, true);
        else if (k == MQI_KIND_LETT_DEN)
            v[s] = lett_hit(t, density, false);
        else if (k == MQI_KIND_EDEP)
            v[s] = edep_hit(t);
        else if (k == MQI_KIND_DM)
            v[s] = dm_hit(t, density, volume);
        else
            v[s] = (k == MQI_KIND_STAT_DOSE2) ? d * d : d;
    }
    if (P.loops == MQI_LOOPS_ONCE) {
        if ((int64_t) cnb != A.cnb || scen_off != A.off || (uint32_t) dij_item != A.item) {
            flush(P, A, pv);
            A.cnb  = (int64_t) cnb;
            A.off  = scen_off;
            A.item = (uint32_t) dij_item;
#pragma unroll
            for (int s = 0; s < MQI_DEFER; ++s) pv[s * blockDim.x] = 0.0;
        }
#pragma unroll
        for (int s = 0; s < MQI_MAX_SCORERS; ++s) {
            if (s >= n) continue;
            const int  k    = P.kinds[s];
            const bool stat = (k == MQI_KIND_STAT_DOSE || k == MQI_KIND_STAT_DOSE2);
            if (!(stat ? in_stat : in_reg) || v[s] <= 0) continue;
            if (s < MQI_DEFER)
                pv[s * blockDim.x] += v[s];
            else if (k == MQI_KIND_DIJ)
                dij_insert(P, dij_item, (uint32_t) cnb, v[s]);
            else
                atomicAdd(P.out[s] + scen_off + cnb, v[s]);
        }
        return;
    }
#pragma unroll
    for (int s = 0; s < MQI_MAX_SCORERS; ++s) {
        if (s >= n) continue;
        const int  k    = P.kinds[s];
        const bool stat = (k == MQI_KIND_STAT_DOSE || k == MQI_KIND_STAT_DOSE2);
        int        reps = 0;
        if (P.loops == MQI_LOOPS_ONCE) {
            reps = (stat ? in_stat : in_reg) ? 1 : 0;
        } else if (P.loops == MQI_LOOPS_PLAIN) {
            // mqi_transport.hpp:204-225: scorers 0..n-3 in both loops
            reps = in_reg ? ((s < n - 2) ? 2 : 1) : 0;
        } else {
            // mqi_transport.hpp:343-364: last two through get_mask_idx >= 0
            if (s < n - 2)
                reps = in_reg ? 1 : 0;
            else if (n >= 2)
                reps = in_stat ? 1 : 0;
        }
        if (v[s] <= 0) continue;
        if (k == MQI_KIND_DIJ) {
            for (int r = 0; r < reps; ++r) dij_insert(P, dij_item, (uint32_t) cnb, v[s]);
            continue;
        }
        for (int r = 0; r < reps; ++r) atomicAdd(P.out[s] + scen_off + cnb, v[s]);
    }
}

// ------------------------------------------------------------------ lane machine
struct lane {
    trk      t;
    bool     active;
    uint32_t item;       // work item (spot of a scenario) of the current history
    uint64_t hist;       // history index within the item
    uint64_t g;          // history index in the flattened list
    int      nsteps;     // trace mode: steps logged for this history
    float    scale;      // scenario density scale
    uint64_t scen_off;   // scenario offset into the output arrays
    uint64_t dij_item;   // Dij table of (scenario, spot): scenario * n_spots + spot
    int      tsteps;     // steps of the current track
    pending  acc;        // deferred deposits of the current voxel
    bool     reenter;    // left a beamline child: enter the next child
};

// pop + enter the child grid: transport loop head (mqi_transport.hpp:158-191)
MQI_DI bool enter(const transport_params& P, ctx& c, trk& t) {
    MQI_NO_CONTRACT
    v3 pp = mk(t.p0.x - P.trans[0], t.p0.y - P.trans[1], t.p0.z - P.trans[2]);
    t.p0  = mv(P.inv, pp);
    t.d0  = mv(P.inv, t.d0);
    normalize(t.d0);
    t.p1 = t.p0;
    t.d1 = t.d0;
    int cx, cy, cz;
    index3(c.g, t.p0, t.d0, cx, cy, cz);
    if (!is_valid(c.g, cx, cy, cz)) {
        int   ox, oy, oz;
        float dist = intersect_out(c.g, t.p0, t.d0, ox, oy, oz);
        t.its      = dist;
        if (dist < 0) return false;
        update_pos(t, dist);
        move(t);
        index3(c.g, t.p0, t.d0, t.cx, t.cy, t.cz);
    } else {
        t.its = 0.0f;
        t.cx  = cx;
        t.cy  = cy;
        t.cz  = cz;
    }
    return is_valid(c.g, t.cx, t.cy, t.cz) && !t.stopped;
}

// ------------------------------------------------------------------ beamline children
// The world's children are transported in order, the patient last
// (mqi_transport.hpp:158-240): the track goes into the child's frame, to its
// first voxel, steps until it leaves or stops, and back to the world frame
// (also when it misses the child) before the next child.  Beamline children
// have no scorers and no scenario shift or density scale; their edges sit in
// LDS after the patient's.
MQI_DI grid_view bl_view(const transport_params& P, const float* shb, int k) {
    const int* m = P.bl_meta + 8 * k;
    grid_view  g;
    g.nx = m[0];
    g.ny = m[1];
    g.nz = m[2];
    g.xe = shb + m[3];
    g.ye = g.xe + g.nx + 1;
    g.ze = g.ye + g.ny + 1;
    g.sx = g.sy = g.sz = 0.0f;
    return g;
}

// into child k's frame and to its first voxel (as enter() for the patient)
MQI_DI bool enter_bl(const transport_params& P, const grid_view& g, int k, trk& t) {
    MQI_NO_CONTRACT
    const float* f = P.bl_xf + 21 * k;
    v3           pp = mk(t.p0.x - f[18], t.p0.y - f[19], t.p0.z - f[20]);
    t.p0            = mv(f + 9, pp);
    t.d0            = mv(f + 9, t.d0);
    normalize(t.d0);
    t.p1 = t.p0;
    t.d1 = t.d0;
    int cx, cy, cz;
    index3(g, t.p0, t.d0, cx, cy, cz);
    if (!is_valid(g, cx, cy, cz)) {
        int   ox, oy, oz;
        float dist = intersect_out(g, t.p0, t.d0, ox, oy, oz);
        t.its      = dist;
        if (dist < 0) return false;
        update_pos(t, dist);
        move(t);
        index3(g, t.p0, t.d0, t.cx, t.cy, t.cz);
    } else {
        t.its = 0.0f;
        t.cx  = cx;
        t.cy  = cy;
        t.cz  = cz;
    }
    return is_valid(g, t.cx, t.cy, t.cz) && !t.stopped;
}

// back from child k's frame to the world: pos -> fwd pos + T, dir -> fwd dir
MQI_DI void leave_bl(const transport_params& P, int k, trk& t) {
    MQI_NO_CONTRACT
    const float* f = P.bl_xf + 21 * k;
    v3           q = mv(f, t.p0);
    t.p0           = mk(q.x + f[18], q.y + f[19], q.z + f[20]);
    t.d0           = mv(f, t.d0);
    t.p1 = t.p0;
    t.d1 = t.d0;
}

// the track (in the world frame) enters the first child it reaches from
// c.child on; false when it reaches none
MQI_DI bool enter_chain(const transport_params& P, ctx& c, trk& t, const float* shb) {
    while (c.child < c.nbl) {
        const grid_view g = bl_view(P, shb, c.child);
        if (enter_bl(P, g, c.child, t)) return true;
        leave_bl(P, c.child, t);
        ++c.child;
    }
    return enter(P, c, t);
}

MQI_DI void pop(ctx& c, trk& t) {
    --c.sp;
    const float* s = c.stack + MQI_STACK_FLOATS * c.sp;
    t.ke0          = s[0];
    t.p0           = mk(s[1], s[2], s[3]);
    t.d0           = mk(s[4], s[5], s[6]);
    if (s[9] >= 0.0f) {   // made in beamline child s[9]: place it now
        const float* f = c.blxf + 21 * (int) s[9];
        t.p0           = place_xf(f, t.p0);
        t.d0           = mv(f, t.d0);
    }
    t.ke1          = s[7];
    t.dE           = s[8];
    t.ldE          = 0.0f;
    t.stopped      = false;
    t.primary      = false;
}

// Vertex of history h of item (scenario, spot): phsp_6d + norm_1d/const_1d
// beamlet, then rotation/translation (mqi_beamlet.hpp:81-90, mqi_phsp6d.hpp:62-79).
MQI_DI void sample_vertex(const transport_params& P, const mqi_spot& sp, uint32_t spot_id, uint32_t scenario,
                          uint64_t h, trk& t) {
    rng_t r;
    r.s.init(P.seed, (uint32_t) h, spot_id, scenario, mqi_rng::STREAM_VERTEX);
    float Ux, Vx, Uy, Vy, Uz;
    if (sp.phsp == MQI_PHSP_UNIFORM) {   // U(-1, 1) draws (phsp_6d_uniform)
        Ux = 2.0f * r.uniform() - 1.0f;
        Vx = 2.0f * r.uniform() - 1.0f;
        Uy = 2.0f * r.uniform() - 1.0f;
        Vy = 2.0f * r.uniform() - 1.0f;
        Uz = 2.0f * r.uniform() - 1.0f;
    } else {
        Ux = r.gauss();
        Vx = r.gauss();
        Uy = r.gauss();
        Vy = r.gauss();
        Uz = r.gauss();
    }
    float ke = sp.energy;
    if (!sp.energy_const) {
        float z = r.gauss();
        ke      = z * sp.sigma_energy + sp.energy;
    }
    float ph[6];
    for (int k = 0; k < 6; ++k) ph[k] = sp.mean[k];
    ph[0] += sp.sigma[0] * Ux;
    ph[1] += sp.sigma[1] * Uy;
    ph[2] += sp.sigma[2] * Uz;
    const float rho0 = sp.corr[0], rho1 = sp.corr[1];
    ph[3] += sp.sigma[3] * (rho0 * Ux + Vx * MQ_SQRTD(1.0 - rho0 * rho0));
    ph[4] += sp.sigma[4] * (rho1 * Uy + Vy * MQ_SQRTD(1.0 - rho1 * rho1));
    ph[5]  = -1.0 * MQ_SQRTD(1.0 - ph[3] * ph[3] - ph[4] * ph[4]);
    v3 pos = mv(sp.rotation, mk(ph[0], ph[1], ph[2]));
    t.p0   = mk(pos.x + sp.translation[0], pos.y + sp.translation[1], pos.z + sp.translation[2]);
    t.d0   = mv(sp.rotation, mk(ph[3], ph[4], ph[5]));
    t.ke0  = ke;
    t.ke1  = ke;
    t.p1   = t.p0;
    t.d1   = t.d0;
    t.dE   = 0;
    t.ldE  = 0;
    t.stopped = false;
    t.primary = true;
}

// history g of the flattened work list -> (item, history-in-item)
MQI_DI void locate(const transport_params& P, uint64_t g, uint32_t& item, uint64_t& h) {
    int lo = 0, hi = P.n_items - 1;   // first item with item_end > g
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (P.item_end[mid] > g)
            hi = mid;
        else
            lo = mid + 1;
    }
    item = (uint32_t) lo;
    h    = g - (lo ? P.item_end[lo - 1] : 0);
}

MQI_DI void trace_record(const transport_params& P, lane& L, uint64_t cnb) {
    // per-history slots: history index g of the flattened list
    const uint64_t g = L.g;
    if (L.nsteps < P.trace_cap) {
        const uint64_t k = g * P.trace_cap + L.nsteps;
        P.trace_cnb[k]   = (int64_t) cnb;
        float* r         = P.trace_rec + 12 * k;
        r[0]             = L.t.its;
        r[1]             = L.t.p0.x;
        r[2]             = L.t.p0.y;
        r[3]             = L.t.p0.z;
        r[4]             = L.t.d0.x;
        r[5]             = L.t.d0.y;
        r[6]             = L.t.d0.z;
        r[7]             = L.t.p1.x;
        r[8]             = L.t.p1.y;
        r[9]             = L.t.p1.z;
        r[10]            = L.t.ke0;
        r[11]            = L.t.dE;
    }
    L.nsteps++;
    P.trace_n[g] = L.nsteps;
}

// one voxel-loop iteration (mqi_transport.hpp:192-233), in the patient or
// in a beamline child (no scoring there; leaving it hands the track to the
// next child).  One stepping call site serves both.
template<bool DBG, bool TRACE>
MQI_DI void step(const transport_params& P, ctx& c, lane& L, const float* shb, double* pv) {
    trk&       t  = L.t;
    const bool bl = c.child < c.nbl;
    uint64_t   cnb;
    float      rho;
    if (bl) {
        // the beamline child's grid view is rebuilt after stepping rather
        // than kept live across it (register pressure of the patient path)
        const grid_view gb = bl_view(P, shb, c.child);
        const int*     m   = P.bl_meta + 8 * c.child;
        const uint64_t off = (uint64_t) (uint32_t) m[4] | ((uint64_t) (uint32_t) m[5] << 32);
        cnb                = (uint64_t) (int32_t) (t.cz * gb.nx * gb.ny + t.cy * gb.nx + t.cx);
        t.its              = intersect_in(gb, t.p0, t.d0, t.cx, t.cy, t.cz);
        rho                = P.bl_density[off + cnb];
    } else {
        cnb              = (uint64_t) (int32_t) (t.cz * c.g.nx * c.g.ny + t.cy * c.g.nx + t.cx);
        const float dens = P.density[cnb];
        t.its            = intersect_in(c.g, t.p0, t.d0, t.cx, t.cy, t.cz);
        rho              = dens * L.scale;
        // The deposits gathered in the voxel the lane has left go out here,
        // after the wait for the density and ahead of the physics of the
        // step: one counter tracks loads, stores and atomics alike, so atomics
        // issued just before the next wait on a load would add their round
        // trip to every step.
        const bool leave = P.loops == MQI_LOOPS_ONCE && ((int64_t) cnb != L.acc.cnb || L.scen_off != L.acc.off ||
                                                         (uint32_t) L.dij_item != L.acc.item);
#if !MQI_EXACT
        if (P.pair == 2)
            flush_pair(P, L.acc, pv, leave);
        else if (P.pair == 4)
            flush_quad(P, L.acc, pv, leave);
        else
#endif
        if (leave)
            flush(P, L.acc, pv);
    }
    float sp;   // stopping power ratio of the step
    stepping<DBG>(c, t, rho, t.its, sp);
    if (bl) {
        const grid_view gb    = bl_view(P, shb, c.child);
        bool            leave = t.its < 0;
        if (!leave && !t.stopped) {
            step_axis(t.p1.x, t.d1.x, gb.xe, 0.0f, t.cx);
            step_axis(t.p1.y, t.d1.y, gb.ye, 0.0f, t.cy);
            step_axis(t.p1.z, t.d1.z, gb.ze, 0.0f, t.cz);
            move(t);
            leave = !is_valid(gb, t.cx, t.cy, t.cz);
        }
        if (t.stopped) {
            // a stopped track does nothing in the children after this one
            L.active = false;
        } else if (leave) {
            leave_bl(P, c.child, t);
            ++c.child;
            L.reenter = true;
        }
        if (++L.tsteps > MQI_MAX_TRACK_STEPS && L.active) {
            L.active = false;
            atomicOr(P.status, MQI_STATUS_TRACK_CAP);
        }
        return;
    }
    if (t.its < 0) {
        L.active = false;
        return;
    }
    if (TRACE) {
        trace_record(P, L, cnb);
    } else {
        // get_volume(cnb) (mqi_grid3d.hpp:471-487) goes through cnb2ijk; the
        // lane's cell indices are that (i, j, k) already
        const int i = t.cx, j = t.cy, k = t.cz;
        float     vol = c.g.X(i + 1) - c.g.X(i);
        vol *= c.g.Y(j + 1) - c.g.Y(j);
        vol *= c.g.Z(k + 1) - c.g.Z(k);
        score<DBG>(P, t, cnb, rho, vol, sp, L.dij_item, L.scen_off, L.acc, pv);
    }
    if (!t.stopped) {
        step_axis(t.p1.x, t.d1.x, c.g.xe, c.g.sx, t.cx);
        step_axis(t.p1.y, t.d1.y, c.g.ye, c.g.sy, t.cy);
        step_axis(t.p1.z, t.d1.z, c.g.ze, c.g.sz, t.cz);
        move(t);
    }
    L.active = is_valid(c.g, t.cx, t.cy, t.cz) && !t.stopped;
    // a finite track takes a few thousand steps at most; the cap only ends
    // tracks whose state went non-finite
    if (++L.tsteps > MQI_MAX_TRACK_STEPS && L.active) {
        L.active = false;
        atomicOr(P.status, MQI_STATUS_TRACK_CAP);
    }
}

// LDS staging of the tables and edges
MQI_DI void stage(const transport_params& P, float* sh) {
    for (int i = threadIdx.x; i < 600; i += blockDim.x) {
        sh[T_CS_ION + i] = kCsPIon[i];
        sh[T_DEDX + i]   = kDedxRestricted[i];
        sh[T_RANGE + i]  = kRangeSteps[i];
        sh[T_CS_PP + i]  = kCsPPElastic[i];
        sh[T_CS_POE + i] = kCsPOElastic[i];
        sh[T_CS_POI + i] = kCsPOInelastic[i];
    }
    float* xe = sh + T_TABLES;
    float* ye = xe + P.nx + 1;
    float* ze = ye + P.ny + 1;
    for (int i = threadIdx.x; i <= P.nx; i += blockDim.x) xe[i] = P.xe[i];
    for (int i = threadIdx.x; i <= P.ny; i += blockDim.x) ye[i] = P.ye[i];
    for (int i = threadIdx.x; i <= P.nz; i += blockDim.x) ze[i] = P.ze[i];
    float* blx = ze + P.nz + 1;
    for (int i = threadIdx.x; i < P.bl_n_edges; i += blockDim.x) blx[i] = P.bl_edges[i];
    __syncthreads();
}

// The transport kernel.  DYNAMIC: histories from the device counter
// (batched persistent form); otherwise each thread takes a contiguous slice
// [start, start+len) of the work list (plain form; start_and_length,
// mqi_utils.hpp:138-147).
// W > 0 asks the register allocator for at least W resident waves per SIMD
// (more memory latency hidden, at the price of spills past the budget).
template<bool DBG, bool DYNAMIC, bool TRACE, int W>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W > 0 ? W : 1)))
k_transport(transport_params P) {
    extern __shared__ float sh[];
    stage(P, sh);
    ctx c;
    c.T     = sh;
    c.g.nx  = P.nx;
    c.g.ny  = P.ny;
    c.g.nz  = P.nz;
    c.g.xe  = sh + T_TABLES;
    c.g.ye  = c.g.xe + P.nx + 1;
    c.g.ze  = c.g.ye + P.ny + 1;
#pragma unroll
    for (int i = 0; i < 9; ++i) c.fwd[i] = P.fwd[i];
#pragma unroll
    for (int i = 0; i < 3; ++i) c.trans[i] = P.trans[i];
    c.cap   = P.stack_cap;
    c.sp    = 0;
    c.drops = 0;
    c.child = 0;
    c.nbl   = P.n_bl;
    c.blxf  = P.bl_xf;
    const float* shb = c.g.ze + P.nz + 1;   // beamline edges in LDS
    // the lane's pending deposit sums (MQI_DEFER doubles, blockDim.x apart),
    // after the beamline edges at an 8-byte boundary
    double* pv = reinterpret_cast<double*>(sh + mqi_pend_offset(P.nx, P.ny, P.nz, P.bl_n_edges)) + threadIdx.x;
    const uint64_t tid      = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t nthreads = (uint64_t) gridDim.x * blockDim.x;
    c.stack                 = P.stack + tid * (uint64_t) P.stack_cap * MQI_STACK_FLOATS;

    // plain: the thread's start_and_length slice; dynamic: the lane's current
    // chunk of P.fetch_chunk consecutive histories from the device counter
    uint64_t s_next = 0, s_end = 0;
    if (!DYNAMIC) {   // start_and_length over the flattened list
        const uint64_t q = P.total / nthreads, r = P.total % nthreads;
        s_next           = q * tid + (tid < r ? tid : r);
        s_end            = s_next + q + (tid < r ? 1 : 0);
    }
    lane     L;
    L.active = false;
    L.item   = 0;
    L.hist   = 0;
    L.nsteps   = 0;
    L.scale    = 1.0f;
    L.scen_off = 0;
    L.dij_item = 0;
    L.acc.cnb  = -1;
    L.acc.off  = 0;
    L.acc.item = 0;
    L.reenter  = false;
    c.g.sx = c.g.sy = c.g.sz = 0.0f;
    unsigned long long done = 0, nstep = 0;
    uint32_t           scen = 0;
    const uint64_t     hoff = P.hist_offset ? *P.hist_offset : 0;
    int                qx   = 0;   // work range this wave fetches from (dynamic form)
    if (DYNAMIC) {
        unsigned xcc;
        asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(xcc));
        qx = (int) (xcc & (MQI_QUEUES - 1));
    }

    for (;;) {
        // A lane whose history and secondaries are done takes a new history.
        // With refill_lanes > 1 (dynamic form) it first waits until that
        // many lanes of its wave need one, so the wave runs the new-history
        // path (sampling, work-list search, grid entry) once for all of
        // them; with no lane left to step it refills at once.
        const bool need   = !L.active && c.sp == 0;
        bool       refill = need;
        if (DYNAMIC && P.refill_lanes > 1) {
            const unsigned long long nm   = __ballot(need);
            const unsigned long long busy = __ballot(!need);
            refill = need && (__popcll(nm) >= (unsigned) P.refill_lanes || busy == 0);
        }
        if (need && !refill) continue;
        if (refill) {
            uint64_t g;
            if (DYNAMIC) {
                if (s_next >= s_end) {
                    // Wave-aggregated fetch of fetch_chunk histories per dry
                    // lane from range qx (initially this XCD's own range; a
                    // range that runs dry sends the wave to the next one).
                    const uint64_t chunk = (uint64_t) P.fetch_chunk;
                    const int      me    = __lane_id();
                    bool           got   = false;
                    for (int r = 0; r < MQI_QUEUES; ++r) {
                        const unsigned long long m = __ballot(!got);
                        if (m == 0) break;
                        const int          lead = __builtin_ctzll(m);
                        const int          rank = __popcll(m & ((1ull << me) - 1ull));
                        // the leader's range for every lane (lanes that were
                        // busy during earlier fetches may hold a stale qx)
                        const int          q    = __shfl(qx, lead);
                        unsigned long long base = 0;
                        if (me == lead)
                            base = atomicAdd(P.queue + q * MQI_QUEUE_STRIDE, (unsigned long long) __popcll(m) * chunk);
                        base               = __shfl(base, lead);
                        const uint64_t end = mqi_range_begin(P.total, q + 1);
                        if (!got) {
                            const uint64_t s0 = base + rank * chunk;
                            if (s0 < end) {
                                s_next = s0;
                                s_end  = s0 + chunk < end ? s0 + chunk : end;
                                got    = true;
                            }
                        }
                        // a range that could not serve every lane is dry: move on
                        qx = (__ballot(!got) != 0) ? ((q + 1) & (MQI_QUEUES - 1)) : q;
                    }
                    if (!got) break;
                }
                g = s_next++;
            } else {
                if (s_next >= s_end) break;
                g = s_next++;
            }
            L.g = g;
            uint32_t spot_id;
            if (P.vertices) {
                // caller-made vertex (upload_vertices path); scenario 0
                spot_id  = P.vertex_spot ? P.vertex_spot[g] : 0;
                L.item   = spot_id;
                L.hist   = g;
                scen     = 0;
            } else {
                locate(P, g, L.item, L.hist);
                spot_id = P.item_spot[L.item];
                scen    = P.item_scenario[L.item];
            }
            c.g.sx                 = P.scen_shift ? P.scen_shift[3 * scen] : 0.0f;
            c.g.sy                 = P.scen_shift ? P.scen_shift[3 * scen + 1] : 0.0f;
            c.g.sz                 = P.scen_shift ? P.scen_shift[3 * scen + 2] : 0.0f;
            L.scale                = P.scen_scale ? P.scen_scale[scen] : 1.0f;
            L.scen_off             = (uint64_t) scen * P.out_scenario_stride;
            L.dij_item             = (uint64_t) scen * P.n_spots + spot_id;
            if (P.vertices) {
                const float* v = P.vertices + 7 * g;
                L.t.ke0        = v[0];
                L.t.p0         = mk(v[1], v[2], v[3]);
                L.t.d0         = mk(v[4], v[5], v[6]);
                L.t.ke1        = L.t.ke0;
                L.t.p1         = L.t.p0;
                L.t.d1         = L.t.d0;
                L.t.dE         = 0;
                L.t.ldE        = 0;
                L.t.stopped    = false;
                L.t.primary    = true;
            } else {
                sample_vertex(P, P.spots[spot_id], spot_id, scen, L.hist + hoff, L.t);
            }
            c.rng.init(P.seed, (uint32_t) (L.hist + hoff), spot_id, scen, mqi_rng::STREAM_TRANSPORT);
            L.nsteps = 0;
            L.tsteps = 0;
            ++done;
            c.child   = 0;
            L.reenter = true;
        } else if (!L.active) {
            pop(c, L.t);
            L.tsteps  = 0;
            c.child   = 0;
            L.reenter = true;
        }
        // a new or popped track, or one that left a beamline child, enters
        // the next child it reaches (one call site for all three)
        if (L.reenter) {
            L.reenter = false;
            L.active  = enter_chain(P, c, L.t, shb);
            if (!L.active) continue;
        }
        step<DBG, TRACE>(P, c, L, shb, pv);
        ++nstep;
    }
    flush(P, L.acc, pv);
    atomicAdd(P.tracked, done);
    atomicAdd(P.steps, nstep);
    if (c.drops) atomicAdd(P.drops, (unsigned long long) c.drops);
}

namespace
{
using launch_fn = void (*)(const transport_params&, int, int, size_t, hipStream_t);

struct variant {
    launch_fn   launch;
    const void* fn;
};

template<bool D, bool Y, bool T, int W>
void launch_one(const transport_params& P, int blocks, int threads, size_t shmem, hipStream_t stream) {
    hipLaunchKernelGGL((k_transport<D, Y, T, W>), dim3(blocks), dim3(threads), shmem, stream, P);
}

template<bool D, bool Y, bool T, int W>
variant make() {
    return {&launch_one<D, Y, T, W>, reinterpret_cast<const void*>(&k_transport<D, Y, T, W>)};
}

template<bool D, bool Y, bool T>
variant pick_waves(int waves) {
#if MQI_EXACT
    (void) waves;   // the verification kernels are not tuned
    return make<D, Y, T, 0>();
#else
    if (T) return make<D, Y, T, 0>();
    // 0: 3 waves/SIMD, the measured optimum of the production kernels with
    // deferred scoring on every plan measured (CT, Dij, robust, slab and
    // water phantoms: 4-16% faster than 4, whose extra waves add atomics in
    // flight and spills); 2: the compiler's own register budget
    switch (waves) {
    case 2: return make<D, Y, T, 0>();
    case 4: return make<D, Y, T, 4>();
    case 5: return make<D, Y, T, 5>();
    case 6: return make<D, Y, T, 6>();
    case 8: return make<D, Y, T, 8>();
    default: return make<D, Y, T, 3>();
    }
#endif
}

variant pick(int debug, int dynamic, int trace, int waves) {
    if (debug) {
        if (dynamic) return trace ? pick_waves<true, true, true>(waves) : pick_waves<true, true, false>(waves);
        return trace ? pick_waves<true, false, true>(waves) : pick_waves<true, false, false>(waves);
    }
    if (dynamic) return trace ? pick_waves<false, true, true>(waves) : pick_waves<false, true, false>(waves);
    return trace ? pick_waves<false, false, true>(waves) : pick_waves<false, false, false>(waves);
}
}   // namespace

int
launch_transport(const transport_params& P, int debug, int dynamic, int trace, int waves, int blocks, int threads,
                 size_t shmem, hipStream_t stream) {
    pick(debug, dynamic, trace, waves).launch(P, blocks, threads, shmem, stream);
    return (int) hipGetLastError();
}

// Verification probes of the material and physics functions as the
// transport kernels evaluate them: fn 0 spr (rho, ek), fn 1 radiation length
// (rho), fn 2 per (rho, ek) the four cross sections at rho, dEdx and the
// relativistic quantities beta^2, gamma, Et, Te_max, momentum (10 floats).
template<bool DBG>
__global__ void __launch_bounds__(256) k_physics_probe(int fn, const float* rho, const float* ek, int64_t n, float* out) {
    __shared__ float sh[T_TABLES];
    for (int i = threadIdx.x; i < 600; i += blockDim.x) {
        sh[T_CS_ION + i] = kCsPIon[i];
        sh[T_DEDX + i]   = kDedxRestricted[i];
        sh[T_RANGE + i]  = kRangeSteps[i];
        sh[T_CS_PP + i]  = kCsPPElastic[i];
        sh[T_CS_POE + i] = kCsPOElastic[i];
        sh[T_CS_POI + i] = kCsPOInelastic[i];
    }
    __syncthreads();
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (fn == 0) {
        out[i] = spr<DBG>(rho[i], ek[i]);
    } else if (fn == 1) {
        out[i] = radiation_length<DBG>(rho[i]);
    } else {
        ctx c;
        c.T           = sh;
        const rel_q r = relq(ek[i]);
        float*      o = out + 10 * i;
        o[0]          = cs_ion(c, r.Ek, rho[i]);
        o[1]          = cs_tab(c, T_CS_PP, r.Ek, rho[i]);
        o[2]          = cs_tab(c, T_CS_POE, r.Ek, rho[i]);
        o[3]          = cs_tab(c, T_CS_POI, r.Ek, rho[i]);
        o[4]          = dedx(c, r.Ek);
        o[5]          = r.beta_sq;
        o[6]          = r.gamma;
        o[7]          = r.Et;
        o[8]          = r.Te_max;
        o[9]          = momentum(r);
    }
}

int
physics_probe(int debug, int fn, const float* rho, const float* ek, int64_t n, float* out, hipStream_t stream) {
    if (n <= 0) return 0;
    const dim3 grid((unsigned) ((n + 255) / 256));
    if (debug)
        hipLaunchKernelGGL(k_physics_probe<true>, grid, dim3(256), 0, stream, fn, rho, ek, n, out);
    else
        hipLaunchKernelGGL(k_physics_probe<false>, grid, dim3(256), 0, stream, fn, rho, ek, n, out);
    return (int) hipGetLastError();
}

// resident workgroups per CU of one instantiation
int
occupancy(int debug, int dynamic, int trace, int waves, int threads, size_t shmem) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, pick(debug, dynamic, trace, waves).fn, threads, shmem) !=
        hipSuccess)
        return 1;
    return n < 1 ? 1 : n;
}

}   // namespace MQI_NS
