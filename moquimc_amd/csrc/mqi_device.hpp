// Device-side transport physics and voxel geometry for gfx950.
//
// The operation order, operand types and float/double promotions follow the
// reference expression by expression (file:line cited per function), so that
// the verification build (MQI_EXACT=1: portable math, no FMA contraction)
// reproduces the CPU restatement bit for bit, history by history.  The
// production build (MQI_EXACT=0) runs the same code with the device math
// library, FMA contraction and hardware float division and square root
// (outside the geometry).
//
// Organisation: one lane owns one track at a time.  Its state lives in
// registers (struct trk); secondaries go to a per-lane LIFO in HBM with the
// reference's capacity and drop-when-full rule; physics tables (6 x 600
// floats) and the grid edges are staged in LDS by each workgroup.
#ifndef MQI_DEVICE_HPP
#define MQI_DEVICE_HPP

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mqi_hip_types.h"
#include "mqi_philox.hpp"
#include "mqi_pmath.hpp"

#ifndef MQI_EXACT
#define MQI_EXACT 0
#endif

#if MQI_EXACT
#define MQ_LOGF(x) mqi_pm::pm_logf(x)
#define MQ_LOGD(x) mqi_pm::pm_log(x)
#define MQ_EXPD(x) mqi_pm::pm_exp(x)
#define MQ_POWF(x, y) mqi_pm::pm_powf(x, y)
#define MQ_SQRTF(x) mqi_pm::pm_sqrtf(x)
#define MQ_SQRTD(x) mqi_pm::pm_sqrt(x)
#define MQ_COSF(x) mqi_pm::pm_cosf(x)
#define MQ_SINF(x) mqi_pm::pm_sinf(x)
#define MQ_ACOSF(x) mqi_pm::pm_acosf(x)
#else
// production: device library, except pow = exp2(y log2 x) on the hardware
// log2 / exp2 (about 2 ulp here; the library's powf costs tens of
// instructions and stopping_power_ratio calls it twice per step)
#define MQ_LOGF(x) logf(x)
#define MQ_LOGD(x) log(x)
#define MQ_EXPD(x) exp(x)
#define MQ_POWF(x, y) __builtin_amdgcn_exp2f((y) * __builtin_amdgcn_logf(x))
#define MQ_SQRTF(x) sqrtf(x)
#define MQ_SQRTD(x) sqrt(x)
#define MQ_COSF(x) cosf(x)
#define MQ_SINF(x) sinf(x)
#define MQ_ACOSF(x) acosf(x)
#endif

#define MQI_DI __device__ __forceinline__
// Geometry is evaluated without FMA contraction in every build: traversal of
// deterministic rays must reproduce the reference bit for bit.
#define MQI_NO_CONTRACT _Pragma("clang fp contract(off)")

namespace mqi_dev
{

// Offsets of the physics tables in the LDS block (floats).
enum { T_CS_ION = 0, T_DEDX = 600, T_RANGE = 1200, T_CS_PP = 1800, T_CS_POE = 2400, T_CS_POI = 3000,
       T_TABLES = 3600 };

// physics_constants<float> (mqi_physics_constants.hpp:15-38), evaluated as the reference does.
struct units {
    static constexpr float cm       = 10.0;
    static constexpr float cm3      = cm * cm * cm;
    static constexpr float Mp       = 938.272046 * 1.0f;
    static constexpr float Mp_sq    = Mp * Mp;
    static constexpr float Me       = 0.510998928 * 1.0f;
    static constexpr float Mo       = 14903.3460795634 * 1.0f;
    static constexpr float MoMp     = Mo / Mp;
    static constexpr float MoMp_sq  = MoMp * MoMp;
    static constexpr float re       = 2.8179403262e-12 * 1.0f;
    static constexpr float re_sq    = re * re;
    static constexpr float two_pi_re2_mc2     = 2.0 * 3.14159265358979323846 * re_sq * Me;
    static constexpr float two_pi_re2_mc2_h2o = two_pi_re2_mc2 * 3.3428e+23 / cm3;
    static constexpr float water_density      = 1.0 / cm3;
    static constexpr float X0_water           = 36.0863 * cm;
};

struct v3 {
    float x, y, z;
};

MQI_DI v3 mk(float a, float b, float c) {
    v3 r;
    r.x = a;
    r.y = b;
    r.z = c;
    return r;
}
// 3x3 matrix times vector, row-major (mqi_matrix.hpp:282-287)
MQI_DI v3 mv(const float* m, v3 r) {
    MQI_NO_CONTRACT
    return mk(m[0] * r.x + m[1] * r.y + m[2] * r.z, m[3] * r.x + m[4] * r.y + m[5] * r.z,
              m[6] * r.x + m[7] * r.y + m[8] * r.z);
}
// Division and square root where the reference's float results are needed
// to the bit (geometry, track set-up).  The production build compiles float
// division and square root to the hardware reciprocal / square root (about
// 1 ulp); these evaluate in double and round once to float, which gives the
// correctly rounded float result (double rounding is innocuous for division
// and square root when the wider format has at least 2 * 24 + 2 significand
// bits).
MQI_DI float geo_div(float a, float b) {
#if MQI_EXACT || !defined(__HIP_DEVICE_COMPILE__)
    return a / b;
#else
    return (float) ((double) a / (double) b);
#endif
}
MQI_DI float geo_sqrt(float a) {
#if MQI_EXACT || !defined(__HIP_DEVICE_COMPILE__)
    return MQ_SQRTF(a);
#else
    return (float) __builtin_sqrt((double) a);
#endif
}

// vec3::normalize (mqi_vec.hpp:205-211), correctly rounded in every build
MQI_DI void normalize(v3& a) {
    MQI_NO_CONTRACT
    float n = geo_sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    a.x     = geo_div(a.x, n);
    a.y     = geo_div(a.y, n);
    a.z     = geo_div(a.z, n);
}
// the same in the physics (scattering rotations): hardware division and
// square root in the production build
MQI_DI void normalize_fast(v3& a) {
    MQI_NO_CONTRACT
    float n = MQ_SQRTF(a.x * a.x + a.y * a.y + a.z * a.z);
    a.x /= n;
    a.y /= n;
    a.z /= n;
}

// relativistic_quantities (mqi_relativistic_quantities.hpp:24-43)
struct rel_q {
    float beta_sq, gamma_sq, gamma, Et, Et_sq, Ek, Te_max;
};
MQI_DI rel_q relq(float Ek) {
    rel_q         r;
    const float   Mp      = 938.272046;
    const float   Me      = 0.510998928;
    r.Ek                  = Ek;
    r.Et                  = Ek + Mp;
    r.Et_sq               = r.Et * r.Et;
    const float   MeMp    = Me / Mp;
    const float   MeMp_sq = MeMp * MeMp;
    r.gamma               = r.Et / Mp;
    r.gamma_sq            = r.gamma * r.gamma;
#if MQI_EXACT
    r.beta_sq = 1.0 - 1.0 / r.gamma_sq;
    r.Te_max  = (2.0 * Me * r.beta_sq * r.gamma_sq);
    r.Te_max /= (1.0 + 2.0 * r.gamma * MeMp + MeMp_sq);
#else
    // production: float, and beta^2 = Ek (Ek + 2 Mp) / Et^2, the same
    // quantity without the cancellation of 1 - 1 / gamma^2 at low energy
    r.beta_sq = Ek * (Ek + 2.0f * Mp) / r.Et_sq;
    r.Te_max  = (2.0f * Me * r.beta_sq * r.gamma_sq) / (1.0f + 2.0f * r.gamma * MeMp + MeMp_sq);
#endif
    return r;
}
MQI_DI float momentum(const rel_q& r) {
#if MQI_EXACT
    return MQ_SQRTF(r.Et * r.Et - units::Mp * units::Mp);
#else
    // production: p^2 = Ek (Ek + 2 Mp), the same quantity without the
    // cancellation of Et^2 - Mp^2 in float at low energy
    return MQ_SQRTF(r.Ek * (r.Ek + 2.0f * units::Mp));
#endif
}

MQI_DI float intpl(float x, float x0, float x1, float y0, float y1) {
    return (x1 == x0) ? y0 : y0 + (x - x0) * (y1 - y0) / (x1 - x0);
}

// spr_default (mqi_patient_materials.hpp:415-450)
template<bool DBG>
MQI_DI float spr(float rho_mass, float Ek) {
    float density_tmp = rho_mass * 1000.0;
    if (DBG) {
        if (fabs(density_tmp - 1.0) < 1e-3) return 1.0f;
    }
    if (density_tmp <= 0.26) {
        if (density_tmp < 0.0012) return 0.0f;
        return intpl(density_tmp, (float) 0.0012, (float) 0.26, (float) 0.8815, (float) 0.9925);
    }
    float rsp = 1.0123 - 3.386e-5 * Ek;
    rsp += 0.291 * (1.0 + MQ_POWF(Ek, (float) -0.3421)) * (MQ_POWF(density_tmp, (float) -0.7) - 1.0);
    if (density_tmp >= 0.9) return rsp;
    return intpl(density_tmp, (float) 0.26, (float) 0.9, (float) 0.9925, rsp);
}

// radiation_length_default (mqi_patient_materials.hpp:452-473)
template<bool DBG>
MQI_DI float radiation_length(float rho_mass) {
    float f       = 0.0;
    float density = rho_mass * 1000.0;
    if (DBG) {
        if (fabs(density - 1.0) < 1e-3) return units::X0_water;
    }
    if (density <= 0.26) {
        f = 0.9857 + 0.0085 * density;
    } else if (density > 0.26 && density <= 0.9) {
        f = 1.0446 - 0.2180 * density;
    } else if (density > 0.9) {
#if MQI_EXACT
        f = 1.19 + 0.44 * MQ_LOGD((double) (density - 0.44));
#else
        f = 1.19f + 0.44f * __logf(density - 0.44f);   // production: float
#endif
    }
#if MQI_EXACT
    return (float) ((units::water_density * units::X0_water) / (density * 0.001 * f));
#else
    return (units::water_density * units::X0_water) / (density * 0.001f * f);
#endif
}

// hu_to_density (mqi_patient_materials.hpp:515-542)
MQI_DI float hu_to_density(int hu_in, const float* corr) {
    int16_t hu = (int16_t) hu_in;
    if (hu < -1000) {
        hu = -1000;
    } else if (hu > 2995) {
        hu = 2995;
    }
    float rho_mass = 0.0;
    if (hu >= -1000 && hu < -98) {
        rho_mass = 0.00121 + 0.001029700665188 * (1000.0 + hu);
    } else if (hu >= -98 && hu < 15) {
        rho_mass = 1.018 + 0.000893 * hu;
    } else if (hu >= 15 && hu < 23) {
        rho_mass = 1.03;
    } else if (hu >= 23 && hu < 101) {
        rho_mass = 1.003 + 0.001169 * hu;
    } else if (hu >= 101 && hu < 2001) {
        rho_mass = 1.017 + 0.000592 * hu;
    } else if (hu >= 2001 && hu < 2995) {
        rho_mass = 2.201 + 0.0005 * (-2000.0 + hu);
    } else {
        rho_mass = 4.54;
    }
    rho_mass *= corr[hu + 1000];
    rho_mass /= 1000.0;
    return rho_mass;
}

// ------------------------------------------------------------------ geometry
// grid3d index(p, dir) (mqi_grid3d.hpp:746-844).  The reference scans every
// edge; here a binary search finds k = max{i : e[i] <= p} and only the edges
// k-1 .. k+1 are tested with the reference's predicates, in the reference's
// order.  With every edge spacing above twice the 1e-3 mm tolerance (checked
// by the host at upload) no other index can satisfy a predicate, so the first
// match, and the -1 when none matches, are the reference's.
// Edges are read as e[i] + shift: a robust setup shift moves the whole grid,
// and one float addition per edge gives exactly the shifted edge array.
MQI_DI int scan_axis(const float* e, float sh, int n, float p, float dir) {
    MQI_NO_CONTRACT
    int lo = -1, hi = n + 1;   // e[lo] <= p < e[hi] (virtual sentinels)
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (e[mid] + sh <= p)
            lo = mid;
        else
            hi = mid;
    }
    int a = lo - 1 < 0 ? 0 : lo - 1;
    int b = lo + 1 > n - 1 ? n - 1 : lo + 1;
    const float tol = 1e-3f;
    for (int ind = a; ind <= b; ++ind) {
        const float e0 = e[ind] + sh, e1 = e[ind + 1] + sh;
        if (fabsf(e0 - p) < tol) {
            if (dir > 0) return ind;
            if (dir < 0) return ind - 1;
            return ind;
        } else if (fabsf(e1 - p) < tol) {
            if (dir > 0) return ind + 1;
            if (dir < 0) return ind;
            return ind;
        } else if (e0 - p < 0 && e1 - p > 0) {
            return ind;
        }
    }
    return -1;
}

struct grid_view {
    int          nx, ny, nz;
    const float* xe;   // LDS
    const float* ye;
    const float* ze;
    float        sx, sy, sz;   // scenario shift of the edges
    MQI_DI float X(int i) const { return xe[i] + sx; }
    MQI_DI float Y(int i) const { return ye[i] + sy; }
    MQI_DI float Z(int i) const { return ze[i] + sz; }
};

MQI_DI bool is_valid(const grid_view& g, int cx, int cy, int cz) {
    if (cx < 0 || cy < 0 || cz < 0) return false;
    if (cx >= g.nx || cy >= g.ny || cz >= g.nz) return false;
    return true;
}

MQI_DI void index3(const grid_view& g, v3 p, v3 d, int& cx, int& cy, int& cz) {
    cx = scan_axis(g.xe, g.sx, g.nx, p.x, d.x);
    cy = scan_axis(g.ye, g.sy, g.ny, p.y, d.y);
    cz = scan_axis(g.ze, g.sz, g.nz, p.z, d.z);
}

// one axis of intersect(p, d, idx) (mqi_grid3d.hpp:519-545); mutates dcomp
MQI_DI void axis_in(float& dcomp, float p, float v1, float v2, int i, int n, float& tmax) {
    MQI_NO_CONTRACT
    const float tol = 1e-3f;
    float       me  = dcomp;
    if (me * me > 1e-7f) {
        if (me < 0) {
            if (fabsf(geo_div(-(p - v1), dcomp)) < tol && i > 0)
                tmax = 1 / tol;
            else
                tmax = geo_div(-(p - v1), dcomp);
        } else {
            if (fabsf(geo_div(v2 - p, dcomp)) < tol && i < n)
                tmax = 1 / tol;
            else
                tmax = geo_div(v2 - p, dcomp);
        }
    } else {
        dcomp = 0;
        tmax  = __builtin_huge_valf();
    }
}

// intersect(p, d, idx): distance to the voxel exit; -1 (cell := -1) when none (mqi_grid3d.hpp:490-626)
MQI_DI float intersect_in(const grid_view& g, v3 p, v3& d, int& cx, int& cy, int& cz) {
    MQI_NO_CONTRACT
    float tx, ty, tz;
    axis_in(d.x, p.x, g.X(cx), g.X(cx + 1), cx, g.nx, tx);
    axis_in(d.y, p.y, g.Y(cy), g.Y(cy + 1), cy, g.ny, ty);
    axis_in(d.z, p.z, g.Z(cz), g.Z(cz + 1), cz, g.nz, tz);
    float u_max;
    if (tx < ty)
        u_max = (tx < tz) ? tx : tz;
    else
        u_max = (ty < tz) ? ty : tz;
    if (u_max > 0) return u_max;
    cx = cy = cz = -1;
    return -1.0f;
}

MQI_DI void axis_out(float& dcomp, float p, float lo, float hi, float& tmin, float& tmax) {
    MQI_NO_CONTRACT
    float me = dcomp;
    if (me * me > 1e-7f) {
        if (me > 0) {
            tmin = geo_div(lo - p, dcomp);
            tmax = geo_div(hi - p, dcomp);
        } else {
            tmax = geo_div(lo - p, dcomp);
            tmin = geo_div(hi - p, dcomp);
        }
    } else {
        dcomp = 0;
        tmin  = -__builtin_huge_valf();
        tmax  = __builtin_huge_valf();
    }
}

// intersect(p, d) from outside the box (mqi_grid3d.hpp:631-743)
MQI_DI float intersect_out(const grid_view& g, v3 p, v3& d, int& cx, int& cy, int& cz) {
    MQI_NO_CONTRACT
    if (p.x >= g.X(0) && p.x <= g.X(g.nx) && p.y >= g.Y(0) && p.y <= g.Y(g.ny) && p.z >= g.Z(0) &&
        p.z <= g.Z(g.nz)) {
        index3(g, p, d, cx, cy, cz);
        return 0.0f;
    }
    cx = cy = cz = -1;
    float dist = -1.0f;
    float tminx, tmaxx, tminy, tmaxy, tminz, tmaxz;
    axis_out(d.x, p.x, g.X(0), g.X(g.nx), tminx, tmaxx);
    axis_out(d.y, p.y, g.Y(0), g.Y(g.ny), tminy, tmaxy);
    axis_out(d.z, p.z, g.Z(0), g.Z(g.nz), tminz, tmaxz);
    float u_min, u_max;
    if (tminx > tminy)
        u_min = (tminx > tminz) ? tminx : tminz;
    else
        u_min = (tminy > tminz) ? tminy : tminz;
    if (tmaxx < tmaxy)
        u_max = (tmaxx < tmaxz) ? tmaxx : tmaxz;
    else
        u_max = (tmaxy < tmaxz) ? tmaxy : tmaxz;
    if ((u_min < u_max || fabsf(u_min - u_max) < 1e-3f) && u_min >= 0 && u_max >= 0) {
        dist = u_min;
        v3 on = mk(p.x + d.x * dist, p.y + d.y * dist, p.z + d.z * dist);
        index3(g, on, d, cx, cy, cz);
    }
    return dist;
}

// index(vtx1, dir1, idx): +-1 cell update (mqi_grid3d.hpp:847-877)
MQI_DI void step_axis(float v, float dir, const float* e, float sh, int& i) {
    MQI_NO_CONTRACT
    const float tol = 1e-3f;
    if (dir < 0) {
        const float e0 = e[i] + sh;
        if (fabsf(v - e0) < tol || v < e0) i -= 1;
    } else if (dir > 0) {
        const float e1 = e[i + 1] + sh;
        if (fabsf(v - e1) < tol || v > e1) i += 1;
    }
}

// ------------------------------------------------------------------ track
struct trk {
    float ke0;
    v3    p0, d0;
    float ke1;
    v3    p1, d1;
    float dE, ldE, its;
    int   cx, cy, cz;
    bool  stopped;
    bool  primary;   // the history's own track (daughters are not)
};

MQI_DI void update_pos(trk& t, float len) {   // mqi_track.hpp:178-180
    MQI_NO_CONTRACT
    t.p1 = mk(t.p0.x + t.d0.x * len, t.p0.y + t.d0.y * len, t.p0.z + t.d0.z * len);
}
MQI_DI void move(trk& t) {   // mqi_track.hpp:193-197
    t.ke0 = t.ke1;
    t.p0  = t.p1;
    t.d0  = t.d1;
    t.dE  = 0;
    t.ldE = 0;
}

// rotate_y / rotate_z on a matrix held as 9 floats (mqi_matrix.hpp:231-273)
MQI_DI void rotate_y(float* m, float a) {
    float c1 = MQ_COSF(a), s1 = MQ_SINF(a);
    float x1 = m[6], y1 = m[7], z1 = m[8];
    m[6]     = c1 * x1 - s1 * m[0];
    m[7]     = c1 * y1 - s1 * m[1];
    m[8]     = c1 * z1 - s1 * m[2];
    m[0]     = s1 * x1 + c1 * m[0];
    m[1]     = s1 * y1 + c1 * m[1];
    m[2]     = s1 * z1 + c1 * m[2];
}
MQI_DI void rotate_z(float* m, float a) {
    float c1 = MQ_COSF(a), s1 = MQ_SINF(a);
    float x1 = m[0], y1 = m[1], z1 = m[2];
    m[0]     = c1 * x1 - s1 * m[3];
    m[1]     = c1 * y1 - s1 * m[4];
    m[2]     = c1 * z1 - s1 * m[5];
    m[3]     = s1 * x1 + c1 * m[3];
    m[4]     = s1 * y1 + c1 * m[4];
    m[5]     = s1 * z1 + c1 * m[5];
}

// track_t::update_post_vertex_direction (mqi_track.hpp:165-172) on direction d:
// m_local = mat3x3(0, theta, phi); d_local = m_local * z; align z to d
// (Moller-Hughes, mqi_matrix.hpp:89-150); d = m_global * d_local, normalised.
MQI_DI void rotate_direction(v3& d, float theta, float phi) {
    float m[9] = { 1.0f, 0, 0, 0, 1.0f, 0, 0, 0, 1.0f };
    if (theta != 0) rotate_y(m, theta);
    if (phi != 0) rotate_z(m, phi);
    const v3 ref = mk(0, 0, 1);
    v3       dl  = mv(m, ref);
    normalize_fast(dl);
    // mat3x3(f = ref, t = d)
    v3    v  = mk(ref.y * d.z - ref.z * d.y, ref.z * d.x - ref.x * d.z, ref.x * d.y - ref.y * d.x);
    float fn = MQ_SQRTF(ref.x * ref.x + ref.y * ref.y + ref.z * ref.z);
    float tn = MQ_SQRTF(d.x * d.x + d.y * d.y + d.z * d.z);
    float c  = (ref.x * d.x + ref.y * d.y + ref.z * d.z) / (fn * tn);
    float g[9];
    if (fabsf(c - 1) < 1e-3f || fabsf(c + 1) < 1e-3f) {
        v3 uu = mk(1 - ref.x, 0 - ref.y, 0 - ref.z);
        v3 vv = mk(1 - d.x, 0 - d.y, 0 - d.z);
        normalize_fast(uu);
        normalize_fast(vv);
        float dot_u  = uu.x * uu.x + uu.y * uu.y + uu.z * uu.z;
        float dot_v  = vv.x * vv.x + vv.y * vv.y + vv.z * vv.z;
        float dot_uv = vv.x * uu.x + vv.y * uu.y + vv.z * uu.z;
        float U[3]   = { uu.x, uu.y, uu.z };
        float V[3]   = { vv.x, vv.y, vv.z };
        // fully unrolled: a runtime index into g, U or V would put them in private memory
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                g[3 * i + j] = (i == j ? 1 : 0) - 2 / dot_u * U[i] * U[j] - 2 / dot_v * V[i] * V[j] +
                               4 * dot_uv / (dot_u * dot_v) * V[i] * U[j];
    } else {
        float h = (1.0 - c) / (1.0 - c * c);
        g[0]    = c + h * v.x * v.x;
        g[1]    = h * v.x * v.y - v.z;
        g[2]    = h * v.x * v.z + v.y;
        g[3]    = h * v.x * v.y + v.z;
        g[4]    = c + h * v.y * v.y;
        g[5]    = h * v.y * v.z - v.x;
        g[6]    = h * v.x * v.z - v.y;
        g[7]    = h * v.y * v.z + v.x;
        g[8]    = c + h * v.z * v.z;
    }
    d = mv(g, dl);
    normalize_fast(d);
}

// ------------------------------------------------------------------ random numbers
// The verification build draws every number from one lazily refilled stream
// in the order the CPU restatement does.  The production build draws the four
// numbers every step takes (mean free path, straggling and scattering
// normals, azimuth) from one Philox block made for all lanes at the start of
// the step, and the first three of an interaction (process, first delta-ray
// trial) from one block made where the step ends in an interaction; the lazy
// stream serves the rest (later delta-ray trials, nuclear interactions).
// With the lazy stream alone, every draw site refilled for some lane of the
// wave (8.3 block evaluations per wave step on the 2000-spot CT plan).
struct rng_t {
    mqi_rng::stream s;
#if !MQI_EXACT
    uint32_t step_blk, int_blk;   // blocks of the two per-step streams used so far
    float    st_u0, st_u1, st_z0, st_z1;   // the step's two uniforms and two normals
    float    it_u0, it_u1, it_u2;          // the interaction's three uniforms
#endif
    MQI_DI void init(uint64_t seed, uint32_t history, uint32_t spot, uint32_t scenario, uint32_t purpose) {
        s.init(seed, history, spot, scenario, purpose);
#if !MQI_EXACT
        step_blk = 0;
        int_blk  = 0;
#endif
    }
    MQI_DI float uniform() { return s.uniform(); }
    // standard normal: the exact build evaluates Box-Muller as the CPU
    // restatement does; the production build in float with the device library
    MQI_DI float gauss() {
#if MQI_EXACT
        return (float) s.gauss();
#else
        if (s.has_spare) {
            s.has_spare = 0;
            return (float) s.spare;
        }
        const float u1 = s.uniform();
        const float u2 = s.uniform();
        const float r  = sqrtf(-2.0f * logf(u1));
        float       sn, cs;
        sincosf(6.28318530717958647f * u2, &sn, &cs);
        s.spare     = r * sn;
        s.has_spare = 1;
        return r * cs;
#endif
    }
    MQI_DI float normal(float avg, float sig) {
        float z = gauss();
        return z * sig + avg;
    }

    // the draws of a step: step_uniform(0) the mean free path, step_normal(0)
    // the energy straggling, step_normal(1) the scattering angle,
    // step_uniform(1) the azimuth, in that order
    MQI_DI void begin_step() {
#if !MQI_EXACT
        const mqi_rng::u32x4 w = s.block_of(mqi_rng::STREAM_STEP, step_blk++);
        st_u0                  = mqi_rng::stream::to_uniform(w.v[0]);
        st_u1                  = mqi_rng::stream::to_uniform(w.v[3]);
        // Box-Muller on the hardware log2, square root, sine and cosine (the
        // last two take revolutions: sin(2 pi u))
        const float u1 = mqi_rng::stream::to_uniform(w.v[1]);
        const float u2 = mqi_rng::stream::to_uniform(w.v[2]);
        const float r  = __builtin_sqrtf(-2.0f * 0.693147180559945309f * __builtin_amdgcn_logf(u1));
        st_z0          = r * __builtin_amdgcn_cosf(u2);
        st_z1          = r * __builtin_amdgcn_sinf(u2);
#endif
    }
    MQI_DI float step_uniform(int k) {
#if MQI_EXACT
        (void) k;
        return uniform();
#else
        return k == 0 ? st_u0 : st_u1;
#endif
    }
    MQI_DI float step_normal(int k, float avg, float sig) {
#if MQI_EXACT
        (void) k;
        return normal(avg, sig);
#else
        return (k == 0 ? st_z0 : st_z1) * sig + avg;
#endif
    }
    // the draws of an interaction: int_uniform(0) the process,
    // int_uniform(1) and (2) the first delta-ray trial; later trials and the
    // nuclear interactions draw uniform()
    MQI_DI void begin_interaction() {
#if !MQI_EXACT
        const mqi_rng::u32x4 w = s.block_of(mqi_rng::STREAM_INTERACTION, int_blk++);
        it_u0                  = mqi_rng::stream::to_uniform(w.v[0]);
        it_u1                  = mqi_rng::stream::to_uniform(w.v[1]);
        it_u2                  = mqi_rng::stream::to_uniform(w.v[2]);
#endif
    }
    MQI_DI float int_uniform(int k) {
#if MQI_EXACT
        (void) k;
        return uniform();
#else
        return k == 0 ? it_u0 : (k == 1 ? it_u1 : it_u2);
#endif
    }
};

}   // namespace mqi_dev

#endif
