// Test infrastructure: CPU restatement ("oracle") of the moqui transport and
// scoring path.  See moqui_oracle.hpp for the policies.  Every function cites
// the reference file:line it restates (paths relative to the reference root).
//
// For prec = 32 each expression keeps the operand types of the reference
// (float members and locals, double literals), so that C++'s usual arithmetic
// conversions promote exactly where the reference's do; with rng = 0 and
// math = 0 the result is bit-identical to the reference compiled for the CPU
// (checked by tests/test_oracle_vs_reference.py).  For prec = 64 the same
// code runs in double (the FP64 oracle).
#include "moqui_oracle.hpp"

#include <cmath>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "moquimc_amd/csrc/mqi_philox.hpp"
#include "moquimc_amd/csrc/mqi_pmath.hpp"
#define MQI_TABLE_QUALIFIER static
#include "moquimc_amd/csrc/mqi_tables.inc"

namespace orc
{

// ---------------------------------------------------------------- math
struct LibmMath {
    static float  ln(float x) { return std::log(x); }
    static double ln(double x) { return std::log(x); }
    static float  ex(float x) { return std::exp(x); }
    static double ex(double x) { return std::exp(x); }
    static float  pw(float x, float y) { return std::pow(x, y); }
    static double pw(double x, double y) { return std::pow(x, y); }
    static float  sq(float x) { return std::sqrt(x); }
    static double sq(double x) { return std::sqrt(x); }
    static float  co(float x) { return std::cos(x); }
    static double co(double x) { return std::cos(x); }
    static float  si(float x) { return std::sin(x); }
    static double si(double x) { return std::sin(x); }
    static float  ac(float x) { return std::acos(x); }
    static double ac(double x) { return std::acos(x); }
};
struct PortMath {
    static float  ln(float x) { return mqi_pm::pm_logf(x); }
    static double ln(double x) { return mqi_pm::pm_log(x); }
    static float  ex(float x) { return mqi_pm::pm_expf(x); }
    static double ex(double x) { return mqi_pm::pm_exp(x); }
    static float  pw(float x, float y) { return mqi_pm::pm_powf(x, y); }
    static double pw(double x, double y) { return mqi_pm::pm_pow(x, y); }
    static float  sq(float x) { return mqi_pm::pm_sqrtf(x); }
    static double sq(double x) { return mqi_pm::pm_sqrt(x); }
    static float  co(float x) { return mqi_pm::pm_cosf(x); }
    static double co(double x) { return mqi_pm::pm_cos(x); }
    static float  si(float x) { return mqi_pm::pm_sinf(x); }
    static double si(double x) { return mqi_pm::pm_sin(x); }
    static float  ac(float x) { return mqi_pm::pm_acosf(x); }
    static double ac(double x) { return mqi_pm::pm_acos(x); }
};

// ---------------------------------------------------------------- RNG
// mqi_math.hpp:454-506 (CPU specialisations): a fresh distribution object
// per draw on one std::default_random_engine.
struct StdRng {
    std::default_random_engine e;
    void                       seed_history(uint64_t, uint32_t, uint32_t, uint32_t, int32_t seed) {
        e.seed((uint32_t) seed);
    }
    template<typename R>
    R
    uniform() {
        std::uniform_real_distribution<R> d;
        return d(e);
    }
    template<typename R>
    R
    normal(R avg, R sig) {
        std::normal_distribution<R> d(avg, sig);
        return d(e);
    }
    // float: untruncated (mqi_math.hpp:487-495); double: truncated at up (:498-505)
    template<typename R, typename M>
    R
    exponential(R avg, R up) {
        std::exponential_distribution<R> d(avg);
        if (sizeof(R) == 4) return d(e);
        R x;
        do {
            x = d(e);
        } while (x > up || x <= 0);
        return x;
    }
};

// The GPU kernel's stream (mqi_philox.hpp).  exponential: -ln(u)/avg, not
// truncated (the CPU reference's semantics).
struct PhiloxRng {
    mqi_rng::stream s;
    void            seed_history(uint64_t master, uint32_t h, uint32_t spot, uint32_t scen, int32_t) {
        s.init(master, h, spot, scen, mqi_rng::STREAM_TRANSPORT);
    }
    template<typename R>
    R
    uniform() {
        return (R) s.uniform();
    }
    template<typename R>
    R
    normal(R avg, R sig) {
        R z = (R) s.gauss();
        return z * sig + avg;
    }
    template<typename R, typename M>
    R
    exponential(R avg, R up) {
        R u = uniform<R>();
        return -M::ln(u) / avg;
    }
};

// ---------------------------------------------------------------- vec/mat
// mqi_vec.hpp:140-270, mqi_matrix.hpp:20-300
template<typename R>
struct V3 {
    R x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(R a, R b, R c) : x(a), y(b), z(c) {}
    V3 operator+(const V3& r) const { return V3(x + r.x, y + r.y, z + r.z); }
    V3 operator-(const V3& r) const { return V3(x - r.x, y - r.y, z - r.z); }
    V3 operator*(const R& r) const { return V3(x * r, y * r, z * r); }
    R  dot(const V3& v) const { return x * v.x + y * v.y + z * v.z; }
    V3 cross(const V3& r) const { return V3(y * r.z - z * r.y, z * r.x - x * r.z, x * r.y - y * r.x); }
    template<typename M>
    R norm() const {
        return M::sq(x * x + y * y + z * z);
    }
    template<typename M>
    void normalize() {
        R n = M::sq(x * x + y * y + z * z);
        x /= n;
        y /= n;
        z /= n;
    }
};

template<typename R>
struct M3 {
    R xx, xy, xz, yx, yy, yz, zx, zy, zz;
    M3() : xx(1.0), xy(0), xz(0), yx(0), yy(1.0), yz(0), zx(0), zy(0), zz(1.0) {}
    M3(R a, R b, R c, R d, R e, R f, R g, R h, R i) :
        xx(a), xy(b), xz(c), yx(d), yy(e), yz(f), zx(g), zy(h), zz(i) {}
    V3<R> operator*(const V3<R>& r) const {
        return V3<R>(xx * r.x + xy * r.y + xz * r.z, yx * r.x + yy * r.y + yz * r.z,
                     zx * r.x + zy * r.y + zz * r.z);
    }
    M3 inverse() const { return M3(xx, yx, zx, xy, yy, zy, xz, yz, zz); }
    // rotate_x / rotate_y / rotate_z (mqi_matrix.hpp:212-273)
    template<typename M>
    void rotate_x(R a) {
        R c1 = M::co(a), s1 = M::si(a);
        R x1 = yx, y1 = yy, z1 = yz;
        yx   = c1 * x1 - s1 * zx;
        yy   = c1 * y1 - s1 * zy;
        yz   = c1 * z1 - s1 * zz;
        zx   = s1 * x1 + c1 * zx;
        zy   = s1 * y1 + c1 * zy;
        zz   = s1 * z1 + c1 * zz;
    }
    // mat3x3(a, b, c): rotations applied only for non-zero angles (mqi_matrix.hpp:69-76)
    template<typename M>
    static M3 euler(R a, R b, R c) {
        M3 m;
        if (a != 0) m.template rotate_x<M>(a);
        if (b != 0) m.template rotate_y<M>(b);
        if (c != 0) m.template rotate_z<M>(c);
        return m;
    }
    M3 operator*(const M3& r) const {
        return M3(xx * r.xx + xy * r.yx + xz * r.zx, xx * r.xy + xy * r.yy + xz * r.zy,
                  xx * r.xz + xy * r.yz + xz * r.zz, yx * r.xx + yy * r.yx + yz * r.zx,
                  yx * r.xy + yy * r.yy + yz * r.zy, yx * r.xz + yy * r.yz + yz * r.zz,
                  zx * r.xx + zy * r.yx + zz * r.zx, zx * r.xy + zy * r.yy + zz * r.zy,
                  zx * r.xz + zy * r.yz + zz * r.zz);
    }
    template<typename M>
    void rotate_y(R a) {
        R c1 = M::co(a), s1 = M::si(a);
        R x1 = zx, y1 = zy, z1 = zz;
        zx   = c1 * x1 - s1 * xx;
        zy   = c1 * y1 - s1 * xy;
        zz   = c1 * z1 - s1 * xz;
        xx   = s1 * x1 + c1 * xx;
        xy   = s1 * y1 + c1 * xy;
        xz   = s1 * z1 + c1 * xz;
    }
    template<typename M>
    void rotate_z(R a) {
        R c1 = M::co(a), s1 = M::si(a);
        R x1 = xx, y1 = xy, z1 = xz;
        xx   = c1 * x1 - s1 * yx;
        xy   = c1 * y1 - s1 * yy;
        xz   = c1 * z1 - s1 * yz;
        yx   = s1 * x1 + c1 * yx;
        yy   = s1 * y1 + c1 * yy;
        yz   = s1 * z1 + c1 * yz;
    }
    // mat3x3(f, t): Moller-Hughes alignment of f to t (mqi_matrix.hpp:89-150)
    template<typename M>
    static M3 align(const V3<R>& f, const V3<R>& t) {
        M3    m;
        V3<R> v = f.cross(t);
        R     c = f.dot(t) / (f.template norm<M>() * t.template norm<M>());
        R     h = 0;
        if (std::abs(c - 1) < (R) 1e-3f || std::abs(c + 1) < (R) 1e-3f) {
            V3<R> x(1, 0, 0);
            V3<R> uu = x - f;
            V3<R> vv = x - t;
            uu.template normalize<M>();
            vv.template normalize<M>();
            R dot_u  = uu.dot(uu);
            R dot_v  = vv.dot(vv);
            R dot_uv = vv.dot(uu);
            m.xx = 1 - 2 / dot_u * uu.x * uu.x - 2 / dot_v * vv.x * vv.x + 4 * dot_uv / (dot_u * dot_v) * vv.x * uu.x;
            m.xy = 0 - 2 / dot_u * uu.x * uu.y - 2 / dot_v * vv.x * vv.y + 4 * dot_uv / (dot_u * dot_v) * vv.x * uu.y;
            m.xz = 0 - 2 / dot_u * uu.x * uu.z - 2 / dot_v * vv.x * vv.z + 4 * dot_uv / (dot_u * dot_v) * vv.x * uu.z;
            m.yx = 0 - 2 / dot_u * uu.y * uu.x - 2 / dot_v * vv.y * vv.x + 4 * dot_uv / (dot_u * dot_v) * vv.y * uu.x;
            m.yy = 1 - 2 / dot_u * uu.y * uu.y - 2 / dot_v * vv.y * vv.y + 4 * dot_uv / (dot_u * dot_v) * vv.y * uu.y;
            m.yz = 0 - 2 / dot_u * uu.y * uu.z - 2 / dot_v * vv.y * vv.z + 4 * dot_uv / (dot_u * dot_v) * vv.y * uu.z;
            m.zx = 0 - 2 / dot_u * uu.z * uu.x - 2 / dot_v * vv.z * vv.x + 4 * dot_uv / (dot_u * dot_v) * vv.z * uu.x;
            m.zy = 0 - 2 / dot_u * uu.z * uu.y - 2 / dot_v * vv.z * vv.y + 4 * dot_uv / (dot_u * dot_v) * vv.z * uu.y;
            m.zz = 1 - 2 / dot_u * uu.z * uu.z - 2 / dot_v * vv.z * vv.z + 4 * dot_uv / (dot_u * dot_v) * vv.z * uu.z;
        } else {
            h    = (1.0 - c) / (1.0 - c * c);
            m.xx = c + h * v.x * v.x;
            m.xy = h * v.x * v.y - v.z;
            m.xz = h * v.x * v.z + v.y;
            m.yx = h * v.x * v.y + v.z;
            m.yy = c + h * v.y * v.y;
            m.yz = h * v.y * v.z - v.x;
            m.zx = h * v.x * v.z - v.y;
            m.zy = h * v.y * v.z + v.x;
            m.zz = c + h * v.z * v.z;
        }
        return m;
    }
};

template<typename R>
struct Vtx {
    R     ke;
    V3<R> pos, dir;
};

// ---------------------------------------------------------------- constants
// physics_constants<R> (mqi_physics_constants.hpp:15-38)
template<typename R>
struct Units {
    const R mm                 = 1.0;
    const R cm                 = 10.0;
    const R cm3                = cm * cm * cm;
    const R MeV                = 1.0;
    const R Mp                 = 938.272046 * MeV;
    const R Mp_sq              = Mp * Mp;
    const R Me                 = 0.510998928 * MeV;
    const R Mo                 = 14903.3460795634 * MeV;
    const R Mo_sq              = Mo * Mo;
    const R MoMp               = Mo / Mp;
    const R MoMp_sq            = MoMp * MoMp;
    const R MeMp               = Me / Mp;
    const R MeMp_sq            = MeMp * MeMp;
    const R re                 = 2.8179403262e-12 * mm;
    const R re_sq              = re * re;
    const R two_pi_re2_mc2     = 2.0 * M_PI * re_sq * Me;
    const R two_pi_re2_mc2_h2o = two_pi_re2_mc2 * 3.3428e+23 / cm3;
    const R water_density      = 1.0 / cm3;
    const R radiation_length_water = 36.0863 * cm;
};

// relativistic_quantities (mqi_relativistic_quantities.hpp:11-54)
template<typename R, typename M>
struct Rel {
    R beta_sq, gamma_sq, gamma, Et, Et_sq, Ek, mc2, tau, Te_max;
    Rel(R kinetic_energy, R rest_mass_MeV) : Ek(kinetic_energy), mc2(rest_mass_MeV) {
        const R Mp      = 938.272046;
        const R Me      = 0.510998928;
        Et              = Ek + Mp;
        Et_sq           = Et * Et;
        const R MeMp    = Me / Mp;
        const R MeMp_sq = MeMp * MeMp;
        gamma           = Et / Mp;
        gamma_sq        = gamma * gamma;
        beta_sq         = 1.0 - 1.0 / gamma_sq;
        Te_max          = (2.0 * Me * beta_sq * gamma_sq);
        Te_max /= (1.0 + 2.0 * gamma * MeMp + MeMp_sq);
        tau = Ek / Mp;
    }
    R momentum() const { return M::sq(Et * Et - mc2 * mc2); }
};

// intpl1d (mqi_math.hpp:28-31)
template<typename T>
inline T
intpl1d(T x, T x0, T x1, T y0, T y1) {
    return (x1 == x0) ? y0 : y0 + (x - x0) * (y1 - y0) / (x1 - x0);
}

// ---------------------------------------------------------------- materials
// spr_default (mqi_patient_materials.hpp:415-450); returns float as the reference.
template<typename R, typename M>
R
spr(R rho_mass, R Ek, bool dbg) {
    R density_tmp = rho_mass * 1000.0;
    if (dbg) {
        if (std::abs(density_tmp - 1.0) < 1e-3) return 1.0;
    }
    if (density_tmp <= 0.26) {
        if (density_tmp < 0.0012) return 0.0;
        return intpl1d<R>(density_tmp, 0.0012, 0.26, 0.8815, 0.9925);
    } else {
        R rsp = 1.0123 - 3.386e-5 * Ek;
        rsp += 0.291 * (1.0 + M::pw(Ek, static_cast<R>(-0.3421))) *
               (M::pw(density_tmp, static_cast<R>(-0.7)) - 1.0);
        if (density_tmp >= 0.9) return rsp;
        return intpl1d<R>(density_tmp, 0.26, 0.9, 0.9925, rsp);
    }
}

// radiation_length_default (mqi_patient_materials.hpp:452-473)
template<typename R, typename M>
R
radiation_length(R rho_mass, R water_density, R radiation_length_water, bool dbg) {
    R radiation_length_mat = 0.0;
    R f                    = 0.0;
    R density              = rho_mass * 1000.0;
    if (dbg) {
        if (std::abs(density - 1.0) < 1e-3) return radiation_length_water;
    }
    if (density <= 0.26) {
        f = 0.9857 + 0.0085 * density;
    } else if (density > 0.26 && density <= 0.9) {
        f = 1.0446 - 0.2180 * density;
    } else if (density > 0.9) {
        f = 1.19 + 0.44 * M::ln((double) (density - 0.44));
    }
    radiation_length_mat = (water_density * radiation_length_water) / (density * 0.001 * f);
    return radiation_length_mat;
}

// hu_to_density (mqi_patient_materials.hpp:515-542), float as the reference.
inline float
hu_to_density(int16_t hu) {
    typedef float R;
    if (hu < -1000) {
        hu = -1000;
    } else if (hu > 2995) {
        hu = 2995;
    }
    R rho_mass = 0.0;
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
    R correction_factor = kDensityCorrection[hu + 1000];
    rho_mass *= correction_factor;
    rho_mass /= 1000.0;
    return rho_mass;
}

// ---------------------------------------------------------------- geometry
// grid3d (mqi_grid3d.hpp:31-890) for one voxel child.
template<typename R>
struct Grid {
    int             nx, ny, nz;
    std::vector<R>  xe, ye, ze;
    const float*    data;
    V3<R>           V000, V111;
    M3<R>           fwd, inv;
    V3<R>           trans;

    template<typename M>
    void init(int nx_, int ny_, int nz_, const float* x, const float* y, const float* z, const float* d) {
        nx = nx_;
        ny = ny_;
        nz = nz_;
        xe.assign(x, x + nx + 1);
        ye.assign(y, y + ny + 1);
        ze.assign(z, z + nz + 1);
        data = d;
        // calculate_bounding_box (mqi_grid3d.hpp:64-89).  The axis normals are
        // (L,0,0)/|L| = (1,0,0) exactly, so d . n100 is d.x bit for bit and
        // the slab tests below use the direction components directly.
        V000 = V3<R>(xe[0], ye[0], ze[0]);
        V111 = V3<R>(xe[nx], ye[ny], ze[nz]);
    }

    uint64_t ijk2cnb(const int32_t* c) const { return (uint64_t) (int32_t) (c[2] * nx * ny + c[1] * nx + c[0]); }
    bool     is_valid(const int32_t* c) const {
        if (c[0] < 0 || c[1] < 0 || c[2] < 0) return false;
        if (c[0] >= nx || c[1] >= ny || c[2] >= nz) return false;
        return true;
    }
    // get_volume (mqi_grid3d.hpp:471-478)
    R volume(uint64_t c) const {
        const uint64_t nxy = (uint64_t) (nx * ny);
        int            k = (int) (c / nxy), j = (int) ((c % nxy) / nx), i = (int) ((c % nxy) % nx);
        R              v = xe[i + 1] - xe[i];
        v *= ye[j + 1] - ye[j];
        v *= ze[k + 1] - ze[k];
        return v;
    }

    // index(p, dir): linear edge scan with tolerance (mqi_grid3d.hpp:746-844)
    static int32_t scan_axis(const std::vector<R>& e, int n, R p, R dir) {
        int32_t idx = 0;
        for (int ind = 0; ind < n; ind++) {
            if (std::abs(e[ind] - p) < (R) 1e-3f) {
                if (dir > 0) return ind;
                if (dir < 0) return ind - 1;
                return ind;
            } else if (std::abs(e[ind + 1] - p) < (R) 1e-3f) {
                if (dir > 0) return ind + 1;
                if (dir < 0) return ind;
                return ind;
            } else if (e[ind] - p < 0 && e[ind + 1] - p > 0) {
                return ind;
            } else {
                idx = -1;
            }
        }
        return idx;
    }
    void index(const V3<R>& p, const V3<R>& dir, int32_t* c) const {
        c[0] = scan_axis(xe, nx, p.x, dir.x);
        c[1] = scan_axis(ye, ny, p.y, dir.y);
        c[2] = scan_axis(ze, nz, p.z, dir.z);
    }

    // intersect(p, d, idx): exit distance from inside voxel idx; mutates d
    // (mqi_grid3d.hpp:490-626).  Returns dist; cell set to -1 on no-hit.
    static void axis_in(R& dcomp, R p, R v1, R v2, int32_t i, int32_t n, R& tmin, R& tmax) {
        const R tol = (R) 1e-3f;
        R       me  = dcomp;   // d . n_axis with a unit axis normal
        if (me * me > (R) 1e-7f) {
            if (me < 0) {
                if (std::abs(-(p - v1) / dcomp) < tol && i > 0) {
                    tmax = 1 / tol;
                } else {
                    tmax = -(p - v1) / dcomp;
                }
                tmin = -(p - v2) / dcomp;
            } else {
                tmin = (v1 - p) / dcomp;
                if (std::abs((v2 - p) / dcomp) < tol && i < n) {
                    tmax = 1 / tol;
                } else {
                    tmax = (v2 - p) / dcomp;
                }
            }
        } else {
            dcomp = 0;
            tmin  = -HUGE_VALF;
            tmax  = HUGE_VALF;
        }
    }
    R intersect_in(const V3<R>& p, V3<R>& d, int32_t* c) const {
        R tmin[3], tmax[3];
        axis_in(d.x, p.x, xe[c[0]], xe[c[0] + 1], c[0], nx, tmin[0], tmax[0]);
        axis_in(d.y, p.y, ye[c[1]], ye[c[1] + 1], c[1], ny, tmin[1], tmax[1]);
        axis_in(d.z, p.z, ze[c[2]], ze[c[2] + 1], c[2], nz, tmin[2], tmax[2]);
        R u_max;
        if (tmax[0] < tmax[1]) {
            u_max = (tmax[0] < tmax[2]) ? tmax[0] : tmax[2];
        } else {
            u_max = (tmax[1] < tmax[2]) ? tmax[1] : tmax[2];
        }
        if (u_max > 0) return u_max;
        c[0] = c[1] = c[2] = -1;
        return -1.0;
    }

    // intersect(p, d): entry from outside (mqi_grid3d.hpp:631-743); mutates d.
    static void axis_out(R& dcomp, R p, R lo, R hi, R& tmin, R& tmax) {
        R me = dcomp;
        if (me * me > (R) 1e-7f) {
            if (me > 0) {
                tmin = (lo - p) / dcomp;
                tmax = (hi - p) / dcomp;
            } else {
                tmax = (lo - p) / dcomp;
                tmin = (hi - p) / dcomp;
            }
        } else {
            dcomp = 0;
            tmin  = -HUGE_VALF;
            tmax  = HUGE_VALF;
        }
    }
    R intersect_out(const V3<R>& p, V3<R>& d, int32_t* c) const {
        if (p.x >= xe[0] && p.x <= xe[nx] && p.y >= ye[0] && p.y <= ye[ny] && p.z >= ze[0] &&
            p.z <= ze[nz]) {
            index(p, d, c);
            return 0;
        }
        c[0] = c[1] = c[2] = -1;
        R dist = -1.0;
        R tmin[3], tmax[3];
        axis_out(d.x, p.x, V000.x, V111.x, tmin[0], tmax[0]);
        axis_out(d.y, p.y, V000.y, V111.y, tmin[1], tmax[1]);
        axis_out(d.z, p.z, V000.z, V111.z, tmin[2], tmax[2]);
        R u_min, u_max;
        if (tmin[0] > tmin[1]) {
            u_min = (tmin[0] > tmin[2]) ? tmin[0] : tmin[2];
        } else {
            u_min = (tmin[1] > tmin[2]) ? tmin[1] : tmin[2];
        }
        if (tmax[0] < tmax[1]) {
            u_max = (tmax[0] < tmax[2]) ? tmax[0] : tmax[2];
        } else {
            u_max = (tmax[1] < tmax[2]) ? tmax[1] : tmax[2];
        }
        if ((u_min < u_max || std::abs(u_min - u_max) < (R) 1e-3f) && u_min >= 0 && u_max >= 0) {
            dist     = u_min;
            V3<R> on = p + d * dist;
            index(on, d, c);
        }
        return dist;
    }

    // index(vtx1, dir1, idx): incremental update (mqi_grid3d.hpp:847-877)
    static void step_axis(R v, R dir, const std::vector<R>& e, int32_t& i) {
        const R tol = (R) 1e-3f;
        if (dir < 0 && (std::abs(v - e[i]) < tol || v < e[i])) {
            i -= 1;
        } else if (dir > 0 && (std::abs(v - e[i + 1]) < tol || v > e[i + 1])) {
            i += 1;
        }
    }
    void index_update(const V3<R>& v, const V3<R>& d, int32_t* c) const {
        step_axis(v.x, d.x, xe, c[0]);
        step_axis(v.y, d.y, ye, c[1]);
        step_axis(v.z, d.z, ze, c[2]);
    }
};

// ---------------------------------------------------------------- track
// track_t (mqi_track.hpp:48-205)
template<typename R>
struct Track {
    Vtx<R>  v0, v1;
    R       dE       = 0;
    R       local_dE = 0;
    bool    stopped  = false;
    bool    primary  = true;
    R       its_dist = 0;
    int32_t cell[3]  = { 0, 0, 0 };

    void deposit(R e) { dE += e; }
    void local_deposit(R e) { local_dE += e; }
    void update_post_vertex_position(const R& len) { v1.pos = v0.pos + v0.dir * len; }
    void update_post_vertex_energy(const R& e) { v1.ke -= e; }
    void move() {
        v0       = v1;
        dE       = 0;
        local_dE = 0;
    }
    // update_post_vertex_direction (mqi_track.hpp:165-172)
    template<typename M>
    void update_post_vertex_direction(const R& theta, const R& phi) {
        M3<R> ml = M3<R>::template euler<M>(0, theta, phi);
        const V3<R> ref(0, 0, 1);
        V3<R>       dl = ml * ref;
        dl.template normalize<M>();
        M3<R> mg = M3<R>::template align<M>(ref, v1.dir);
        v1.dir   = mg * dl;
        v1.dir.template normalize<M>();
    }
};

// track_stack_t: LIFO with silent drop when full (mqi_track_stack.hpp:12-66)
template<typename R>
struct Stack {
    std::vector<Track<R>> t;
    int                   limit, idx = 0;
    explicit Stack(int lim) : t(lim), limit(lim) {}
    void push_secondary(const Track<R>& k) {
        if (idx < limit) t[idx++] = k;
    }
    void push_primary(const Track<R>& k) {
        t[0] = k;
        idx  = 1;
    }
    bool     empty() const { return idx == 0; }
    Track<R> pop() { return t[--idx]; }
};

// ---------------------------------------------------------------- physics
template<typename R, typename Rng, typename M, bool DBG>
struct Physics {
    Units<R>     u;
    const Grid<R>* g;
    Rng*         rng;
    const R      Tp_cut   = 0.5 * 1.0;                  // physics_list / interaction Tp_cut
    const R      T_cut    = DBG ? 0.08511 * 1.0 : 0.0815 * 1.0;   // mqi_interaction.hpp:24-28
    const float  max_step = 1.0;                        // fippel_physics::max_step (mm)
    // p_ionization_tabulated(0.1, 299.6, 0.5, ...) (mqi_fippel_physics.hpp:30-35)
    const R Ei = 0.1, Ef = 299.6, E_step = 0.5, Es = 13.9;
    // pp/po tables: Ek_min 0.5, Ek_max 300, dEk 0.5
    const R Ek_min = 0.5, Ek_max = 300.0, dEk = 0.5;

    R uni() { return rng->template uniform<R>(); }
    R nrm(R avg, R sig) { return rng->template normal<R>(avg, sig); }

    R rsp(R rho, R Ek) const { return spr<R, M>(rho, Ek, DBG); }

    // p_ionization_tabulated::cross_section (mqi_p_ionization.hpp:254-268)
    R cs_ion(const Rel<R, M>& rel, R rho) const {
        R cs = 0;
        if (rel.Ek >= Ei && rel.Ek <= Ef) {
            uint16_t idx0 = uint16_t((rel.Ek - Ei) / E_step);
            uint16_t idx1 = idx0 + 1;
            R        x0   = Ei + idx0 * E_step;
            R        x1   = x0 + E_step;
            cs            = intpl1d<R>(rel.Ek, x0, x1, kCsPIon[idx0], kCsPIon[idx1]);
        }
        cs *= rho;
        return cs;
    }
    // pp/po cross sections (mqi_pp_elastic.hpp:221-235, po_elastic:236-249, po_inelastic:143-155)
    R cs_tab(const float* tab, const Rel<R, M>& rel, R rho) const {
        R cs = 0;
        if (rel.Ek >= Ek_min && rel.Ek <= Ek_max) {
            uint16_t idx0 = uint16_t((rel.Ek - Ek_min) / dEk);
            uint16_t idx1 = idx0 + 1;
            R        x0   = Ek_min + idx0 * dEk;
            R        x1   = x0 + 0.5;
            cs            = intpl1d<R>(rel.Ek, x0, x1, tab[idx0], tab[idx1]);
        }
        cs *= rho;
        return cs;
    }
    // dEdx (mqi_p_ionization.hpp:271-286)
    R dEdx(const Rel<R, M>& rel) const {
        R pw = 0;
        if (rel.Ek >= Ei && rel.Ek <= Ef) {
            uint16_t idx0 = uint16_t((rel.Ek - Ei) / E_step);
            uint16_t idx1 = idx0 + 1;
            R        x0   = Ei + idx0 * E_step;
            R        x1   = x0 + E_step;
            pw            = intpl1d<R>(rel.Ek, x0, x1, kDedxRestricted[idx0], kDedxRestricted[idx1]);
        } else if (rel.Ek < Ei && rel.Ek > 0) {
            pw = kDedxRestricted[0];
        }
        return -1.0 * pw;
    }
    // energy_straggling (mqi_p_ionization.hpp:334-345)
    R straggling(const Rel<R, M>& rel, R step_length, R rho) const {
        R Te   = (rel.Te_max >= 0.08511) ? 0.08511 : rel.Te_max;
        R O_sq = u.two_pi_re2_mc2_h2o * rho / u.water_density * step_length;
        O_sq *= Te / rel.beta_sq * (1.0 - 0.5 * rel.beta_sq);
        return O_sq;
    }
    // energy_loss (mqi_p_ionization.hpp:297-331); the residual-range walk
    // stops at index 0 where the reference's uint16 counter would wrap.
    R energy_loss(const Rel<R, M>& rel, R rho, const R step_length) {
        R        L_w = step_length * rsp(rho, rel.Ek) * rho / u.water_density;
        uint16_t n   = uint16_t((rel.Ek - Ei) / E_step);
        R        x0  = Ei + n * E_step;
        R        x1  = x0 + E_step;
        if (x0 > rel.Ek) n -= 1;
        if (x1 < rel.Ek) n += 1;
        R r = intpl1d<R>(rel.Ek, x0, x1, kRangeSteps[n], kRangeSteps[n + 1]);
        if (r < L_w) return rel.Ek;
        r -= L_w;
        for (;;) {
            if (r >= kRangeSteps[n]) break;
            if (n == 0) break;
            --n;
            if (n == 0) break;
        }
        x0        = Ei + n * E_step;
        x1        = x0 + E_step;
        R dE_mean = rel.Ek - intpl1d<R>(r, kRangeSteps[n], kRangeSteps[n + 1], x0, x1);
        R dE_var  = straggling(rel, L_w, rho);
        R ret     = nrm(dE_mean, M::sq(dE_var));
        if (ret < 0) ret *= -1.0;
        return ret;
    }

    // along_step (mqi_p_ionization.hpp:348-420)
    void along_step(Track<R>& trk, const R len, R rho) {
        Rel<R, M> rel(trk.v0.ke, u.Mp);
        R         dE = energy_loss(rel, rho, len);
        R         r  = 1.0;
        if (dE >= trk.v0.ke) {
            r           = trk.v0.ke / dE;
            trk.stopped = true;
        }
        R P     = rel.momentum();
        R X0    = radiation_length<R, M>(rho, u.water_density, u.radiation_length_water, DBG);
        R th_sq = ((Es / P) * (Es / P) / rel.beta_sq) * len / X0;
        R th    = M::sq(th_sq);
        th      = nrm(0, M::sq((R) 2.0f) * th);
        if (th < 0) th *= -1.0;
        R phi = 2.0 * M_PI * uni();
        trk.template update_post_vertex_direction<M>(th, phi);
        trk.deposit(dE * r);
        trk.update_post_vertex_position(r * len);
        trk.update_post_vertex_energy(dE * r);
    }

    // daughter placement used by every secondary:
    // pos -> fwd * (pos - T) + T, dir -> fwd * dir (e.g. mqi_pp_elastic.hpp:190-196)
    V3<R> place(const V3<R>& p) const { return g->fwd * (p - g->trans) + g->trans; }
    V3<R> turn(const V3<R>& d) const { return g->fwd * d; }

    // p_ionization_tabulated::post_step (mqi_p_ionization.hpp:424-477)
    void post_ion(Track<R>& trk, Stack<R>& stk) {
        Rel<R, M> rel(trk.v1.ke, u.Mp);
        R         Te, n;
        while (true) {
            n  = uni();
            Te = T_cut * rel.Te_max;
            Te /= ((1.0 - n) * rel.Te_max + n * T_cut);
            if (uni() < 1.0 - rel.beta_sq * Te / rel.Te_max + Te * Te / (2.0 * rel.Et_sq)) break;
        }
        if (DBG) {
            Track<R> d(trk);
            d.dE      = Te;
            d.primary = false;
            d.v0.ke   = 0;
            d.v1.ke   = 0;
            d.stopped = false;
            d.v0.pos  = place(d.v0.pos);
            d.v0.dir  = turn(d.v0.dir);
            d.v1.pos  = place(d.v1.pos);
            d.v1.dir  = turn(d.v1.dir);
            stk.push_secondary(d);
        } else {
            trk.deposit(Te);
        }
        trk.update_post_vertex_energy(Te);
    }

    // pp_elastic_tabulated::post_step (mqi_pp_elastic.hpp:119-219)
    void post_pp(Track<R>& trk, Stack<R>& stk) {
        Rel<R, M> rel(trk.v1.ke, u.Mp);
        R         min_value = Tp_cut / rel.Ek;
        R         uu        = uni() * (1.0 - 2.0 * min_value) + min_value;
        R         E1        = rel.Et;
        R         dE        = rel.Ek * uu;
        R         E3        = (rel.Ek - dE) + u.Mp;
        R         E4        = dE + u.Mp;
        R         P1        = rel.momentum();
        R         P3        = M::sq(E3 * E3 - u.Mp_sq);
        R         P4        = M::sq(E4 * E4 - u.Mp_sq);
        R         cos_th3   = E1 * E3 - u.Mp_sq - u.Mp * (E1 - E3);
        cos_th3 /= (P1 * P3);
        R cos_th34 = E3 * E4 - E1 * u.Mp;
        cos_th34 /= (P3 * P4);
        if (cos_th3 > 1.0)
            cos_th3 = 1.0;
        else if (cos_th3 < -1.0)
            cos_th3 = -1.0;
        if (cos_th34 > 1.0)
            cos_th34 = 1.0;
        else if (cos_th34 < -1.0)
            cos_th34 = -1.0;
        R th3 = M::ac(cos_th3);
        R th4 = th3 - M::ac(cos_th34);
        R phi = 2.0 * M_PI * uni();
        trk.update_post_vertex_energy(dE);
        trk.template update_post_vertex_direction<M>(th3, phi);
        Track<R> d = trk;
        d.dE       = 0;
        d.local_dE = 0;
        d.primary  = false;
        d.v0.ke    = dE;
        d.v1.ke    = dE;
        d.stopped  = false;
        d.template update_post_vertex_direction<M>(th4, phi);
        d.v0.pos = place(d.v1.pos);
        d.v0.dir = turn(d.v1.dir);
        d.v1.pos = place(d.v1.pos);
        d.v1.dir = turn(d.v1.dir);
        stk.push_secondary(d);
    }

    // po_elastic::post_step (mqi_po_elastic.hpp:96-217)
    void post_poe(Track<R>& trk, Stack<R>& stk) {
        Rel<R, M> rel(trk.v1.ke, u.Mp);
        if (rel.Ek <= 5.5) {
            R dE = rel.Ek;
            if (DBG) {
                Track<R> d(trk);
                d.dE      = dE;
                d.primary = false;
                d.v0.ke   = dE;
                d.v1.ke   = 0;
                d.v0.pos  = place(d.v1.pos);
                d.v0.dir  = turn(d.v1.dir);
                d.v1.pos  = place(d.v1.pos);
                d.v1.dir  = turn(d.v1.dir);
                d.stopped = false;
                stk.push_secondary(d);
            } else {
                trk.local_deposit(dE);
            }
            trk.update_post_vertex_energy(dE);
            trk.stopped = true;
        } else {
            R Tp_avg = 0.65 * M::ex(-0.0013 * rel.Ek);
            Tp_avg -= 0.71 * M::ex(-0.0177 * rel.Ek);
            R Tp_max = (2.0 * u.Mo * rel.beta_sq * rel.gamma_sq);
            Tp_max /= (1.0 + 2.0 * rel.gamma * u.MoMp + u.MoMp_sq);
            R dE      = rng->template exponential<R, M>((R) (1.0 / Tp_avg), Tp_max);
            R E1      = rel.Ek * (rel.Ek + 2.0 * u.Mp);
            R E3      = (rel.Ek - dE) * (rel.Ek - dE + 2.0 * u.Mp);
            R cos_th3 = (E1 + E3 - dE * (dE + 2.0 * u.Mo)) / 2.0 / M::sq(E1 * E3);
            if (cos_th3 > 1.0) cos_th3 = 1.0;
            if (cos_th3 < -1.0) cos_th3 = -1.0;
            R th3 = M::ac(cos_th3);
            R phi = 2.0 * M_PI * uni();
            if (DBG) {
                Track<R> d(trk);
                d.dE      = dE;
                d.primary = false;
                d.v0.ke   = dE;
                d.v1.ke   = 0;
                d.stopped = false;
                d.v0.pos  = place(d.v0.pos);
                d.v0.dir  = turn(d.v0.dir);
                d.v1.pos  = place(d.v1.pos);
                d.v1.dir  = turn(d.v1.dir);
                stk.push_secondary(d);
            } else {
                trk.local_deposit(dE);
            }
            trk.update_post_vertex_energy(dE);
            trk.template update_post_vertex_direction<M>(th3, phi);
        }
    }

    // po_inelastic_tabulated::post_step (mqi_po_inelastic.hpp:158-288)
    void post_poi(Track<R>& trk, Stack<R>& stk) {
        const R Ek     = trk.v1.ke;
        R       Eb     = 5.0;   // E_bind
        R       Er     = Ek;
        const R E_mini = 2.0, E_ratio = 0.65;
        R       Prob_2nd, Prob_long, power;
        if (Ek <= 215 && Ek > 200) {
            Prob_2nd  = 0.78;
            Prob_long = Prob_2nd + (1 - Prob_2nd) * 0.9;
            power     = 0.4;
        } else if (Ek > 215) {
            Prob_2nd  = 0.78;
            Prob_long = Prob_2nd + (1 - Prob_2nd) * 1.0;
            power     = 0.4;
        } else if (Ek <= 200 && Ek > 150) {
            Prob_2nd  = 0.72;
            Prob_long = Prob_2nd + (1 - Prob_2nd) * 0.83;
            power     = 0.45;
        } else {
            Prob_2nd  = 0.7;
            Prob_long = Prob_2nd + (1 - Prob_2nd) * 0.83;
            power     = 0.52;
        }
        while ((Er - Eb) > E_mini) {
            Er -= Eb;
            R uu = uni();
            R dE = M::pw(uu, power) * (Er - E_mini) + E_mini;
            if (dE >= Er) dE = Er;
            Er -= dE;
            trk.update_post_vertex_energy(dE + Eb);
            R zeta = uni();
            if (zeta < Prob_2nd) {
                R cos_th = (2.0 * dE / Ek - 1.0) + 2.0 * (1 - dE / Ek) * uni();
                if (cos_th < -1) cos_th = -1;
                if (cos_th > 1) cos_th = 1;
                R        th  = M::ac(cos_th);
                R        phi = 2.0 * M_PI * uni();
                Track<R> d(trk);
                d.dE       = 0;
                d.local_dE = 0;
                d.primary  = false;
                d.v0.ke    = dE;
                d.v1.ke    = dE;
                d.stopped  = false;
                d.template update_post_vertex_direction<M>(th, phi);
                d.v0.pos = place(d.v1.pos);
                d.v0.dir = turn(d.v1.dir);
                d.v1.pos = place(d.v1.pos);
                d.v1.dir = turn(d.v1.dir);
                stk.push_secondary(d);
            } else if (zeta < Prob_long) {
                // long-range energy leaves the track
            } else {
                if (DBG) {
                    Track<R> d(trk);
                    d.dE      = dE;
                    d.primary = false;
                    d.v0.ke   = dE;
                    d.v1.ke   = 0;
                    d.stopped = false;
                    d.v0.pos  = place(d.v0.pos);
                    d.v0.dir  = turn(d.v0.dir);
                    d.v1.pos  = place(d.v1.pos);
                    d.v1.dir  = turn(d.v1.dir);
                    stk.push_secondary(d);
                }
            }
            Eb *= E_ratio;
        }
        trk.deposit(Er);
        trk.update_post_vertex_energy(Er);
        trk.stopped = true;
    }

    // last_step (mqi_p_ionization.hpp:481-490)
    void last_step(Track<R>& trk, R rho) {
        Rel<R, M> rel(trk.v0.ke, u.Mp);
        R         L_w = 0;
        if (trk.dE > 0) L_w = -trk.dE / dEdx(rel);
        R step_length = L_w * u.water_density / (rsp(rho, trk.v0.ke) * rho);
        trk.update_post_vertex_position(step_length);
    }

    // fippel_physics::stepping (mqi_fippel_physics.hpp:68-216)
    void stepping(Track<R>& trk, Stack<R>& stk, const R rho, const R d2b) {
        if (rho < 1.0e-7) {
            trk.update_post_vertex_position(d2b);
            return;
        } else if (rho > 99.9) {
            trk.stopped = true;
            return;
        }
        if (trk.v0.ke <= Tp_cut) {
            if (trk.v0.ke < 0) trk.v0.ke = 0;
            trk.deposit(trk.v0.ke);
            trk.update_post_vertex_energy(trk.v0.ke);
            last_step(trk, rho);
            trk.stopped = true;
            return;
        }
        Rel<R, M> rel(trk.v0.ke, u.Mp);
        R         cms = max_step;
        cms           = cms * rsp(rho, rel.Ek) * rho / u.water_density;
        R max_loss_energy = -1.0 * cms * dEdx(rel);
        R cs1[4]  = { cs_ion(rel, rho), cs_tab(kCsPPElastic, rel, rho), cs_tab(kCsPOElastic, rel, rho),
                      cs_tab(kCsPOInelastic, rel, rho) };
        R cs1_sum = cs1[0] + cs1[1] + cs1[2] + cs1[3];
        Rel<R, M> rel_de(trk.v0.ke - max_loss_energy, u.Mp);
        R cs2[4]  = { cs_ion(rel_de, rho), cs_tab(kCsPPElastic, rel_de, rho),
                      cs_tab(kCsPOElastic, rel_de, rho), cs_tab(kCsPOInelastic, rel_de, rho) };
        R cs2_sum = cs2[0] + cs2[1] + cs2[2] + cs2[3];
        R  cs_sum = (cs1_sum >= cs2_sum) ? cs1_sum : cs2_sum;
        R* cs     = (cs1_sum >= cs2_sum) ? cs1 : cs2;
        R  prob   = uni();
        R  mfp;
        if (sizeof(R) == 4)
            mfp = -1.0f * M::ln((float) prob) / cs_sum;
        else
            mfp = -1.0 * M::ln(prob) / cs_sum;
        R step_limit = cms * u.water_density / (rsp(rho, rel.Ek) * rho);
        const R tol  = (R) 1e-3f;
        if (d2b < mfp && d2b < step_limit) {
            along_step(trk, d2b, rho);
        } else if ((mfp < d2b || std::abs(mfp - d2b) < tol) &&
                   (mfp < step_limit || std::abs(mfp - step_limit) < tol)) {
            along_step(trk, mfp, rho);
            if (trk.v1.ke <= Tp_cut) return;
            R uu       = cs_sum * uni();
            trk.v1.dir = trk.v0.dir;
            if (uu < cs[0]) {
                post_ion(trk, stk);
            } else if (uu < (cs[0] + cs[1])) {
                post_pp(trk, stk);
            } else if (uu < (cs[0] + cs[1] + cs[2])) {
                post_poe(trk, stk);
            } else if (uu < (cs[0] + cs[1] + cs[2] + cs[3])) {
                post_poi(trk, stk);
            }
        } else {
            along_step(trk, step_limit, rho);
        }
    }
};

// ---------------------------------------------------------------- scoring
// dose_to_water / dose_to_water_square / LETd_weight1 / LETd_weight2
// (mqi_scorer_energy_deposit.hpp:43-139)
template<typename R, typename M, bool DBG>
double
dose_hit(const Track<R>& trk, uint64_t cnb, const Grid<R>& g) {
    R density = g.data[cnb];
    R volume  = g.volume(cnb);
    if (density < 1.0e-7) return 0.0;
    return (trk.dE + trk.local_dE) * 1.60218e-10 / (volume * density * spr<R, M>(density, trk.v0.ke, DBG));
}
template<typename R, typename M>
double
letd_hit(const Track<R>& trk, uint64_t cnb, const Grid<R>& g, bool numer) {
    R density = g.data[cnb];
    density *= 1000.0;
    double length = (trk.v1.pos.x - trk.v0.pos.x) * (trk.v1.pos.x - trk.v0.pos.x);
    length += (trk.v1.pos.y - trk.v0.pos.y) * (trk.v1.pos.y - trk.v0.pos.y);
    length += (trk.v1.pos.z - trk.v0.pos.z) * (trk.v1.pos.z - trk.v0.pos.z);
    length = M::sq(length);
    if (length <= 0) return 0.0;
    double let = trk.dE / length / density;
    if (let >= 25.0) return 0;
    return numer ? trk.dE * let : trk.dE * 1.0;
}

// LETt_weight1 / LETt_weight2 (mqi_scorer_energy_deposit.hpp:141-177)
template<typename R, typename M>
double
lett_hit(const Track<R>& trk, uint64_t cnb, const Grid<R>& g, bool numer) {
    R density = g.data[cnb];
    density *= 1000.0;
    double length = (trk.v1.pos.x - trk.v0.pos.x) * (trk.v1.pos.x - trk.v0.pos.x);
    length += (trk.v1.pos.y - trk.v0.pos.y) * (trk.v1.pos.y - trk.v0.pos.y);
    length += (trk.v1.pos.z - trk.v0.pos.z) * (trk.v1.pos.z - trk.v0.pos.z);
    length = M::sq(length);
    if (length <= 0) return 0.0;
    if (!numer) return length;
    double let = trk.dE / length / density;
    return length * let;
}
// energy_deposit (:14-20) and dose_to_medium (:80-92)
template<typename R>
double
edep_hit(const Track<R>& trk) {
    return trk.dE + trk.local_dE;
}
template<typename R>
double
dm_hit(const Track<R>& trk, uint64_t cnb, const Grid<R>& g) {
    R density = g.data[cnb];
    R volume  = g.volume(cnb);
    return trk.primary ? trk.dE * 1.60218e-13 * 1000.0 / (volume * density) : 0.0;
}

enum { SC_DOSE = 1, SC_LETD = 2, SC_STAT = 4, SC_LOG = 8, SC_LETT = 16, SC_EDEP = 32, SC_DM = 64 };

struct Scorers {
    int                               mask = 0;
    // 0 dose, 1 letd num, 2 letd den, 3 stat dose, 4 stat dose^2, 5 lett num,
    // 6 lett den, 7 energy deposit, 8 dose to medium; laid out as tps_env
    // does: dose, letd, lett, edep, dm, the statistics pair last
    std::vector<int>                  kinds;
    std::vector<std::vector<double>>  data;
    void init(int m, uint64_t nvox) {
        mask = m;
        kinds.clear();
        if (m & SC_DOSE) kinds.push_back(0);
        if (m & SC_LETD) {
            kinds.push_back(1);
            kinds.push_back(2);
        }
        if (m & SC_LETT) {
            kinds.push_back(5);
            kinds.push_back(6);
        }
        if (m & SC_EDEP) kinds.push_back(7);
        if (m & SC_DM) kinds.push_back(8);
        if (m & SC_STAT) {
            kinds.push_back(3);
            kinds.push_back(4);
        }
        data.assign(kinds.size(), std::vector<double>(nvox, 0.0));
    }
};

// The scorers' ROI ([0]) and the statistics ROI ([1]) set by oracle_set_roi:
// empty is DIRECT (roi_t(mqi::DIRECT, n), idx(v) = v), otherwise the CONTOUR
// runs of mask_reader::mask_to_roi (mqi_file_handler.hpp:177-217).
struct RoiRuns {
    std::vector<uint32_t> start, stride;
};
RoiRuns g_roi[2];

// roi_t::lower_bound_cpp (mqi_roi.hpp:147-166): the first run whose start
// exceeds v
inline int32_t
roi_lower_bound(const RoiRuns& r, uint32_t v) {
    int32_t first = 0, count = (int32_t) r.start.size();
    while (count > 0) {
        int32_t step = count / 2, it = first + step;
        if (r.start[it] <= v) {
            first = it + 1;
            count -= step + 1;
        } else {
            count = step;
        }
    }
    return first;
}

// idx(v) > 0 (scorers) and get_mask_idx(v) >= 0 (statistics pair) of a
// CONTOUR ROI (idx_contour, get_contour_idx, mqi_roi.hpp:71-84,129-138):
// the run before lower_bound holds v when v - start < stride.  For v before
// the first run get_contour_idx answers -1; idx_contour reads start_[-1]
// there, which is taken as outside as well.
inline bool
roi_contour_in(const RoiRuns& r, uint32_t v) {
    const int32_t c = roi_lower_bound(r, v) - 1;
    return c >= 0 && v - r.start[c] < r.stride[c];
}

struct StepLog {
    int64_t  cap = 0, n = 0;
    int64_t* cnb = nullptr;
    float*   rec = nullptr;
};

// ---------------------------------------------------------------- transport
// transport_particles_patient(_stat) for one history (mqi_transport.hpp:146-249)
template<typename R, typename Rng, typename M, bool DBG>
struct Transport {
    const Grid<R>*               g;    // the patient/phantom: the last child, the one with scorers
    std::vector<const Grid<R>*>  bl;   // beamline children before it (range shifter, aperture)
    Physics<R, Rng, M, DBG>   ph;
    Stack<R>                  stk;
    Scorers*                  sc;
    int                       kernel;
    StepLog*                  log;
    int64_t                   cur_hist = 0;

    Transport(const Grid<R>* grid, Rng* r, Scorers* s, int k, StepLog* l) :
        g(grid), stk(DBG ? 200 : 10), sc(s), kernel(k), log(l) {
        ph.g   = grid;
        ph.rng = r;
    }

    double hit(int kind, const Track<R>& t, uint64_t cnb) const {
        switch (kind) {
        case 0:
        case 3: return dose_hit<R, M, DBG>(t, cnb, *g);
        case 1: return letd_hit<R, M>(t, cnb, *g, true);
        case 2: return letd_hit<R, M>(t, cnb, *g, false);
        case 5: return lett_hit<R, M>(t, cnb, *g, true);
        case 6: return lett_hit<R, M>(t, cnb, *g, false);
        case 7: return edep_hit<R>(t);
        case 8: return dm_hit<R>(t, cnb, *g);
        default: {
            double d = dose_hit<R, M, DBG>(t, cnb, *g);
            return d * d;
        }
        }
    }
    void insert(int s, uint64_t slot, double value) {
        if (value <= 0) return;
        sc->data[s][slot] += value;
    }
    void score(const Track<R>& t, uint64_t cnb) {
        if (log) {
            if (log->n < log->cap) {
                log->cnb[log->n] = (int64_t) cnb;
                float* r         = log->rec + 13 * log->n;
                r[0]             = (float) t.its_dist;
                r[1]             = (float) t.v0.pos.x;
                r[2]             = (float) t.v0.pos.y;
                r[3]             = (float) t.v0.pos.z;
                r[4]             = (float) t.v0.dir.x;
                r[5]             = (float) t.v0.dir.y;
                r[6]             = (float) t.v0.dir.z;
                r[7]             = (float) t.v1.pos.x;
                r[8]             = (float) t.v1.pos.y;
                r[9]             = (float) t.v1.pos.z;
                r[10]            = (float) t.v0.ke;
                r[11]            = (float) t.dE;
                r[12]            = (float) cur_hist;
            }
            log->n++;
            return;
        }
        const int n  = (int) sc->kinds.size();
        const int32_t roi = (int32_t) (uint32_t) cnb;   // roi_t DIRECT: idx(v) = v
        // regular scorers go through roi_->idx(cnb) > 0, the statistics pair
        // through get_mask_idx(cnb) >= 0
        const bool in_reg  = g_roi[0].start.empty() ? roi > 0 : roi_contour_in(g_roi[0], (uint32_t) cnb);
        const bool in_stat = g_roi[1].start.empty() ? roi >= 0 : roi_contour_in(g_roi[1], (uint32_t) cnb);
        if (kernel == 2) {
            // each scorer once: regular scorers through idx(cnb) > 0, the two
            // stat scorers through get_mask_idx(cnb) >= 0 (what the _stat
            // kernel does; for fewer than three scorers also the plain kernel)
            for (int s = 0; s < n; ++s) {
                const bool stat = sc->kinds[s] == 3 || sc->kinds[s] == 4;
                if (stat ? in_stat : in_reg) insert(s, cnb, hit(sc->kinds[s], t, cnb));
            }
            return;
        }
        // first loop: scorers 0 .. n-3 (mqi_transport.hpp:204-214 / 343-352)
        for (int s = 0; s < n - 2; ++s)
            if (in_reg) insert(s, cnb, hit(sc->kinds[s], t, cnb));
        if (kernel == 0) {
            // plain kernel: every scorer again (mqi_transport.hpp:215-225)
            for (int s = 0; s < n; ++s)
                if (in_reg) insert(s, cnb, hit(sc->kinds[s], t, cnb));
        } else if (n >= 2) {
            // _stat kernel: the last two scorers through get_mask_idx >= 0 (:353-364);
            // with fewer than two scorers the uint8_t start index wraps and the loop is empty
            for (int s = n - 2; s < n; ++s)
                if (in_stat) insert(s, cnb, hit(sc->kinds[s], t, cnb));
        }
    }

    // The world's children in order, the patient last (mqi_transport.hpp:158-240):
    // into the child's frame, to its first voxel, steps until the track
    // leaves or stops, back to the world frame (also when the child is
    // missed), then the next child.  Only the patient scores.  Secondaries
    // are placed with the transform of the child they were made in.
    void run_history(const Vtx<R>& vtx) {
        Track<R> primary;
        primary.v0 = vtx;
        primary.v1 = vtx;
        stk.push_primary(primary);
        while (!stk.empty()) {
            Track<R> t  = stk.pop();
            const int nc = (int) bl.size() + 1;
            for (int k = 0; k < nc; ++k) {
                const Grid<R>* G       = k < (int) bl.size() ? bl[k] : g;
                const bool     patient = k == nc - 1;
                ph.g                   = G;
                t.v0.pos               = G->inv * (t.v0.pos - G->trans);
                t.v0.dir               = G->inv * (t.v0.dir);
                t.v0.dir.template normalize<M>();
                t.v1.pos = t.v0.pos;
                t.v1.dir = t.v0.dir;
                int32_t chk[3];
                G->index(t.v0.pos, t.v0.dir, chk);
                bool inside = true;
                if (!G->is_valid(chk)) {
                    int32_t c[3];
                    R       dist = G->intersect_out(t.v0.pos, t.v0.dir, c);
                    t.its_dist   = dist;
                    if (dist < 0) {
                        inside = false;
                    } else {
                        t.update_post_vertex_position(dist);
                        t.move();
                        G->index(t.v0.pos, t.v0.dir, t.cell);
                    }
                } else {
                    t.its_dist = 0.0;
                    t.cell[0]  = chk[0];
                    t.cell[1]  = chk[1];
                    t.cell[2]  = chk[2];
                }
                while (inside && G->is_valid(t.cell) && !t.stopped) {
                    uint64_t cnb = G->ijk2cnb(t.cell);
                    t.its_dist   = G->intersect_in(t.v0.pos, t.v0.dir, t.cell);
                    ph.stepping(t, stk, (R) G->data[cnb], t.its_dist);
                    if (t.its_dist < 0) break;
                    if (patient) score(t, cnb);
                    if (!t.stopped) {
                        G->index_update(t.v1.pos, t.v1.dir, t.cell);
                        t.move();
                    }
                }
                t.v0.pos = G->fwd * (t.v0.pos) + G->trans;
                t.v0.dir = G->fwd * (t.v0.dir);
                t.v1.pos = t.v0.pos;
                t.v1.dir = t.v0.dir;
            }
        }
    }
};

// ---------------------------------------------------------------- sampling
// beamlet::operator() with phsp_6d or phsp_6d_uniform + norm_1d/const_1d
// (mqi_beamlet.hpp:81-90, mqi_phsp6d.hpp:62-79, mqi_phsp6d_uniform.hpp:68-82,
// mqi_norm_1d.hpp:54-56): R = float.
template<typename R>
Vtx<R>
phsp_vertex(const mqi_spot& sp, R Ux, R Vx, R Uy, R Vy, R Uz, R ke) {
    R phsp[6];
    for (int k = 0; k < 6; ++k) phsp[k] = sp.mean[k];
    phsp[0] += (R) sp.sigma[0] * Ux;
    phsp[1] += (R) sp.sigma[1] * Uy;
    phsp[2] += (R) sp.sigma[2] * Uz;
    const R rho0 = sp.corr[0], rho1 = sp.corr[1];
    phsp[3] += (R) sp.sigma[3] * (rho0 * Ux + Vx * std::sqrt(1.0 - rho0 * rho0));
    phsp[4] += (R) sp.sigma[4] * (rho1 * Uy + Vy * std::sqrt(1.0 - rho1 * rho1));
    phsp[5] = -1.0 * std::sqrt(1.0 - phsp[3] * phsp[3] - phsp[4] * phsp[4]);
    M3<R> rot(sp.rotation[0], sp.rotation[1], sp.rotation[2], sp.rotation[3], sp.rotation[4],
              sp.rotation[5], sp.rotation[6], sp.rotation[7], sp.rotation[8]);
    V3<R> tr(sp.translation[0], sp.translation[1], sp.translation[2]);
    Vtx<R> v;
    v.ke  = ke;
    v.pos = rot * V3<R>(phsp[0], phsp[1], phsp[2]) + tr;
    v.dir = rot * V3<R>(phsp[3], phsp[4], phsp[5]);
    return v;
}

// Reference sampling sequence for one spot, as run_by_spot does it
// (mqi_tps_env.hpp:1537-1550): phsp_6d and norm_1d keep their
// std::normal_distribution objects (and cached variates) across histories.
void
sample_spot_std(uint64_t master_seed, uint32_t spot_index, const mqi_spot& sp, int64_t n,
                Vtx<float>* out, int32_t* seeds) {
    std::default_random_engine             beam_rng, tmp_rng;
    std::normal_distribution<float>        phsp_func(0, 1);
    std::uniform_real_distribution<float>  phsp_uni(-1.0, 1.0);   // phsp_6d_uniform (mqi_phsp6d_uniform.hpp:46)
    std::normal_distribution<float>        e_func(sp.energy, sp.sigma_energy);
    const bool                             uni = sp.phsp == MQI_PHSP_UNIFORM;
    auto draw = [&]() { return uni ? phsp_uni(beam_rng) : phsp_func(beam_rng); };
    beam_rng.seed((uint32_t) (master_seed + spot_index));
    tmp_rng.seed((uint32_t) (master_seed + spot_index));
    for (int64_t h = 0; h < n; ++h) {
        float Ux = draw();
        float Vx = draw();
        float Uy = draw();
        float Vy = draw();
        float Uz = draw();
        float ke = sp.energy_const ? sp.energy : e_func(beam_rng);
        out[h]   = phsp_vertex<float>(sp, Ux, Vx, Uy, Vy, Uz, ke);
        if (seeds) seeds[h] = (int32_t) tmp_rng();
    }
}

// The GPU's sampling for history h of a spot: one vertex stream per history.
template<typename R>
Vtx<R>
sample_history_philox(uint64_t master_seed, uint32_t scenario, uint32_t spot_index, uint32_t h,
                      const mqi_spot& sp) {
    mqi_rng::stream s;
    s.init(master_seed, h, spot_index, scenario, mqi_rng::STREAM_VERTEX);
    const bool uni  = sp.phsp == MQI_PHSP_UNIFORM;
    auto       draw = [&]() { return uni ? (R) (2.0f * s.uniform() - 1.0f) : (R) s.gauss(); };
    R          Ux   = draw();
    R          Vx   = draw();
    R          Uy   = draw();
    R          Vy   = draw();
    R          Uz   = draw();
    R ke = sp.energy;
    if (!sp.energy_const) {
        R z = (R) s.gauss();
        ke  = z * (R) sp.sigma_energy + (R) sp.energy;
    }
    return phsp_vertex<R>(sp, Ux, Vx, Uy, Vy, Uz, ke);
}

// ---------------------------------------------------------------- drivers
struct Job {
    int             prec, rng, math, dbg;
    uint64_t        master_seed;
    uint32_t        scenario;
    int             nx, ny, nz;
    const float *   xe, *ye, *ze, *density, *rot9, *t3;
    // histories: either explicit vertices ...
    int64_t         n_hist;
    const float*    vtx;
    const int32_t*  seeds;
    const uint32_t* spot_ids;
    // ... or spots
    const mqi_spot* spots;
    int             n_spots;
    int             mask, kernel, n_threads;
    double**        out;
    uint64_t*       tracked;
    StepLog*        log;
};

// Beamline children set by oracle_set_beamline (test set-up: one world at a
// time): dims, edges, densities and the fwd rotation + translation of each.
struct BeamlineDesc {
    std::vector<int>                dims;    // 3 per child
    std::vector<std::vector<float>> xe, ye, ze, dens;
    std::vector<float>              xf;      // 12 per child: fwd[9], translation[3]
};
BeamlineDesc g_beamline;

template<typename R, typename Rng, typename M, bool DBG>
void
run_job(const Job& J) {
    Grid<R> g;
    g.template init<M>(J.nx, J.ny, J.nz, J.xe, J.ye, J.ze, J.density);
    const int              nbl = (int) g_beamline.dims.size() / 3;
    std::vector<Grid<R>>   bgrids(nbl);
    for (int k = 0; k < nbl; ++k) {
        Grid<R>& b = bgrids[k];
        b.template init<M>(g_beamline.dims[3 * k], g_beamline.dims[3 * k + 1], g_beamline.dims[3 * k + 2],
                           g_beamline.xe[k].data(), g_beamline.ye[k].data(), g_beamline.ze[k].data(),
                           g_beamline.dens[k].data());
        const float* f = g_beamline.xf.data() + 12 * k;
        b.fwd          = M3<R>(f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8]);
        b.inv          = b.fwd.inverse();
        b.trans        = V3<R>(f[9], f[10], f[11]);
    }
    if (J.rot9) {
        g.fwd = M3<R>(J.rot9[0], J.rot9[1], J.rot9[2], J.rot9[3], J.rot9[4], J.rot9[5], J.rot9[6],
                      J.rot9[7], J.rot9[8]);
        g.inv = g.fwd.inverse();
    }
    if (J.t3) g.trans = V3<R>(J.t3[0], J.t3[1], J.t3[2]);
    const uint64_t nvox = (uint64_t) J.nx * J.ny * J.nz;

    // flatten the history list: (spot index, history-in-spot)
    std::vector<std::pair<uint32_t, uint64_t>> hist;
    std::vector<uint64_t>                      spot_base;
    int64_t                                    total = J.n_hist;
    std::vector<Vtx<float>>                    std_vtx;
    std::vector<int32_t>                       std_seeds;
    if (J.spots) {
        total = 0;
        for (int s = 0; s < J.n_spots; ++s) {
            spot_base.push_back(total);
            total += (int64_t) J.spots[s].histories;
        }
        if (J.rng == 0) {   // reference sampling, sequential per spot
            std_vtx.resize(total);
            std_seeds.resize(total);
            for (int s = 0; s < J.n_spots; ++s)
                sample_spot_std(J.master_seed, (uint32_t) s, J.spots[s], (int64_t) J.spots[s].histories,
                                std_vtx.data() + spot_base[s], std_seeds.data() + spot_base[s]);
        }
    }
    int nt = J.n_threads < 1 ? 1 : J.n_threads;
    if (J.log) nt = 1;
    std::vector<Scorers>  scs(nt);
    std::vector<uint64_t> counts(nt, 0);
    auto work = [&](int t) {
        scs[t].init(J.mask, nvox);
        Rng rng;
        Transport<R, Rng, M, DBG> tr(&g, &rng, &scs[t], J.kernel, J.log);
        for (auto& b : bgrids) tr.bl.push_back(&b);
        int64_t q = total / nt, rmd = total % nt;
        int64_t h0 = q * t + (t < rmd ? t : rmd);
        int64_t h1 = h0 + q + (t < rmd ? 1 : 0);
        int     sp = 0;
        for (int64_t h = h0; h < h1; ++h) {
            Vtx<R>   v;
            uint32_t spot = 0, hin = (uint32_t) h;
            int32_t  seed = 0;
            if (J.spots) {
                while (sp + 1 < J.n_spots && (int64_t) spot_base[sp + 1] <= h) ++sp;
                while (sp > 0 && (int64_t) spot_base[sp] > h) --sp;
                spot = (uint32_t) sp;
                hin  = (uint32_t) (h - (int64_t) spot_base[sp]);
                if (J.rng == 0) {
                    const Vtx<float>& f = std_vtx[h];
                    v.ke                = f.ke;
                    v.pos               = V3<R>(f.pos.x, f.pos.y, f.pos.z);
                    v.dir               = V3<R>(f.dir.x, f.dir.y, f.dir.z);
                    seed                = std_seeds[h];
                } else {
                    v = sample_history_philox<R>(J.master_seed, J.scenario, spot, hin, J.spots[sp]);
                }
            } else {
                const float* f = J.vtx + 7 * h;
                v.ke           = f[0];
                v.pos          = V3<R>(f[1], f[2], f[3]);
                v.dir          = V3<R>(f[4], f[5], f[6]);
                seed           = J.seeds ? J.seeds[h] : 0;
                spot           = J.spot_ids ? J.spot_ids[h] : 0;
            }
            rng.seed_history(J.master_seed, hin, spot, J.scenario, seed);
            tr.cur_hist = h;
            tr.run_history(v);
            counts[t]++;
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto& th : pool) th.join();
    uint64_t tot = 0;
    for (auto c : counts) tot += c;
    if (J.tracked) *J.tracked = tot;
    if (J.out && !J.log) {
        const int n_sc = (int) scs[0].kinds.size();
        for (int k = 0; k < n_sc; ++k) {
            if (!J.out[k]) continue;
            double* o = J.out[k];
            for (uint64_t i = 0; i < nvox; ++i) o[i] = 0.0;
            for (int t = 0; t < nt; ++t) {
                const std::vector<double>& d = scs[t].data[k];
                for (uint64_t i = 0; i < nvox; ++i) o[i] += d[i];
            }
        }
    }
}

template<typename R>
void
dispatch_prec(const Job& J) {
    if (J.rng == 0 && J.math == 0) {
        if (J.dbg) run_job<R, StdRng, LibmMath, true>(J); else run_job<R, StdRng, LibmMath, false>(J);
    } else if (J.rng == 0) {
        if (J.dbg) run_job<R, StdRng, PortMath, true>(J); else run_job<R, StdRng, PortMath, false>(J);
    } else if (J.math == 0) {
        if (J.dbg) run_job<R, PhiloxRng, LibmMath, true>(J); else run_job<R, PhiloxRng, LibmMath, false>(J);
    } else {
        if (J.dbg) run_job<R, PhiloxRng, PortMath, true>(J); else run_job<R, PhiloxRng, PortMath, false>(J);
    }
}

int
n_scorers(int mask) {
    return ((mask & SC_DOSE) ? 1 : 0) + ((mask & SC_LETD) ? 2 : 0) + ((mask & SC_STAT) ? 2 : 0) +
           ((mask & SC_LETT) ? 2 : 0) + ((mask & SC_EDEP) ? 1 : 0) + ((mask & SC_DM) ? 1 : 0);
}

}   // namespace orc

using namespace orc;

template<typename M>
static void
physics_probe(const float* ek, int64_t n, float rho, float* out) {
    Physics<float, PhiloxRng, M, false> ph;
    for (int64_t i = 0; i < n; ++i) {
        Rel<float, M> rel(ek[i], ph.u.Mp);
        float*        o = out + 10 * i;
        o[0]            = ph.cs_ion(rel, rho);
        o[1]            = ph.cs_tab(kCsPPElastic, rel, rho);
        o[2]            = ph.cs_tab(kCsPOElastic, rel, rho);
        o[3]            = ph.cs_tab(kCsPOInelastic, rel, rho);
        o[4]            = ph.dEdx(rel);
        o[5]            = rel.beta_sq;
        o[6]            = rel.gamma;
        o[7]            = rel.Et;
        o[8]            = rel.Te_max;
        o[9]            = rel.momentum();
    }
}


extern "C" {

// Beamline children of the next transports (n = 0: none): dims 3 per child,
// edges concatenated (nx+1, ny+1, nz+1 per child), densities concatenated,
// xf 12 per child (fwd rotation row-major, translation).
void
oracle_set_beamline(int n, const int* dims, const float* edges, const float* dens, const float* xf) {
    g_beamline = BeamlineDesc{};
    size_t eo = 0, dofs = 0;
    for (int k = 0; k < n; ++k) {
        const int nx = dims[3 * k], ny = dims[3 * k + 1], nz = dims[3 * k + 2];
        g_beamline.dims.insert(g_beamline.dims.end(), {nx, ny, nz});
        g_beamline.xe.emplace_back(edges + eo, edges + eo + nx + 1);
        eo += nx + 1;
        g_beamline.ye.emplace_back(edges + eo, edges + eo + ny + 1);
        eo += ny + 1;
        g_beamline.ze.emplace_back(edges + eo, edges + eo + nz + 1);
        eo += nz + 1;
        const size_t nv = (size_t) nx * ny * nz;
        g_beamline.dens.emplace_back(dens + dofs, dens + dofs + nv);
        dofs += nv;
        g_beamline.xf.insert(g_beamline.xf.end(), xf + 12 * k, xf + 12 * k + 12);
    }
}

// ROI of the scorers (target 0) or the statistics pair (target 1) for the
// next transports: mask = NULL is DIRECT; otherwise mask_to_roi's runs
// (mqi_file_handler.hpp:185-203): in voxel order a 1 opens a run, a 0 closes
// it, other values change nothing.  A run open at the last voxel is closed
// there (the reference leaves that stride unset).
void
oracle_set_roi(int target, int64_t nvox, const uint8_t* mask) {
    RoiRuns& r = g_roi[target];
    r          = RoiRuns{};
    if (!mask) return;
    bool started = false;
    for (int64_t v = 0; v < nvox; ++v) {
        if (mask[v] == 1 && !started) {
            started = true;
            r.start.push_back((uint32_t) v);
        }
        if (mask[v] == 0 && started) {
            started = false;
            r.stride.push_back((uint32_t) v - r.start.back());
        }
    }
    if (started) r.stride.push_back((uint32_t) nvox - r.start.back());
    if (r.start.empty()) {   // nothing scores: one empty run keeps the ROI CONTOUR
        r.start.push_back(0);
        r.stride.push_back(0);
    }
}

int
oracle_transport(int prec, int rng, int math, int debug, uint64_t master_seed, int nx, int ny, int nz,
                 const float* xe, const float* ye, const float* ze, const float* density,
                 const float* rot9, const float* t3, int64_t n_hist, const float* vtx,
                 const int32_t* seeds, const uint32_t* spot_ids, int mask, int kernel, int n_threads,
                 double** out, uint64_t* tracked, int64_t* log_cnb, float* log_rec, int64_t log_cap,
                 int64_t* log_n) {
    StepLog L;
    L.cap = log_cap;
    L.cnb = log_cnb;
    L.rec = log_rec;
    Job J{};
    J.prec        = prec;
    J.rng         = rng;
    J.math        = math;
    J.dbg         = debug;
    J.master_seed = master_seed;
    J.nx          = nx;
    J.ny          = ny;
    J.nz          = nz;
    J.xe          = xe;
    J.ye          = ye;
    J.ze          = ze;
    J.density     = density;
    J.rot9        = rot9;
    J.t3          = t3;
    J.n_hist      = n_hist;
    J.vtx         = vtx;
    J.seeds       = seeds;
    J.spot_ids    = spot_ids;
    J.mask        = (mask & SC_LOG) ? 0 : mask;
    J.kernel      = kernel;
    J.n_threads   = n_threads;
    J.out         = out;
    J.tracked     = tracked;
    J.log         = (mask & SC_LOG) ? &L : nullptr;
    if (prec == 64)
        dispatch_prec<double>(J);
    else
        dispatch_prec<float>(J);
    if (log_n) *log_n = L.n;
    return n_scorers(J.mask);
}

int
oracle_run_spots(int prec, int rng, int math, int debug, uint64_t master_seed, uint32_t scenario, int nx,
                 int ny, int nz, const float* xe, const float* ye, const float* ze, const float* density,
                 const mqi_spot* spots, int n_spots, int mask, int kernel, int n_threads, double** out,
                 uint64_t* tracked) {
    Job J{};
    J.prec        = prec;
    J.rng         = rng;
    J.math        = math;
    J.dbg         = debug;
    J.master_seed = master_seed;
    J.scenario    = scenario;
    J.nx          = nx;
    J.ny          = ny;
    J.nz          = nz;
    J.xe          = xe;
    J.ye          = ye;
    J.ze          = ze;
    J.density     = density;
    J.spots       = spots;
    J.n_spots     = n_spots;
    J.mask        = mask;
    J.kernel      = kernel;
    J.n_threads   = n_threads;
    J.out         = out;
    J.tracked     = tracked;
    if (prec == 64)
        dispatch_prec<double>(J);
    else
        dispatch_prec<float>(J);
    return n_scorers(mask);
}

void
oracle_sample_spot(int rng, int math, uint64_t master_seed, uint32_t scenario, uint32_t spot_index,
                   const mqi_spot* spot, int64_t n, float* vtx_out, int32_t* seed_out) {
    (void) math;
    std::vector<Vtx<float>> v(n);
    if (rng == 0) {
        sample_spot_std(master_seed, spot_index, *spot, n, v.data(), seed_out);
    } else {
        for (int64_t h = 0; h < n; ++h)
            v[h] = sample_history_philox<float>(master_seed, scenario, spot_index, (uint32_t) h, *spot);
    }
    for (int64_t h = 0; h < n; ++h) {
        float* o = vtx_out + 7 * h;
        o[0]     = v[h].ke;
        o[1]     = v[h].pos.x;
        o[2]     = v[h].pos.y;
        o[3]     = v[h].pos.z;
        o[4]     = v[h].dir.x;
        o[5]     = v[h].dir.y;
        o[6]     = v[h].dir.z;
    }
}

// coordinate_transform rotation = iec2dicom * couch * gantry * collimator
// (mqi_coordinate_transform.hpp:42-60); angles in degrees, float as the reference.
void
oracle_coordinate_rotation(int math, const float* angles4, float* out9) {
    const float deg2rad = M_PI / 180.0;
    M3<float>   r;
    if (math) {
        r = M3<float>::euler<PortMath>(angles4[3] * deg2rad, 0, 0) *
            M3<float>::euler<PortMath>(0, 0, angles4[2] * deg2rad) *
            M3<float>::euler<PortMath>(0, angles4[1] * deg2rad, 0) *
            M3<float>::euler<PortMath>(0, 0, angles4[0] * deg2rad);
    } else {
        r = M3<float>::euler<LibmMath>(angles4[3] * deg2rad, 0, 0) *
            M3<float>::euler<LibmMath>(0, 0, angles4[2] * deg2rad) *
            M3<float>::euler<LibmMath>(0, angles4[1] * deg2rad, 0) *
            M3<float>::euler<LibmMath>(0, 0, angles4[0] * deg2rad);
    }
    const float m[9] = { r.xx, r.xy, r.xz, r.yx, r.yy, r.yz, r.zx, r.zy, r.zz };
    for (int k = 0; k < 9; ++k) out9[k] = m[k];
}

void
oracle_hu_to_density(const int16_t* hu, int64_t n, float* out) {
    for (int64_t i = 0; i < n; ++i) out[i] = hu_to_density(hu[i]);
}

void
oracle_spr(int math, int debug, const float* rho, const float* ek, int64_t n, float* out) {
    for (int64_t i = 0; i < n; ++i)
        out[i] = math ? spr<float, PortMath>(rho[i], ek[i], debug) : spr<float, LibmMath>(rho[i], ek[i], debug);
}

void
oracle_radiation_length(int math, int debug, const float* rho, int64_t n, float* out) {
    Units<float> u;
    for (int64_t i = 0; i < n; ++i)
        out[i] = math ? radiation_length<float, PortMath>(rho[i], u.water_density, u.radiation_length_water, debug)
                      : radiation_length<float, LibmMath>(rho[i], u.water_density, u.radiation_length_water, debug);
}

void
oracle_physics_probe(int math, const float* ek, int64_t n, float rho, float* out) {
    if (math)
        physics_probe<PortMath>(ek, n, rho, out);
    else
        physics_probe<LibmMath>(ek, n, rho, out);
}

void
oracle_rotate_direction(int math, const float* dir, const float* theta, const float* phi, int64_t n,
                        float* out) {
    for (int64_t i = 0; i < n; ++i) {
        Track<float> t;
        t.v1.dir = V3<float>(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]);
        if (math)
            t.update_post_vertex_direction<PortMath>(theta[i], phi[i]);
        else
            t.update_post_vertex_direction<LibmMath>(theta[i], phi[i]);
        out[3 * i]     = t.v1.dir.x;
        out[3 * i + 1] = t.v1.dir.y;
        out[3 * i + 2] = t.v1.dir.z;
    }
}

void
oracle_grid_probe(int op, int nx, int ny, int nz, const float* xe, const float* ye, const float* ze,
                  const float* p, const float* d, const int32_t* cell_in, int64_t n, float* dist_out,
                  float* d_out, int32_t* cell_out) {
    Grid<float> g;
    g.init<LibmMath>(nx, ny, nz, xe, ye, ze, nullptr);
    for (int64_t i = 0; i < n; ++i) {
        V3<float> pp(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
        V3<float> dd(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
        int32_t   c[3] = { cell_in ? cell_in[3 * i] : 0, cell_in ? cell_in[3 * i + 1] : 0,
                           cell_in ? cell_in[3 * i + 2] : 0 };
        float     dist = 0;
        if (op == 0) {
            g.index(pp, dd, c);
        } else if (op == 1) {
            dist = g.intersect_in(pp, dd, c);
            if (dist < 0) {
                // intersect returns its_non with cell -1 (mqi_grid3d.hpp:498-502)
            }
        } else if (op == 2) {
            dist = g.intersect_out(pp, dd, c);
        } else {
            g.index_update(pp, dd, c);
        }
        dist_out[i]         = dist;
        d_out[3 * i]        = dd.x;
        d_out[3 * i + 1]    = dd.y;
        d_out[3 * i + 2]    = dd.z;
        cell_out[3 * i]     = c[0];
        cell_out[3 * i + 1] = c[1];
        cell_out[3 * i + 2] = c[2];
    }
}

// fn: 0 log, 1 exp, 2 pow(x,y), 3 sin, 4 cos, 5 acos, 6 sqrt (portable, double)
void
oracle_math_probe(int fn, const double* x, const double* y, int64_t n, double* out) {
    for (int64_t i = 0; i < n; ++i) {
        switch (fn) {
        case 0: out[i] = mqi_pm::pm_log(x[i]); break;
        case 1: out[i] = mqi_pm::pm_exp(x[i]); break;
        case 2: out[i] = mqi_pm::pm_pow(x[i], y[i]); break;
        case 3: out[i] = mqi_pm::pm_sin(x[i]); break;
        case 4: out[i] = mqi_pm::pm_cos(x[i]); break;
        case 5: out[i] = mqi_pm::pm_acos(x[i]); break;
        default: out[i] = mqi_pm::pm_sqrt(x[i]); break;
        }
    }
}

void
oracle_philox(const uint32_t* ctr, const uint32_t* key, int64_t n, uint32_t* out) {
    for (int64_t i = 0; i < n; ++i) {
        mqi_rng::u32x4 c;
        for (int k = 0; k < 4; ++k) c.v[k] = ctr[4 * i + k];
        mqi_rng::u32x4 o = mqi_rng::philox4x32_10(c, key[2 * i], key[2 * i + 1]);
        for (int k = 0; k < 4; ++k) out[4 * i + k] = o.v[k];
    }
}

}   // extern "C"
