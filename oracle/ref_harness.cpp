// Test infrastructure only: a C-ABI harness around the reference moqui CPU
// transport, compiled from the reference headers where they lie
// (g++ -I<reference root>, see oracle/Makefile) into oracle/_ref/.
// Nothing in the product links or loads this file; tests and the bench's
// cpu_baseline leg call it through ctypes as the checker.
//
// What it wraps (reference file:line):
//   mc::transport_particles_patient / _stat   moqui/kernel_functions/mqi_transport.hpp:113-389
//   mqi::grid3d index / intersect             moqui/base/mqi_grid3d.hpp:490-877
//   mqi::patient_material_t::hu_to_density    moqui/base/materials/mqi_patient_materials.hpp:515-542
//   mqi::spr_default / radiation_length_default  mqi_patient_materials.hpp:415-473
//   physics tables                            mqi_p_ionization.hpp:11-221, mqi_pp_elastic.hpp:11,
//                                             mqi_po_elastic.hpp:12, mqi_po_inelastic.hpp:11
//   beamlet / phsp_6d / norm_1d sampling      mqi_beamlet.hpp:81-90, distributions/mqi_phsp6d.hpp:62-79
//   per-spot seeding as in run_by_spot        moqui/base/environments/mqi_tps_env.hpp:1537-1550
//
// The one behavioural deviation from the reference sources: the member
// p_ionization_tabulated<float>::energy_loss is given an explicit
// specialisation below that stops the residual-range table walk at index 0.
// The reference loop (mqi_p_ionization.hpp:318-320) decrements a uint16_t
// past zero when the step starts below 0.6 MeV and the residual range falls
// under the first table entry; it then reads r_steps[65535] (out of bounds,
// undefined behaviour, a segfault on the CPU build).  Everywhere the reference
// is well defined the specialisation computes the same value.
#include <chrono>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include <moqui/kernel_functions/mqi_transport.hpp>
#include <moqui/base/materials/mqi_patient_materials.hpp>
#include <moqui/base/scorers/mqi_scorer_energy_deposit.hpp>
#include <moqui/base/mqi_distributions.hpp>
#include <moqui/base/mqi_beamlet.hpp>
#include <moqui/base/mqi_beamsource.hpp>

namespace mqi
{
template<>
float
p_ionization_tabulated<float>::energy_loss(const relativistic_quantities<float>& rel,
                                           material_t<float>*&                   mat,
                                           float                                 rho_mass,
                                           const float                           step_length,
                                           mqi_rng*                              rng) {
    float length_in_water =
      step_length * mat->compute_rsp_(rho_mass, rel.Ek) * rho_mass / this->units.water_density;
    uint16_t n  = uint16_t((rel.Ek - this->Ei) / this->E_step);
    float    x0 = this->Ei + n * this->E_step;
    float    x1 = x0 + this->E_step;
    if (x0 > rel.Ek) n -= 1;
    if (x1 < rel.Ek) n += 1;
    float r = mqi::intpl1d(rel.Ek, x0, x1, r_steps[n], r_steps[n + 1]);
    if (r < length_in_water) return rel.Ek;
    r -= length_in_water;
    for (;;) {   // same walk as the reference, but never below index 0
        if (r >= r_steps[n]) break;
        if (n == 0) break;
        --n;
        if (n == 0) break;
    }
    x0             = this->Ei + n * this->E_step;
    x1             = x0 + this->E_step;
    float dE_mean  = rel.Ek - mqi::intpl1d(r, r_steps[n], r_steps[n + 1], x0, x1);
    float dE_var   = this->energy_straggling(rel, mat, length_in_water, rho_mass);
    float ret      = mqi::mqi_normal(rng, dE_mean, mqi::mqi_sqrt(dE_var));
    if (ret < 0) ret *= -1.0;
    return ret;
}
}   // namespace mqi

namespace
{
typedef float R;

// ---- step logger used by the deterministic ray tests -----------------------
struct step_log {
    int64_t  cap = 0, n = 0;
    int64_t* cnb  = nullptr;   // voxel of the step
    float*   rec  = nullptr;   // 13 floats per step: its.dist, vtx0.pos(3), vtx0.dir(3), vtx1.pos(3), ke0, dE, history
};
thread_local step_log* g_log  = nullptr;
thread_local int64_t   g_hist = 0;

double
log_hit(const mqi::track_t<R>& trk, const mqi::cnb_t& cnb, mqi::grid3d<mqi::density_t, R>&, mqi::material_t<R>*&) {
    step_log* L = g_log;
    if (L && L->n < L->cap) {
        L->cnb[L->n] = (int64_t) cnb;
        float* r     = L->rec + 13 * L->n;
        r[0]         = trk.its.dist;
        r[1]         = trk.vtx0.pos.x;
        r[2]         = trk.vtx0.pos.y;
        r[3]         = trk.vtx0.pos.z;
        r[4]         = trk.vtx0.dir.x;
        r[5]         = trk.vtx0.dir.y;
        r[6]         = trk.vtx0.dir.z;
        r[7]         = trk.vtx1.pos.x;
        r[8]         = trk.vtx1.pos.y;
        r[9]         = trk.vtx1.pos.z;
        r[10]        = trk.vtx0.ke;
        r[11]        = trk.dE;
        r[12]        = (float) g_hist;
    }
    if (L) L->n++;
    return 0.0;
}

// Beamline children set by ref_set_beamline (test set-up, one world at a
// time): the reference's own node_t/grid3d, no scorers, placed before the
// phantom as tps_env::setup_world does (mqi_tps_env.hpp:729-753).
struct beamline_child {
    std::vector<float>              xe, ye, ze, dens;
    mqi::grid3d<mqi::density_t, R>* geo = nullptr;
    mqi::node_t<R>                  node;
};
std::vector<beamline_child*> g_beamline;

// ROIs set by ref_set_roi: [0] for the regular scorers, [1] for the
// statistics pair; unset is roi_t(mqi::DIRECT, nvox).  The runs are what
// mask_reader::mask_to_roi builds (mqi_file_handler.hpp:177-217; that file
// pulls in GDCM, so the loop is restated in ref_set_roi), the lookups are the
// reference's own roi_t.
struct roi_runs {
    bool                  set = false;
    std::vector<uint32_t> start, stride, acc;
};
roi_runs g_roi[2];

struct world_t {
    mqi::node_t<R>                     world;
    mqi::node_t<R>                     child;
    std::vector<mqi::node_t<R>*>       children;
    mqi::grid3d<mqi::density_t, R>*    geo = nullptr;
    std::vector<mqi::scorer<R>*>       scorers;
    std::vector<mqi::key_value*>       tables;
    std::vector<mqi::roi_t*>           rois;
    std::vector<uint32_t>              ones;
    mqi::roi_t*                        stat_roi = nullptr;
};

mqi::grid3d<mqi::density_t, R>*
make_grid(int nx, int ny, int nz, const float* xe, const float* ye, const float* ze,
          const float* density, const float* rot9, const float* t3) {
    auto* g = new mqi::grid3d<mqi::density_t, R>(xe, nx + 1, ye, ny + 1, ze, nz + 1);
    if (rot9) {
        mqi::mat3x3<R> m(rot9[0], rot9[1], rot9[2], rot9[3], rot9[4], rot9[5], rot9[6], rot9[7], rot9[8]);
        g->rotation_matrix_fwd = m;
        g->rotation_matrix_inv = m.inverse();
    }
    if (t3) g->translation_vector = mqi::vec3<R>(t3[0], t3[1], t3[2]);
    g->set_data(const_cast<float*>(density));
    return g;
}

enum { SC_DOSE = 1, SC_LETD = 2, SC_STAT = 4, SC_LOG = 8, SC_LETT = 16, SC_EDEP = 32, SC_DM = 64 };

void
add_scorer(world_t& w, const char* name, uint32_t nvox, mqi::fp_compute_hit<R> fp, mqi::roi_t* roi) {
    auto* s  = new mqi::scorer<R>(name, nvox, fp);
    auto* kv = new mqi::key_value[nvox];
    mqi::init_table(kv, nvox);
    s->data_ = kv;
    s->roi_  = roi;
    w.scorers.push_back(s);
    w.tables.push_back(kv);
}

// Scorers are laid out in the order tps_env::setup_world uses
// (mqi_tps_env.hpp:788-960): Dose, LETd numerator, LETd denominator,
// Dose_stat, DoseSquare_stat.  SC_LOG replaces them with a single logger
// whose ROI maps every voxel to 1 (so voxel 0 is logged too).
void
build_world(world_t& w, mqi::grid3d<mqi::density_t, R>* geo, int mask) {
    const uint32_t nvox = geo->get_nxyz().x * geo->get_nxyz().y * geo->get_nxyz().z;
    w.geo               = geo;
    auto* direct        = new mqi::roi_t(mqi::DIRECT, nvox);
    w.rois.push_back(direct);
    mqi::roi_t* roi[2]  = { direct, direct };
    for (int k = 0; k < 2; ++k)
        if (g_roi[k].set) {
            roi[k] = new mqi::roi_t(mqi::CONTOUR, nvox, (int32_t) g_roi[k].start.size(), g_roi[k].start.data(),
                                    g_roi[k].stride.data(), g_roi[k].acc.data());
            w.rois.push_back(roi[k]);
        }
    w.stat_roi = roi[1];
    if (mask & SC_LOG) {
        w.ones.assign(nvox, 1u);
        auto* ind = new mqi::roi_t(mqi::INDIRECT, nvox, nvox, w.ones.data());
        w.rois.push_back(ind);
        add_scorer(w, "log", nvox, log_hit, ind);
    } else {
        if (mask & SC_DOSE) add_scorer(w, "Dose", nvox, mqi::dose_to_water<R>, roi[0]);
        if (mask & SC_LETD) {
            add_scorer(w, "LETd_numer", nvox, mqi::LETd_weight1<R>, roi[0]);
            add_scorer(w, "LETd_denom", nvox, mqi::LETd_weight2<R>, roi[0]);
        }
        if (mask & SC_LETT) {
            add_scorer(w, "LETt_numer", nvox, mqi::LETt_weight1<R>, roi[0]);
            add_scorer(w, "LETt_denom", nvox, mqi::LETt_weight2<R>, roi[0]);
        }
        if (mask & SC_EDEP) add_scorer(w, "Edep", nvox, mqi::energy_deposit<R>, roi[0]);
        if (mask & SC_DM) add_scorer(w, "Dm", nvox, mqi::dose_to_medium<R>, roi[0]);
        if (mask & SC_STAT) {
            add_scorer(w, "Dose_stat", nvox, mqi::dose_to_water<R>, roi[1]);
            add_scorer(w, "DoseSquare_stat", nvox, mqi::dose_to_water_square<R>, roi[1]);
        }
    }
    w.child.geo        = geo;
    w.child.n_scorers  = (uint16_t) w.scorers.size();
    w.child.scorers    = w.scorers.data();
    w.child.n_children = 0;
    w.children.clear();
    for (beamline_child* b : g_beamline) w.children.push_back(&b->node);
    w.children.push_back(&w.child);
    w.world.geo        = nullptr;
    w.world.n_scorers  = 0;
    w.world.n_children = (uint16_t) w.children.size();
    w.world.children   = w.children.data();
}

void
free_world(world_t& w) {
    for (auto* s : w.scorers) {
        s->data_ = nullptr;
        delete s;
    }
    for (auto* t : w.tables) delete[] t;
    for (auto* r : w.rois) delete r;
}

// ROI of the scorers (target 0) or the statistics pair (target 1) of the
// next transports; mask = NULL restores DIRECT.  mask_to_roi's loop: in voxel
// order a 1 opens a run, a 0 closes it and appends its stride and the
// accumulated stride.  Returns -1 for a mask whose last run is still open at
// the last voxel (the reference leaves that stride unset) or that has no run.
extern "C" int
ref_set_roi(int target, int64_t nvox, const uint8_t* mask) {
    roi_runs& r = g_roi[target];
    r           = roi_runs{};
    if (!mask) return 0;
    bool started = false;
    for (int64_t v = 0; v < nvox; ++v) {
        if (mask[v] == 1 && !started) {
            started = true;
            r.start.push_back((uint32_t) v);
        }
        if (mask[v] == 0 && started) {
            started = false;
            r.stride.push_back((uint32_t) v - r.start.back());
            r.acc.push_back((r.acc.empty() ? 0u : r.acc.back()) + r.stride.back());
        }
    }
    if (started || r.start.empty()) {
        r = roi_runs{};
        return -1;
    }
    r.set = true;
    return 0;
}

extern "C" void
ref_set_beamline(int n, const int* dims, const float* edges, const float* dens, const float* xf) {
    for (beamline_child* b : g_beamline) {
        delete b->geo;
        delete b;
    }
    g_beamline.clear();
    size_t eo = 0, dofs = 0;
    for (int k = 0; k < n; ++k) {
        auto*     b  = new beamline_child;
        const int nx = dims[3 * k], ny = dims[3 * k + 1], nz = dims[3 * k + 2];
        b->xe.assign(edges + eo, edges + eo + nx + 1);
        eo += nx + 1;
        b->ye.assign(edges + eo, edges + eo + ny + 1);
        eo += ny + 1;
        b->ze.assign(edges + eo, edges + eo + nz + 1);
        eo += nz + 1;
        b->dens.assign(dens + dofs, dens + dofs + (size_t) nx * ny * nz);
        dofs += (size_t) nx * ny * nz;
        const float* f = xf + 12 * k;
        b->geo         = make_grid(nx, ny, nz, b->xe.data(), b->ye.data(), b->ze.data(), b->dens.data(), f, f + 9);
        b->node.geo        = b->geo;
        b->node.n_scorers  = 0;
        b->node.scorers    = nullptr;
        b->node.n_children = 0;
        b->node.children   = nullptr;
        g_beamline.push_back(b);
    }
}

mqi::patient_material_t<R>*
material() {
    static mqi::patient_material_t<R>* m = new mqi::patient_material_t<R>();
    return m;
}

}   // namespace

extern "C" {

int
ref_physics_debug() {
#ifdef __PHYSICS_DEBUG__
    return 1;
#else
    return 0;
#endif
}

// Physics and calibration tables as compiled: 600 floats each, then 3996.
void
ref_tables(float* cs_p_ion, float* dedx, float* range, float* cs_pp, float* cs_po_e, float* cs_po_i,
           float* dens_corr) {
    std::memcpy(cs_p_ion, mqi::cs_p_ion_table, 600 * sizeof(float));
    std::memcpy(dedx, mqi::restricted_stopping_power_table, 600 * sizeof(float));
    std::memcpy(range, mqi::range_steps, 600 * sizeof(float));
    std::memcpy(cs_pp, mqi::cs_pp_e_g4_table, 600 * sizeof(float));
    std::memcpy(cs_po_e, mqi::cs_po_e_g4_table, 600 * sizeof(float));
    std::memcpy(cs_po_i, mqi::cs_po_i_g4_table, 600 * sizeof(float));
    std::memcpy(dens_corr, mqi::density_correction, 3996 * sizeof(float));
}

void
ref_hu_to_density(const int16_t* hu, int64_t n, float* out) {
    for (int64_t i = 0; i < n; ++i) out[i] = material()->hu_to_density(hu[i]);
}

void
ref_spr(const float* rho, const float* ek, int64_t n, float* out) {
    for (int64_t i = 0; i < n; ++i) out[i] = mqi::spr_default<R>(rho[i], ek[i]);
}

void
ref_radiation_length(const float* rho, int64_t n, float* out) {
    mqi::physics_constants<R> u;
    for (int64_t i = 0; i < n; ++i)
        out[i] = mqi::radiation_length_default<R>(rho[i], u.water_density, u.radiation_length_water);
}

// Per energy: cs_ion, cs_pp, cs_po_e, cs_po_i (each at rho), dEdx, then the
// relativistic quantities beta_sq, gamma, Et, Te_max, momentum.
void
ref_physics_probe(const float* ek, int64_t n, float rho, float* out) {
    mqi::fippel_physics<R> f;
    mqi::material_t<R>*    mat = material();
    for (int64_t i = 0; i < n; ++i) {
        mqi::relativistic_quantities<R> rel(ek[i], f.units.Mp);
        float*                          o = out + 10 * i;
        o[0]                              = f.p_ion.cross_section(rel, mat, rho);
        o[1]                              = f.pp_e.cross_section(rel, mat, rho);
        o[2]                              = f.po_e.cross_section(rel, mat, rho);
        o[3]                              = f.po_i.cross_section(rel, mat, rho);
        o[4]                              = f.p_ion.dEdx(rel, mat);
        o[5]                              = rel.beta_sq;
        o[6]                              = rel.gamma;
        o[7]                              = rel.Et;
        o[8]                              = rel.Te_max;
        o[9]                              = rel.momentum();
    }
}

// track_t::update_post_vertex_direction (mqi_track.hpp:165-172) on n inputs:
// in dir[3n], theta[n], phi[n]; out dir[3n].
void
ref_rotate_direction(const float* dir, const float* theta, const float* phi, int64_t n, float* out) {
    for (int64_t i = 0; i < n; ++i) {
        mqi::track_t<R> t;
        t.vtx1.dir = mqi::vec3<R>(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]);
        t.update_post_vertex_direction(theta[i], phi[i]);
        out[3 * i]     = t.vtx1.dir.x;
        out[3 * i + 1] = t.vtx1.dir.y;
        out[3 * i + 2] = t.vtx1.dir.z;
    }
}

// grid3d probes.  op 0: index(p,dir) scan -> cell; op 1: intersect(p,d,cell)
// -> dist (+ mutated d); op 2: intersect(p,d) from outside -> dist, cell,
// mutated d; op 3: index(vtx1,dir1,cell) incremental update -> cell.
// io layout per ray: in p[3], d[3], cell[3] (int as float bits not used:
// cells are passed separately); out dist, d[3], cell[3].
void
ref_grid_probe(int op, int nx, int ny, int nz, const float* xe, const float* ye, const float* ze,
               const float* p, const float* d, const int32_t* cell_in, int64_t n, float* dist_out,
               float* d_out, int32_t* cell_out) {
    mqi::grid3d<mqi::density_t, R> g(xe, nx + 1, ye, ny + 1, ze, nz + 1);
    for (int64_t i = 0; i < n; ++i) {
        mqi::vec3<R>           pp(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
        mqi::vec3<R>           dd(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
        mqi::vec3<mqi::ijk_t>  c(cell_in ? cell_in[3 * i] : 0, cell_in ? cell_in[3 * i + 1] : 0,
                                cell_in ? cell_in[3 * i + 2] : 0);
        float                  dist = 0;
        if (op == 0) {
            c = g.index(pp, dd);
        } else if (op == 1) {
            mqi::intersect_t<R> its = g.intersect(pp, dd, c);
            dist                    = its.dist;
        } else if (op == 2) {
            mqi::intersect_t<R> its = g.intersect(pp, dd);
            dist                    = its.dist;
            c                       = its.cell;
        } else {
            g.index(pp, dd, c);
        }
        dist_out[i]         = dist;
        d_out[3 * i]        = dd.x;
        d_out[3 * i + 1]    = dd.y;
        d_out[3 * i + 2]    = dd.z;
        cell_out[3 * i]     = c.x;
        cell_out[3 * i + 1] = c.y;
        cell_out[3 * i + 2] = c.z;
    }
}

// Vertex sampling for one spot, as run_by_spot does it
// (mqi_tps_env.hpp:1537-1550): beam_rng and tmp_rng are both seeded with
// master_seed + spot; each history draws a vertex from the beamlet
// (phsp_6d Gaussian fluence, norm_1d or const_1d energy, coordinate
// transform) and a transport seed from tmp_rng.
// mean6/sigma6: x,y,z,x',y',z' (z' mean and sigma ignored by phsp_6d).
void
ref_sample_spot(float energy, float sigma_e, int const_energy, int uniform, const float* mean6,
                const float* sigma6, const float* corr2, const float* angles4, const float* iso3,
                uint32_t master_seed, uint32_t spot, int64_t n, float* vtx_out, int32_t* seed_out) {
    std::array<R, 6> m, s;
    std::array<R, 2> r = { corr2 ? corr2[0] : 0.0f, corr2 ? corr2[1] : 0.0f };
    for (int k = 0; k < 6; ++k) {
        m[k] = mean6[k];
        s[k] = sigma6[k];
    }
    mqi::phsp_6d<R>         gauss(m, s, r);
    mqi::phsp_6d_uniform<R> flat(m, s, r);
    mqi::pdf_Md<R, 6>*      fluence = uniform ? (mqi::pdf_Md<R, 6>*) &flat : (mqi::pdf_Md<R, 6>*) &gauss;
    mqi::pdf_Md<R, 1>*  e;
    if (const_energy)
        e = new mqi::const_1d<R>({ energy }, { sigma_e });
    else
        e = new mqi::norm_1d<R>({ energy }, { sigma_e });
    std::array<R, 4>             ang = { angles4[0], angles4[1], angles4[2], angles4[3] };
    mqi::vec3<R>                 iso(iso3[0], iso3[1], iso3[2]);
    mqi::coordinate_transform<R> pc(ang, iso);
    mqi::beamlet<R>              bl(e, fluence);
    bl.set_coordinate_transform(pc);
    std::default_random_engine beam_rng, tmp_rng;
    beam_rng.seed(master_seed + spot);
    tmp_rng.seed(master_seed + spot);
    for (int64_t h = 0; h < n; ++h) {
        mqi::vertex_t<R> v = bl(&beam_rng);
        float*           o = vtx_out + 7 * h;
        o[0]               = v.ke;
        o[1]               = v.pos.x;
        o[2]               = v.pos.y;
        o[3]               = v.pos.z;
        o[4]               = v.dir.x;
        o[5]               = v.dir.y;
        o[6]               = v.dir.z;
        seed_out[h]        = tmp_rng();
    }
    delete e;
}

// Transport n histories through one voxel child of the world with the
// reference kernel on the CPU.  Each history re-seeds the thread RNG with
// its transport seed first, which is what curand_init(transport_seed[i])
// does per history in the CUDA build (mqi_transport.hpp:147-149).
// kernel 0: transport_particles_patient, 1: transport_particles_patient_stat.
// out[k] receives the raw table value of the k-th scorer in the layout of
// build_world (nvox doubles each).  With SC_LOG, log_cnb/log_rec receive up
// to log_cap steps and *log_n the number of logged steps.
int
ref_transport(int nx, int ny, int nz, const float* xe, const float* ye, const float* ze,
              const float* density, const float* rot9, const float* t3, int64_t n_hist,
              const float* vtx, const int32_t* seeds, const uint32_t* spot_ids, int mask, int kernel,
              int n_threads, double** out, uint64_t* tracked, int64_t* log_cnb, float* log_rec,
              int64_t log_cap, int64_t* log_n, double* seconds) {
    if (n_threads < 1) n_threads = 1;
    if (mask & SC_LOG) n_threads = 1;
    auto*          geo  = make_grid(nx, ny, nz, xe, ye, ze, density, rot9, t3);
    const uint32_t nvox = (uint32_t) nx * ny * nz;
    std::vector<world_t> worlds(n_threads);
    for (auto& w : worlds) build_world(w, geo, mask);
    const int n_sc = (int) worlds[0].scorers.size();
    std::vector<uint64_t> counts(n_threads, 0);
    step_log              L;
    L.cap = log_cap;
    L.cnb = log_cnb;
    L.rec = log_rec;

    auto work = [&](int t) {
        if (mask & SC_LOG) g_log = &L;
        mqi::thrd_t th;
        int64_t     q = n_hist / n_threads, rmd = n_hist % n_threads;
        int64_t     h0 = q * t + (t < rmd ? t : rmd);
        int64_t     h1 = h0 + q + (t < rmd ? 1 : 0);
        uint32_t    tracked_local = 0;
        for (int64_t h = h0; h < h1; ++h) {
            th.rnd_generator.seed((uint32_t) seeds[h]);
            g_hist = h;
            mqi::vertex_t<R> v;
            v.ke  = vtx[7 * h];
            v.pos = mqi::vec3<R>(vtx[7 * h + 1], vtx[7 * h + 2], vtx[7 * h + 3]);
            v.dir = mqi::vec3<R>(vtx[7 * h + 4], vtx[7 * h + 5], vtx[7 * h + 6]);
            uint32_t  sid    = spot_ids ? spot_ids[h] : 0;
            uint32_t* sidp   = spot_ids ? &sid : nullptr;
            int32_t   seed_h = seeds[h];
            if (kernel == 1)
                mc::transport_particles_patient_stat<R>(&th, &worlds[t].world, &v, material(), 1,
                                                        &tracked_local, &seed_h, sidp, true, 1, 0);
            else
                mc::transport_particles_patient<R>(&th, &worlds[t].world, &v, material(), 1,
                                                   &tracked_local, &seed_h, sidp, true, 1, 0);
        }
        counts[t] = tracked_local;
        g_log     = nullptr;
    };
    auto                     t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> pool;
    for (int t = 1; t < n_threads; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto& th : pool) th.join();
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    uint64_t tot = 0;
    for (auto c : counts) tot += c;
    if (tracked) *tracked = tot;
    if (log_n) *log_n = L.n;
    if (out && !(mask & SC_LOG)) {
        for (int k = 0; k < n_sc; ++k) {
            if (!out[k]) continue;
            double* o = out[k];
            for (uint32_t i = 0; i < nvox; ++i) o[i] = 0.0;
            // the statistics pair of a CONTOUR ROI is keyed by the compact
            // index get_mask_idx gave; get_org_idx maps it back to the voxel
            const bool compact = (mask & SC_STAT) && k >= n_sc - 2 && g_roi[1].set;
            for (auto& w : worlds) {
                const mqi::key_value* kv = w.tables[k];
                for (uint32_t i = 0; i < nvox; ++i)
                    if (kv[i].key1 != mqi::empty_pair) o[compact ? (uint32_t) w.stat_roi->get_org_idx(i) : i] += kv[i].value;
            }
        }
    }
    for (auto& w : worlds) free_world(w);
    delete geo;   // ~grid3d frees neither edges nor data; the density stays the caller's
    return n_sc;
}

}   // extern "C"
