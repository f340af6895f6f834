// ct_env: the tps_env environment without DICOM, on the C ABI.
#ifndef MQI_HIP_CT_ENV_HPP
#define MQI_HIP_CT_ENV_HPP

#include <array>
#include <string>
#include <vector>

#include "mqi_hip_beamline.hpp"
#include "mqi_hip_env.hpp"
#include "mqi_hip_io.hpp"

namespace mqi_hip
{

// The scale tps_env applies to every written scorer: particles_per_history *
// rbe * n_fractions evaluated in float (float * float * int, the members'
// types, mqi_tps_env.hpp:117,133-134 and :1873-1889), then widened.
inline double
output_scale(float particles_per_history, float rbe, int n_fractions) {
    const float s = particles_per_history * rbe * n_fractions;
    return (double) s;
}

// tps_env without DICOM (mqi_tps_env.hpp): the inputs of its parameter file
// that do not need GDCM, under the same names, and its world, run and output
// steps on a MetaImage CT with a caller-made beam source.
//   CTVolumeName          ct_path: MetaImage CT, float32 HU after the header
//                         (read_ct_image's mha branch, :615-701)
//   X/Y/ZShift, DensityScaling   volume_shift, density_scale (:253-257, :475-477, :766-769)
//   Scorer                scorer_names: Dose, LETd, LETt or Dij (Dij alone, :209-217)
//   StoppingStatistics, StoppingCriteria, StatThreshold, SaveStoppingStatistics
//                         the statistics pair and its stopping rounds (:281-302)
//   ScoringMask / Mask, StatROIMaskFilename   mask_files, stat_mask_files (:224-232, :290)
//   OutputDir, OutputFormat (raw, mhd, mha; Dij as npz)   output_dir, output_format
//   ParticlesPerHistory, RBE, NumberOfFraction   the output scale (:201-203)
// Beamline objects (range shifter, aperture) are built by the caller
// (mqi_hip_beamline.hpp) and placed before the patient, as setup_world does
// (:713-753); the beam source is the caller's beamlet list (for a machine
// file and a spot map: pbs_model::create_beamsource).  Outputs are named
// <beam_name>_<child index>_<scorer name> as save_reshaped_files and
// save_sparse_file name them (:1840-1930): the patient's child index is the
// number of beamline objects, the scorer names are tps_env's (Dose / Dij,
// LETd_numer, LETd_denom, LETt_numer, LETt_denom, Dose_stat,
// DoseSquare_stat; the statistics pair only with SaveStoppingStatistics).
class ct_env : public x_environment
{
public:
    std::string                ct_path;
    std::array<float, 3>       volume_shift = {0.0f, 0.0f, 0.0f};
    std::vector<std::string>   scorer_names = {"Dose"};
    bool                       record_statistics = false;
    float                      stat_criteria     = -1.0f;
    float                      stat_threshold    = 0.0f;
    bool                       save_statistics   = false;
    std::vector<std::string>   mask_files, stat_mask_files;
    std::vector<beamline_grid> beamline;
    std::vector<mqi_spot>      beamlets;
    std::string                output_dir, output_format = "raw", beam_name = "beam";
    float                      particles_per_history = 1.0f, rbe = 1.1f;
    int                        n_fractions           = 1;
    uint64_t                   random_seed           = 0;
    // -1: the kernel tps_env runs (:1118-1135): transport_particles_patient_stat
    // with the statistics pair, transport_particles_patient otherwise (which
    // scores scorers 0 .. n-3 twice when there are three or more,
    // mqi_transport.hpp:204-225); MQI_LOOPS_ONCE: each scorer once per step
    int                        scoring_loops = -1;

    void
    setup_world() override {
        io::ct_image ct = io::read_mha_ct(ct_path, volume_shift.data());
        xe              = std::move(ct.xe);
        ye              = std::move(ct.ye);
        ze              = std::move(ct.ze);
        hu              = std::move(ct.hu);
    }

    void
    setup_beamsource() override {
        spots = beamlets;
    }

    void
    initialize() override {
        scorer_mask = 0;
        dij_        = false;
        for (const std::string& n : scorer_names) {
            if (io::iequal(n, "Dose"))
                scorer_mask |= MQI_SCORE_DOSE;
            else if (io::iequal(n, "LETd"))
                scorer_mask |= MQI_SCORE_LETD;
            else if (io::iequal(n, "LETt"))
                scorer_mask |= MQI_SCORE_LETT;
            else if (io::iequal(n, "Dij"))
                dij_ = true;
            else
                throw error(MQI_E_ARG, "unrecognized scorer name " + n);
        }
        if (dij_) {
            if (scorer_names.size() > 1) throw error(MQI_E_ARG, "Dij cannot be scored with the other quantities");
            scorer_mask = MQI_SCORE_DOSE;
        }
        if (record_statistics) {
            if (!(stat_criteria >= 0.0f)) throw error(MQI_E_ARG, "Statistical criteria must be positive float");
            scorer_mask |= MQI_SCORE_STAT;
        }
        x_environment::initialize();
        std::vector<mqi_grid_desc> d;
        for (const beamline_grid& b : beamline) d.push_back(b.desc());
        check(ctx_, mqi_set_beamline(ctx_, (int) d.size(), d.data()), "mqi_set_beamline");
        set_roi(0, mask_files);
        set_roi(1, stat_mask_files);
    }

    void
    run() override {
        config.seed          = random_seed;
        config.physics_debug = 0;   // the tps build
        config.scoring_loops = scoring_loops >= 0 ? scoring_loops
                                                  : (record_statistics ? MQI_LOOPS_REFERENCE_STAT : MQI_LOOPS_REFERENCE_PLAIN);
        config.dij           = dij_ ? 1 : 0;
        if (dij_ && config.dij_slots_sqrt == 0 && config.dij_capacity == 0) config.dij_slots_sqrt = 4096;
        config.stat_criterion = record_statistics ? stat_criteria : 0.0f;
        config.stat_threshold = stat_threshold;
        x_environment::run();
    }

    void
    finalize() override {
        x_environment::finalize();
        if (output_dir.empty()) return;
        const double      scale  = output_scale(particles_per_history, rbe, n_fractions);
        const std::string prefix = beam_name + "_" + std::to_string(beamline.size()) + "_";
        if (dij_) {
            check(ctx_, mqi_dij_build(ctx_, 0.0), "mqi_dij_build");
            const std::string path = output_dir + "/" + prefix + scorer_names[0] + ".npz";
            check(ctx_, mqi_dij_save_npz(ctx_, 0, path.c_str(), scale), "mqi_dij_save_npz");
            return;
        }
        int fmt = MQI_FORMAT_RAW;
        if (io::iequal(output_format, "mhd"))
            fmt = MQI_FORMAT_MHD;
        else if (io::iequal(output_format, "mha"))
            fmt = MQI_FORMAT_MHA;
        else if (!io::iequal(output_format, "raw"))
            throw error(MQI_E_ARG, "output format " + output_format + " (raw, mhd, mha; Dij writes npz)");
        std::vector<std::string> names;
        if (scorer_mask & MQI_SCORE_DOSE) names.push_back(scorer_names_dose());
        if (scorer_mask & MQI_SCORE_LETD) names.insert(names.end(), {"LETd_numer", "LETd_denom"});
        if (scorer_mask & MQI_SCORE_LETT) names.insert(names.end(), {"LETt_numer", "LETt_denom"});
        if (scorer_mask & MQI_SCORE_STAT) names.insert(names.end(), {"Dose_stat", "DoseSquare_stat"});
        for (int k = 0; k < (int) names.size(); ++k) {
            const bool stat = (scorer_mask & MQI_SCORE_STAT) && k >= (int) names.size() - 2;
            if (stat && !save_statistics) continue;
            check(ctx_, mqi_save_scorer(ctx_, k, 0, output_dir.c_str(), (prefix + names[k]).c_str(), fmt, scale),
                  "mqi_save_scorer");
        }
    }

private:
    bool dij_ = false;

    std::string
    scorer_names_dose() const {
        for (const std::string& n : scorer_names)
            if (io::iequal(n, "Dose")) return n;
        return "Dose";
    }

    void
    set_roi(int target, const std::vector<std::string>& files) {
        std::vector<const char*> p;
        for (const std::string& f : files) p.push_back(f.c_str());
        check(ctx_, mqi_set_roi_mha(ctx_, target, (int) p.size(), p.data()), "mqi_set_roi_mha");
    }
};

}   // namespace mqi_hip

#endif
