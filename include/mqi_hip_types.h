/* Plain data types shared by the C ABI of libmqi_hip.so (include/mqi_hip.h)
 * and its callers.  No torch or HIP types appear here. */
#ifndef MQI_HIP_TYPES_H
#define MQI_HIP_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A voxel grid placed in the world: a beamline child (range shifter,
 * aperture) of mqi_set_beamline.  Edges ascend, n + 1 per axis, in the
 * child's frame; density in g/mm^3, x fastest; rotation (row-major) and
 * translation map the child's frame to the world (grid3d's
 * rotation_matrix_fwd and translation_vector, mqi_grid3d.hpp:31-890). */
typedef struct mqi_grid_desc {
    int32_t      nx, ny, nz;
    int32_t      reserved;
    const float* xe;
    const float* ye;
    const float* ze;
    const float* density;
    float        rotation[9];
    float        translation[3];
} mqi_grid_desc;

/* One pencil-beam spot (a "beamlet" in the reference).  The phase-space and
 * energy model is the reference's phsp_6d + norm_1d/const_1d beamlet
 * (moqui/base/distributions/mqi_phsp6d.hpp:62-79, mqi_norm_1d.hpp:54-56,
 * mqi_const_1d.hpp:55-57, mqi_beamlet.hpp:81-90): in the beam frame the
 * position is mean + sigma * N(0,1) per axis, the direction x', y' are
 * correlated with x, y through corr, and z' = -sqrt(1 - x'^2 - y'^2), so the
 * beam travels along -z of its frame; the vertex is then mapped to the world
 * by rotation * v + translation (coordinate_transform, mqi_coordinate_transform.hpp:27-122). */
typedef struct mqi_spot {
    float    energy;        /* mean kinetic energy, MeV */
    float    sigma_energy;  /* Gaussian energy spread, MeV (ignored when energy_const) */
    int32_t  energy_const;  /* 1: const_1d energy (no random draw), 0: norm_1d */
    uint32_t phsp;          /* MQI_PHSP_GAUSS (phsp_6d) or MQI_PHSP_UNIFORM (phsp_6d_uniform) */
    float    mean[6];       /* x, y, z (mm), x', y', z' in the beam frame */
    float    sigma[6];      /* sigmas of the same; sigma[5] is ignored */
    float    corr[2];       /* x-x' and y-y' correlation coefficients */
    float    rotation[9];   /* beam frame -> world, row-major */
    float    translation[3];/* mm */
    uint64_t histories;     /* histories of this spot */
} mqi_spot;

/* Phase-space distribution of a spot: GAUSS draws N(0,1) per variable
 * (phsp_6d, mqi_phsp6d.hpp:62-79); UNIFORM draws U(-1,1) (phsp_6d_uniform,
 * mqi_phsp6d_uniform.hpp:68-82, the distribution phantom_env uses,
 * mqi_phantom_env.hpp:216); both combine the draws as described above. */
enum {
    MQI_PHSP_GAUSS   = 0,
    MQI_PHSP_UNIFORM = 1
};

/* Scorer bits.  DOSE: dose-to-water (Gy per history sum,
 * mqi_scorer_energy_deposit.hpp:43-60); LETD: LETd numerator and
 * denominator (:93-139); STAT: per-step dose and dose^2 for the
 * statistical stopping criterion (mqi_tps_env.hpp:939-959). */
enum {
    MQI_SCORE_DOSE = 1,
    MQI_SCORE_LETD = 2,
    MQI_SCORE_STAT = 4,
    MQI_SCORE_LETT = 16,   /* LETt numerator and denominator (LETt_weight1/2, :141-177) */
    MQI_SCORE_EDEP = 32,   /* energy deposit dE + local dE, MeV (energy_deposit, :14-20) */
    MQI_SCORE_DM   = 64    /* dose to medium of primaries, Gy (dose_to_medium, :80-92) */
};

#ifdef __cplusplus
}
#endif

#endif
