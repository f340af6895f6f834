/* libmqi_hip.so: MI355X (gfx950) transport and scoring for moqui's hot path.
 *
 * The C ABI replaces the reference's device boundary: the kernel
 * mc::transport_particles_patient<R> / _stat<R>
 * (moqui/kernel_functions/mqi_transport.hpp:113-126, :252-265), the marshalling
 * functions mc::upload_node / upload_materials / upload_vertices /
 * upload_scorer_offset_vector / download_node
 * (moqui/kernel_functions/mqi_upload_data.hpp:148-450, mqi_download_data.hpp:37-121),
 * the statistics kernels (moqui/kernel_functions/mqi_variables.hpp:21-66) and
 * the run loops that drive them (moqui/base/environments/mqi_tps_env.hpp:1040-1603,
 * mqi_phantom_env.hpp:303-428).
 *
 * Conventions: every call returns 0 on success or a nonzero MQI_E_* code and
 * leaves a message for mqi_last_error(); nothing exits the process (the
 * reference calls exit(), mqi_error_check.hpp:11-42).  A context owns one
 * device and one stream; contexts are independent, so several may be used
 * from different threads.  Host pointers are plain C arrays; the only device
 * pointers in the ABI are the ones mqi_*_device_ptr hand out and the x / y
 * vectors of mqi_dij_matvec.
 */
#ifndef MQI_HIP_H
#define MQI_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "mqi_hip_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mqi_ctx mqi_ctx;

enum {
    MQI_OK           = 0,
    MQI_E_ARG        = 1,   /* invalid argument */
    MQI_E_STATE      = 2,   /* call order: grid / spots / scorers missing */
    MQI_E_HIP        = 3,   /* HIP runtime error */
    MQI_E_MEMORY     = 4,   /* device allocation failed */
    MQI_E_CAPACITY   = 5,   /* a fixed-size table (Dij hash) overflowed */
    MQI_E_IO         = 6    /* an output file could not be written */
};

/* scoring-loop semantics of a run */
enum {
    MQI_LOOPS_REFERENCE_PLAIN = 0, /* transport_particles_patient: scorers 0..n-3 scored twice when n >= 3
                                      (mqi_transport.hpp:204-225) */
    MQI_LOOPS_REFERENCE_STAT  = 1, /* transport_particles_patient_stat (mqi_transport.hpp:343-364) */
    MQI_LOOPS_ONCE            = 2  /* each scorer once per step (default) */
};

typedef struct mqi_run_config {
    uint64_t seed;             /* master seed of the per-history Philox streams */
    int32_t  physics_debug;    /* 1: the __PHYSICS_DEBUG__ physics of the phantom build, 0: release */
    int32_t  exact;            /* 1: verification kernels (portable math, no FMA), 0: production */
    int32_t  batched;          /* 1: one persistent launch over all items; 0: one plain launch per spot */
    int32_t  scoring_loops;    /* MQI_LOOPS_* */
    int32_t  stack_depth;      /* secondary stack per lane; 0: the reference's (200 debug, 10 release) */
    int32_t  block_size;       /* threads per workgroup; 0: 256 */
    int32_t  blocks;           /* workgroups of the persistent launch; 0: fill the device */
    int32_t  waves_per_simd;   /* register budget of the production kernels in waves/SIMD: 2 (compiler's), 3-6, 8; 0: 3 */
    int32_t  fetch_chunk;      /* histories a lane takes per work-queue fetch (batched); 0: 1 */
    int32_t  refill_lanes;     /* finished lanes a wave gathers before fetching new histories (batched); 0: 4 */
    int32_t  dij_slots_sqrt;   /* > 0: the Dij table of a (scenario, spot) with h histories gets
                                  ceil(dij_slots_sqrt * sqrt(h)) slots (at least 1024, at most dij_capacity
                                  or 2^24); 0: every table gets dij_capacity slots */
    int32_t  use_graph;        /* 1: capture the batched round in a hipGraph and replay it */
    float    stat_criterion;   /* > 0: repeat rounds until mean(sd/mean) <= criterion (a fraction) */
    float    stat_threshold;   /* voxels with mean > threshold * max mean enter the criterion */
    int32_t  max_rounds;       /* cap on statistical rounds (0: 64) */
    int32_t  dij;              /* 1: score the per-spot dose-influence tables */
    uint64_t dij_capacity;     /* hash slots per (scenario, spot) (the cap when dij_slots_sqrt > 0); 0: 131072 */
    uint64_t hist_offset;      /* added to every history index (independent repeat runs) */
} mqi_run_config;

typedef struct mqi_run_result {
    uint64_t histories;   /* histories transported (tracked_particles) */
    uint64_t dropped;     /* secondaries dropped by full stacks (the reference drops them silently) */
    uint64_t steps;       /* voxel steps taken by all tracks */
    uint32_t status;      /* bit 0: a Dij table overflowed; bit 1: a track hit the step cap (non-finite state) */
    int32_t  rounds;      /* statistical rounds run */
    double   kernel_ms;   /* device time of the transport launches (HIP events) */
    double   stat;        /* last mean(sd/mean) as a fraction; -1 when not computed */
} mqi_run_result;

/* Context.  Replaces the process-global mc::mc_world / mc_vertices /
 * mc_materials (mqi_variables.hpp:11-16) and the GPUID selection
 * (mqi_tps_env.hpp:161-164). */
int         mqi_create(int device, mqi_ctx** out);
int         mqi_destroy(mqi_ctx* ctx);
const char* mqi_last_error(const mqi_ctx* ctx);
int         mqi_device_info(mqi_ctx* ctx, int* compute_units, size_t* total_memory);

/* Voxel geometry of the patient/phantom child node: edges (n+1 per axis,
 * ascending, mm) and the density in g/mm^3 with x fastest
 * (grid3d::ijk2cnb, mqi_grid3d.hpp:403-413).  rot_fwd (row-major) and
 * translation are grid3d::rotation_matrix_fwd and translation_vector; NULL
 * means identity / zero.  Replaces add_node_geometry (mqi_upload_data.hpp:70-103). */
int mqi_set_grid(mqi_ctx* ctx, int nx, int ny, int nz, const float* xe, const float* ye, const float* ze,
                 const float* density, const float* rot_fwd, const float* translation);

/* Same from a CT: density = hu_to_density(HU) * density_scale on the device
 * (patient_material_t::hu_to_density, mqi_patient_materials.hpp:515-542, and
 * the DensityScaling loop of tps_env::setup_world, mqi_tps_env.hpp:766-769). */
int mqi_set_grid_hu(mqi_ctx* ctx, int nx, int ny, int nz, const float* xe, const float* ye, const float* ze,
                    const int16_t* hu, float density_scale, const float* rot_fwd, const float* translation);

/* Read back the density the device holds (nx*ny*nz floats). */
int mqi_get_density(mqi_ctx* ctx, float* out);

/* Spot table (the beamsource's beamlets with their history counts,
 * mqi_beamsource.hpp:49-112).  Vertices are sampled on the device per history
 * instead of on the host (phantom_env::setup_beamsource :205-232,
 * tps_env::run_by_spot :1530-1552). */
int mqi_set_spots(mqi_ctx* ctx, const mqi_spot* spots, int n_spots);

/* Robust scenarios in one launch: per scenario an edge shift (mm, x y z) and
 * a density scale (the XShift/YShift/ZShift and DensityScaling knobs,
 * mqi_tps_env.hpp:475-477, 766-769), and optionally a history budget factor
 * (each spot's histories times budget[s], rounded; NULL: 1).  n = 0 restores
 * the single nominal scenario. */
int mqi_set_scenarios(mqi_ctx* ctx, int n, const float* shift_xyz, const float* density_scale, const float* budget);

/* Scorers (MQI_SCORE_* bits of mqi_hip_types.h), in the reference's order
 * Dose, LETd numerator, LETd denominator, LETt numerator, LETt denominator,
 * energy deposit, dose to medium, Dose_stat, DoseSquare_stat
 * (mqi_tps_env.hpp:788-960; the statistics pair last, as the _stat kernel
 * expects); at most 6 arrays (5 with cfg.dij); per_scenario != 0 keeps one
 * array per scenario.
 * Clears them.  Replaces add_node_scorers (mqi_upload_data.hpp:106-142). */
int mqi_set_scorers(mqi_ctx* ctx, int mask, int per_scenario);
int mqi_zero_scorers(mqi_ctx* ctx);
int mqi_scorer_count(mqi_ctx* ctx);

/* Transport all histories of all spots (and scenarios).  Replaces
 * run_simulation / run_by_beam / run_by_beam_stat / run_by_spot
 * (mqi_tps_env.hpp:1040-1603) and phantom_env::run (mqi_phantom_env.hpp:303-428). */
int mqi_run(mqi_ctx* ctx, const mqi_run_config* cfg, mqi_run_result* res);

/* Autotuner (no counterpart in the reference, which fixes 512 threads per
 * block and one history per thread, mqi_tps_env.hpp:1058-1083): times
 * batched passes of the current plan with every spot's histories scaled by
 * `sample` (0 < sample <= 1): waves_per_simd {3, 4, 5} x block_size {128, 256},
 * then refill_lanes {1, 2, 4, 8, 16, 32, 60, 62, 64}, then fetch_chunk {1, 2, 4} (coordinate
 * descent), and writes the fastest into *cfg.  With
 * cfg->stack_depth < 0 it also picks the smallest secondary stack (16 .. 512)
 * that drops nothing on the sample (the reference's depth drops silently,
 * mqi_track_stack.hpp:36-41).  The trial passes score into the scorers, which
 * are cleared at the end.  *best_ms: device time of the fastest trial. */
int mqi_autotune(mqi_ctx* ctx, mqi_run_config* cfg, double sample, double* best_ms);

/* Transport caller-made vertices (7 floats each: ke, x, y, z, u, v, w), the
 * upload_vertices + transport_particles_patient path.  spot_ids may be NULL
 * (the scorer_offset_vector, mqi_transport.hpp:150-154). */
int mqi_run_vertices(mqi_ctx* ctx, const float* vertices, uint64_t n, const uint32_t* spot_ids,
                     const mqi_run_config* cfg, mqi_run_result* res);

/* Per-step record of the transport for verification: for history h, up to
 * max_steps records of (voxel, its.dist, vtx0.pos, vtx0.dir, vtx1.pos, vtx0.ke, dE)
 * at the point where the reference scores the step.  vertices (7 floats per
 * history, may be NULL) replace the spot table as in mqi_run_vertices.
 * cnb: n*max_steps, rec: n*max_steps*12, n_steps: n. */
int mqi_trace(mqi_ctx* ctx, const mqi_run_config* cfg, const float* vertices, uint64_t n_histories, int max_steps,
              int64_t* cnb, float* rec, int32_t* n_steps);

/* Scorer k (0-based in the order above) of a scenario: copy to host
 * (download_node, mqi_download_data.hpp:37-121) or hand out the device array. */
int mqi_get_scorer(mqi_ctx* ctx, int k, int scenario, double* out);
int mqi_scorer_device_ptr(mqi_ctx* ctx, int k, int scenario, double** dptr);

/* mean(sd/mean) over voxels with mean > threshold * max mean from the Dose_stat
 * and DoseSquare_stat scorers after n_histories histories
 * (calculate_standard_deviation + calculate_stat, mqi_variables.hpp:21-48,
 * mqi_tps_env.hpp:1340-1426). */
int mqi_stat(mqi_ctx* ctx, int scenario, uint64_t n_histories, float threshold, double* stat);

/* Dose-influence matrices.  mqi_dij_build turns the per-spot tables of the
 * last run with cfg.dij = 1 into one CSR matrix per scenario, rows = spots,
 * columns = voxels (the layout io::save_to_npz writes, mqi_io.hpp:249-320),
 * float32 values and int32 indices, keeping entries >= rel_threshold times
 * the row maximum; columns sorted; a voxel-major copy is kept for D w. */
int mqi_dij_build(mqi_ctx* ctx, double rel_threshold);
/* A caller's CSR as scenario `scenario`'s Dij (host arrays: rows + 1 row
 * pointers from 0, nondecreasing; each row's columns in [0, cols), strictly
 * ascending; float32 values), prepared for mqi_dij_matvec like a built one
 * and replacing any Dij of that scenario.  No counterpart in the reference,
 * which only writes its Dij (io::save_to_npz, mqi_io.hpp:249-320): this
 * takes such a matrix back (read by the caller, values as float32) so the
 * products and their gradients run on it.  MQI_E_ARG names the first
 * offending row or entry. */
int mqi_dij_set_csr(mqi_ctx* ctx, int scenario, int64_t rows, int64_t cols, const int64_t* indptr,
                    const int32_t* indices, const float* values);
int mqi_dij_info(mqi_ctx* ctx, int scenario, int64_t* nnz, int64_t* rows, int64_t* cols);
int mqi_dij_copy(mqi_ctx* ctx, int scenario, int64_t* indptr, int32_t* indices, float* values);
int mqi_dij_device_ptrs(mqi_ctx* ctx, int scenario, const int64_t** indptr, const int32_t** indices,
                        const float** values);

/* y = alpha * A x (+ y if accumulate) on the device, A = D (voxels x spots,
 * transpose = 0, the forward dose D w) or D^T (transpose = 1, the adjoint
 * D^T g).  x: float32 device vector; y: float64 if y_double else float32;
 * FP64 accumulation in a fixed order (results do not depend on scheduling).
 * stream: a hipStream_t or NULL for the context's stream.  The call only
 * enqueues work: y is ready when that stream reaches it.  Products of the same
 * scenario share a device staging buffer (the compacted x of D^T g): do not
 * run two of them concurrently on different streams. */
int mqi_dij_matvec(mqi_ctx* ctx, int scenario, int transpose, const float* x, void* y, int y_double, double alpha,
                   int accumulate, void* stream);

/* The quadratic objective of the dose d = alpha * D w of scenario `scenario`
 * in one pass over the D w matrix, without d: *f = mean_v c_v (d_v - p_v)^2
 * (device double; c NULL: all 1) and, when g is set, g_v = 2 c_v (d_v - p_v) /
 * n_voxels (device float32), the input of the D^T g that gives the gradient
 * alpha D^T g.  w: device float32 (spots); p: device double (voxels).
 * Enqueued on stream like mqi_dij_matvec, in a fixed order.  No counterpart
 * in the reference (its Dij is only written out): the objective of the
 * differentiable layer over the Dij. */
int mqi_dij_quadratic(mqi_ctx* ctx, int scenario, const float* w, const double* p, const float* c, double alpha,
                      float* g, double* f, void* stream);

/* Region of interest of the scorers (target 0: Dose, LETd, LETt, Edep, Dm
 * and Dij) or of the statistics pair (target 1), from a voxel mask (x
 * fastest, the grid's size; the reference's mask_total, the sum of the
 * structure masks, mqi_file_handler.hpp:94-113): the CONTOUR mapping that
 * mask_reader::mask_to_roi builds (mqi_file_handler.hpp:177-217, roi_t
 * mqi_roi.hpp:18-196).  In voxel order a 1 opens a run and a 0 closes it;
 * other values continue the current run, so a 0/1 mask scores exactly its
 * 1 voxels.  A run open at the last voxel ends there.  Scoring goes through
 * roi_->idx(v) > 0 for the scorers and get_mask_idx(v) >= 0 for the
 * statistics pair (mqi_transport.hpp:342-364), i.e. a voxel scores when it
 * is in a run.  mask = NULL restores DIRECT (every voxel, voxel 0 excluded
 * from the scorers).  Arrays stay dense over the grid: the reference stores
 * the statistics pair compactly over its ROI, here voxels outside hold 0
 * and the stopping statistic covers the same voxels.  The runs are expanded
 * into one byte per voxel on the device (nvox bytes while a mask is set).
 * Set after mqi_set_grid: setting a grid clears both masks. */
int mqi_set_roi(mqi_ctx* ctx, int target, const uint8_t* mask);

/* The same from structure mask files: tps_env's ScoringMask "Mask" list
 * (target 0) or StatROIMaskFilename list (target 1), mqi_tps_env.hpp:
 * 224-232,290-291,776-805.  Each file is a MetaImage with one byte per voxel
 * after the header (ElementDataFile = LOCAL), read as
 * mask_reader::read_mha_file does (mqi_file_handler.hpp:39-89); the masks
 * are summed into mask_total (read_mask_files, :92-113), which then goes
 * through mqi_set_roi.  A file whose DimSize differs from the grid's, whose
 * ElementType is not MET_UCHAR / MET_CHAR or which holds a value other than
 * 0 or 1 is refused with MQI_E_IO.  n = 0 restores DIRECT. */
int mqi_set_roi_mha(mqi_ctx* ctx, int target, int n, const char* const* paths);

/* Beamline children transported before the patient grid, in order (the
 * world of tps_env::setup_world, mqi_tps_env.hpp:713-753: range shifter and
 * voxelised aperture nodes, then the phantom; transport loop
 * mqi_transport.hpp:158-240).  Every track (primary or popped secondary)
 * goes through the children from the first: into the child's frame, to its
 * first voxel, steps with the same physics until it leaves or stops, back to
 * the world frame; then the next child, the patient last.  Beamline children
 * have no scorers; an aperture blocks with density above 99.9 g/mm^3
 * (stepping stops the track), passes with density below 1e-7.  Robust
 * scenarios shift and scale the patient only.  n = 0 removes them; at most
 * 8 children. */
int mqi_set_beamline(mqi_ctx* ctx, int n, const mqi_grid_desc* children);

/* Loads a CT from a MetaImage file with the voxels after the header
 * (ElementDataFile = LOCAL, float32 HU, tps_env::read_ct_image,
 * mqi_tps_env.hpp:615-701) and sets it as the grid: edges from Offset (the
 * first voxel's centre) and ElementSpacing (mqi_tps_env.hpp:469-524), HU
 * truncated to int16, then as mqi_set_grid_hu with the density scale.  shift:
 * NULL or the volume shift (mm) added to the first edge of each axis. */
int mqi_set_grid_ct_mha(mqi_ctx* ctx, const char* path, float density_scale, const float* shift);

/* Writes the Dij of a scenario as the reference's scipy-style CSR .npz
 * (io::save_to_npz, mqi_io.hpp:249-320, zip writer mqi_sparse_io.hpp:166-400):
 * an uncompressed zip of indices.npy and indptr.npy (uint32), shape.npy
 * (uint32 [spots, voxels]), data.npy (float64, value * scale) and format.npy
 * ("csr").  scale: the reference's particles_per_history * rbe * n_fractions.
 * Rows hold mqi_dij_build's entries (columns sorted, threshold applied). */
int mqi_dij_save_npz(mqi_ctx* ctx, int scenario, const char* path, double scale);

/* output formats of mqi_save_scorer */
enum {
    MQI_FORMAT_RAW = 0,   /* <dir>/<name>.raw, float64 (io::save_to_bin, mqi_io.hpp:165-183) */
    MQI_FORMAT_MHD = 1,   /* <dir>/<name>.mhd + .raw (io::save_to_mhd, mqi_io.hpp:493-537) */
    MQI_FORMAT_MHA = 2    /* <dir>/<name>.mha, data after the header (io::save_to_mha, mqi_io.hpp:539-591) */
};

/* Writes scorer k of a scenario, times scale, in one of the reference's dense
 * formats (the save_reshaped_files loop, mqi_tps_env.hpp:1840-1895); the
 * MetaImage header carries the grid's spacing and first voxel centre. */
int mqi_save_scorer(mqi_ctx* ctx, int k, int scenario, const char* dir, const char* name, int format,
                    double scale);

/* Device time (ms, HIP events) of the last mqi_dij_matvec, for measurement;
 * waits for that product to finish. */
int mqi_last_matvec_ms(mqi_ctx* ctx, double* ms);

/* Verification: the portable math functions on the device (fn 0 log, 1 exp,
 * 2 pow, 3 sin, 4 cos, 5 acos, 6 sqrt, 7 float divide, 8 double divide,
 * 9 float sqrt) on n host inputs. */
int mqi_math_probe(mqi_ctx* ctx, int fn, const double* x, const double* y, int64_t n, double* out);

/* Verification: the material and physics functions as the transport kernels
 * evaluate them on the device, in the verification (exact = 1) or production
 * (exact = 0) build, with phantom-build (physics_debug = 1) or release
 * physics, on n host inputs:
 *   fn 0  spr_default(rho[i], ek[i]) (mqi_patient_materials.hpp:415-450), n floats out;
 *   fn 1  radiation_length_default(rho[i]) (:452-473; ek may be NULL), n floats out;
 *   fn 2  per i, 10 floats: the cross sections of ionization, p-p elastic,
 *         p-O elastic and p-O inelastic at (ek[i], rho[i])
 *         (mqi_p_ionization.hpp:254-268, mqi_pp_elastic.hpp:221-235,
 *         mqi_po_elastic.hpp, mqi_po_inelastic.hpp), dEdx
 *         (mqi_p_ionization.hpp:271-286), and beta^2, gamma, Et, Te_max,
 *         momentum (mqi_relativistic_quantities.hpp:24-54).
 * rho in g/mm^3, ek in MeV. */
int mqi_physics_probe(mqi_ctx* ctx, int exact, int physics_debug, int fn, const float* rho, const float* ek, int64_t n,
                      float* out);

/* Measurement probe (no counterpart in the reference): the time of n_updates
 * FP64 atomic adds to uniformly hashed slots of an n_slots array
 * (1 <= n_slots <= 2^32), the transport's scoring stream without the rest of
 * the step; *ms is the device time of one pass after a warm-up pass.  The
 * throughput ceiling the transport's scoring atomics are measured against.
 * window > 0 (a power of two below n_slots): each wave's 64 adds fall in one
 * window of that many slots placed at random, to measure how clustering
 * raises the rate; 0: uniform. */
int mqi_atomic_probe(mqi_ctx* ctx, int64_t n_slots, int64_t n_updates, int64_t window, double* ms);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): mqi_run_vertices with the vertices (7 floats each) and
 * the optional spot ids already in DEVICE memory, and nothing allocated.  The
 * one-item work list, the counters, the work-range queue and the per-lane
 * secondary stacks live in a static per-GPU scratch array of
 * MQI_NOALLOC_STACK_BYTES (+ 8 KiB), held from library load to process end;
 * the launch takes as many lanes as that holds stacks for (stack_depth x 40 B
 * per lane), at most the launch mqi_run_vertices makes.  A stack_depth x
 * block_size above the scratch is refused with MQI_E_CAPACITY.  Scores into
 * the scorers set with mqi_set_scorers, as mqi_run_vertices does; results
 * equal mqi_run_vertices' to FP64 summation order.  Every call on one GPU
 * shares that GPU's array (counters are cleared at the start of each call):
 * calls that overlap on one GPU race. */
#define MQI_NOALLOC_STACK_BYTES (32u << 20)
int mqi_run_vertices_noalloc(mqi_ctx* ctx, const float* d_vertices, uint64_t n, const uint32_t* d_spot_ids,
                             const mqi_run_config* cfg, mqi_run_result* res);

#ifdef __cplusplus
}
#endif

#endif
