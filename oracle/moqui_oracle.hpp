// Test infrastructure: the CPU restatement ("oracle") of moqui's transport and
// scoring path.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load it (through ctypes, as oracle/_ref/liboracle.so).
//
// Policies selected per call:
//   prec     32 -> float arithmetic with the reference's float/double
//            promotions reproduced expression by expression; 64 -> double.
//   rng      0 -> the reference CPU generator: std::default_random_engine
//            with fresh libstdc++ uniform/normal/exponential distributions per
//            draw (mqi_math.hpp:454-506), re-seeded per history with its
//            transport seed; 1 -> the Philox per-history stream of
//            moquimc_amd/csrc/mqi_philox.hpp (what the GPU kernel draws).
//   math     0 -> libm (as the reference's CPU build); 1 -> the portable
//            functions of moquimc_amd/csrc/mqi_pmath.hpp (bit-identical to
//            the GPU verification build).
//   debug    1 -> __PHYSICS_DEBUG__ physics (phantom build), 0 -> release.
#ifndef MOQUI_ORACLE_HPP
#define MOQUI_ORACLE_HPP

#include <cstdint>

#include "include/mqi_hip_types.h"

extern "C" {

void oracle_set_beamline(int n, const int* dims, const float* edges, const float* dens, const float* xf);

void oracle_set_roi(int target, int64_t nvox, const uint8_t* mask);

int oracle_transport(int prec, int rng, int math, int debug, uint64_t master_seed, int nx, int ny,
                     int nz, const float* xe, const float* ye, const float* ze,
                     const float* density, const float* rot9, const float* t3, int64_t n_hist,
                     const float* vtx, const int32_t* seeds, const uint32_t* spot_ids, int mask,
                     int kernel, int n_threads, double** out, uint64_t* tracked, int64_t* log_cnb,
                     float* log_rec, int64_t log_cap, int64_t* log_n);

int oracle_run_spots(int prec, int rng, int math, int debug, uint64_t master_seed,
                     uint32_t scenario, int nx, int ny, int nz, const float* xe, const float* ye,
                     const float* ze, const float* density, const mqi_spot* spots, int n_spots,
                     int mask, int kernel, int n_threads, double** out, uint64_t* tracked);

void oracle_sample_spot(int rng, int math, uint64_t master_seed, uint32_t scenario, uint32_t spot_index,
                        const mqi_spot* spot, int64_t n, float* vtx_out, int32_t* seed_out);

void oracle_coordinate_rotation(int math, const float* angles4, float* out9);
void oracle_hu_to_density(const int16_t* hu, int64_t n, float* out);
void oracle_spr(int math, int debug, const float* rho, const float* ek, int64_t n, float* out);
void oracle_radiation_length(int math, int debug, const float* rho, int64_t n, float* out);
void oracle_physics_probe(int math, const float* ek, int64_t n, float rho, float* out);
void oracle_rotate_direction(int math, const float* dir, const float* theta, const float* phi,
                             int64_t n, float* out);
void oracle_grid_probe(int op, int nx, int ny, int nz, const float* xe, const float* ye,
                       const float* ze, const float* p, const float* d, const int32_t* cell_in,
                       int64_t n, float* dist_out, float* d_out, int32_t* cell_out);
void oracle_math_probe(int fn, const double* x, const double* y, int64_t n, double* out);
void oracle_philox(const uint32_t* ctr, const uint32_t* key, int64_t n, uint32_t* out);
}

#endif
