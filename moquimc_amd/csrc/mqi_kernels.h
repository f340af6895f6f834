// Launch parameters shared by the host runtime (mqi_api.cpp) and the
// transport kernels (mqi_transport.hip).  Device pointers only.
#ifndef MQI_KERNELS_H
#define MQI_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mqi_hip.h"

#define MQI_MAX_SCORERS 6

// scorer kinds, in the order tps_env lays scorers out (mqi_tps_env.hpp:788-960)
enum {
    MQI_KIND_DOSE        = 0,   // dose_to_water
    MQI_KIND_LETD_NUM    = 1,   // LETd_weight1
    MQI_KIND_LETD_DEN    = 2,   // LETd_weight2
    MQI_KIND_STAT_DOSE   = 3,   // Dose_stat (dose_to_water)
    MQI_KIND_STAT_DOSE2  = 4,   // DoseSquare_stat (dose_to_water_square)
    MQI_KIND_DIJ         = 5,   // dose_to_water keyed by (voxel, spot)
    MQI_KIND_LETT_NUM    = 6,   // LETt_weight1
    MQI_KIND_LETT_DEN    = 7,   // LETt_weight2
    MQI_KIND_EDEP        = 8,   // energy_deposit
    MQI_KIND_DM          = 9    // dose_to_medium (primaries)
};

// scoring loop semantics (MQI_LOOPS_* of mqi_hip.h)
enum {
    MQI_LOOPS_PLAIN = MQI_LOOPS_REFERENCE_PLAIN,   // scorers 0..n-3 twice (mqi_transport.hpp:204-225)
    MQI_LOOPS_STAT  = MQI_LOOPS_REFERENCE_STAT     // (mqi_transport.hpp:343-364)
};

// Scorer slots whose deposits a lane sums per voxel before one atomic add
// (the rest are added per step), and where those sums sit in the dynamic LDS
// of the transport kernel: after the six 600-entry tables, the three edge
// arrays and the beamline edges, at an 8-byte boundary, MQI_DEFER doubles
// per thread.
#define MQI_DEFER 4
#define MQI_LDS_TABLE_FLOATS 3600
__host__ __device__ inline uint32_t mqi_pend_offset(int nx, int ny, int nz, int bl_edges) {
    const uint32_t f = MQI_LDS_TABLE_FLOATS + (uint32_t) (nx + 1 + ny + 1 + nz + 1 + bl_edges);
    return (f + 1u) & ~1u;
}
__host__ __device__ inline size_t mqi_shmem_bytes(int nx, int ny, int nz, int bl_edges, int threads) {
    return sizeof(float) * (size_t) mqi_pend_offset(nx, ny, nz, bl_edges) + sizeof(double) * MQI_DEFER * (size_t) threads;
}

// floats per secondary-stack entry: ke0, pos[3], dir[3], ke1, dE, and the
// beamline child whose transform still has to place it (-1: placed)
#define MQI_STACK_FLOATS 10

#define MQI_STATUS_DIJ_FULL 1u
#define MQI_STATUS_TRACK_CAP 2u
#define MQI_MAX_TRACK_STEPS 200000

// Work ranges of the dynamic form: one per XCD of the MI355X, counters a
// 128-byte line apart.
#define MQI_QUEUES 8
#define MQI_QUEUE_STRIDE 16

// first history of range q of a list of `total` histories (balanced split)
__host__ __device__ inline uint64_t mqi_range_begin(uint64_t total, int q) {
    const uint64_t base = total / MQI_QUEUES, rem = total % MQI_QUEUES;
    return base * (uint64_t) q + ((uint64_t) q < rem ? (uint64_t) q : rem);
}

struct transport_params {
    // geometry (one voxel child of the world)
    int          nx, ny, nz;
    const float* xe;
    const float* ye;
    const float* ze;
    const float* density;   // g/mm^3, x fastest
    float        fwd[9], inv[9], trans[3];
    // robust scenarios: per-scenario edge shift (mm, 3 per scenario) and density scale; may be null
    const float* scen_shift;
    const float* scen_scale;
    // work list: items are (scenario, spot) pairs with item_end = inclusive prefix sum of histories
    const mqi_spot* spots;
    const float*    vertices;      // optional: 7 floats per history (ke, pos, dir), replaces sampling
    const uint32_t* vertex_spot;   // optional: spot id per vertex (the scorer_offset_vector)
    int             n_spots;
    const uint64_t* item_end;
    const uint32_t* item_spot;
    const uint32_t* item_scenario;
    int             n_items;
    uint64_t        total;
    // dynamic form: MQI_QUEUES contiguous ranges of the flattened work list,
    // one per XCD, each with its own counter at queue[q * MQI_QUEUE_STRIDE]
    // (initialised to the range's first history)
    unsigned long long* queue;
    int             fetch_chunk;   // histories a lane takes per fetch (dynamic form), >= 1
    int             refill_lanes;  // idle lanes a wave gathers before refilling them (dynamic form), >= 1
    const uint64_t* hist_offset;   // device scalar added to every history index (graph-replay safe)
    uint64_t        seed;
    // scorers
    int      n_scorers;
    int      kinds[MQI_MAX_SCORERS];
    double*  out[MQI_MAX_SCORERS];
    uint64_t out_scenario_stride;   // 0: scenarios share the arrays; else per-scenario offset (elements)
    int      loops;
    // Grouped scoring (production kernels, each-scorer-once, no Dij): the
    // leading group_ns scorers (2 .. 4) add into one interleaved array with
    // pair (2 or 4; 0: off) doubles per element, element i of scorer s at
    // pair_out[pair * i + s], folded into their own arrays after the pass
    double*  pair_out;
    int      pair;
    int      group_ns;
    uint32_t* dij_keys;
    double*   dij_vals;
    const uint64_t* dij_off;  // item i owns slots [dij_off[i], dij_off[i + 1]) of dij_keys/dij_vals
    // secondary stack
    float* stack;
    int    stack_cap;
    // counters
    unsigned long long* tracked;
    unsigned long long* drops;
    unsigned long long* steps;   // voxel steps taken (all tracks)
    unsigned int*       status;
    // beamline children transported before the patient, in order (range
    // shifter, aperture, ...): per child bl_meta = {nx, ny, nz, offset of its
    // edges in bl_edges (x, y, z concatenated), offset of its densities in
    // bl_density (two ints, low and high 32 bits), 0} and bl_xf = {fwd[9],
    // inv[9], translation[3]}; no scorers
    int          n_bl;
    const int*   bl_meta;
    const float* bl_xf;
    const float* bl_edges;
    int          bl_n_edges;
    const float* bl_density;
    // ROIs: NULL is DIRECT for both (the regular scorers score voxels v > 0,
    // the statistics pair every voxel); otherwise one byte per voxel, bit 0
    // set where the regular scorers score, bit 1 where the statistics pair
    // does (the host expands CONTOUR runs, or DIRECT for a target without one)
    const uint8_t*  roi_flags;
    // per-step trace (verification)
    int64_t* trace_cnb;
    float*   trace_rec;
    int      trace_cap;
    int*     trace_n;
};

namespace mqi_fast
{
int launch_transport(const transport_params& P, int debug, int dynamic, int trace, int waves, int blocks,
                     int threads, size_t shmem, hipStream_t stream);
int occupancy(int debug, int dynamic, int trace, int waves, int threads, size_t shmem);
int physics_probe(int debug, int fn, const float* rho, const float* ek, int64_t n, float* out, hipStream_t stream);
}
namespace mqi_exact
{
int launch_transport(const transport_params& P, int debug, int dynamic, int trace, int waves, int blocks,
                     int threads, size_t shmem, hipStream_t stream);
int occupancy(int debug, int dynamic, int trace, int waves, int threads, size_t shmem);
int physics_probe(int debug, int fn, const float* rho, const float* ek, int64_t n, float* out, hipStream_t stream);
}

#endif
