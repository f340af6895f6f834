// Auxiliary kernels: buffer fills, HU -> density calibration, the
// statistical-uncertainty reduction, Dij compaction into CSR, and the Dij
// forward (D w) / adjoint (D^T g) sparse products.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "mqi_aux.h"
#include "mqi_device.hpp"
#include "mqi_kernels.h"

#define MQI_TABLE_QUALIFIER static __constant__
#include "mqi_tables.inc"

namespace mqi_aux
{

// ------------------------------------------------------------------ fills
// Grid-stride fills; used instead of hipMemsetAsync so a captured graph
// re-clears its buffers on every replay.
__global__ void k_fill_u64(unsigned long long* p, unsigned long long v, uint64_t n) {
    for (uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t) gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
    for (uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t) gridDim.x * blockDim.x)
        p[i] = v;
}

static int
grid_for(uint64_t n) {
    uint64_t b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return (int) b;
}

__global__ void k_queue_init(unsigned long long* q, uint64_t total) {
    const int i = threadIdx.x;
    if (i < MQI_QUEUES) q[i * MQI_QUEUE_STRIDE] = mqi_range_begin(total, i);
}

int
queue_init(unsigned long long* q, uint64_t total, hipStream_t s) {
    hipLaunchKernelGGL(k_queue_init, dim3(1), dim3(64), 0, s, q, total);
    return (int) hipGetLastError();
}

int
fill_u64(void* p, unsigned long long v, uint64_t n, hipStream_t s) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n)), dim3(256), 0, s, (unsigned long long*) p, v, n);
    return (int) hipGetLastError();
}
int
fill_u32(void* p, uint32_t v, uint64_t n, hipStream_t s) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n)), dim3(256), 0, s, (uint32_t*) p, v, n);
    return (int) hipGetLastError();
}

// ------------------------------------------------------------------ calibration
// rho = hu_to_density(HU) * scale (mqi_tps_env.hpp:766-769)
__global__ void k_hu_to_density(const int16_t* hu, float* rho, uint64_t n, float scale) {
    __shared__ float corr[3996];
    for (int i = threadIdx.x; i < 3996; i += blockDim.x) corr[i] = kDensityCorrection[i];
    __syncthreads();
    for (uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t) gridDim.x * blockDim.x)
        rho[i] = mqi_dev::hu_to_density(hu[i], corr) * scale;
}

int
hu_to_density(const int16_t* hu, float* rho, uint64_t n, float scale, hipStream_t s) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_hu_to_density, dim3(grid_for(n)), dim3(256), 0, s, hu, rho, n, scale);
    return (int) hipGetLastError();
}

// ------------------------------------------------------------------ math probe
// The portable functions evaluated on the device (verification of the exact build).
__global__ void k_math_probe(int fn, const double* x, const double* y, double* out, int64_t n) {
    int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double r;
    switch (fn) {
    case 0: r = mqi_pm::pm_log(x[i]); break;
    case 1: r = mqi_pm::pm_exp(x[i]); break;
    case 2: r = mqi_pm::pm_pow(x[i], y[i]); break;
    case 3: r = mqi_pm::pm_sin(x[i]); break;
    case 4: r = mqi_pm::pm_cos(x[i]); break;
    case 5: r = mqi_pm::pm_acos(x[i]); break;
    case 6: r = mqi_pm::pm_sqrt(x[i]); break;
    case 7: r = (double) ((float) x[i] / (float) y[i]); break;   // IEEE float division
    case 8: r = x[i] / y[i]; break;                              // IEEE double division
    default: r = (double) __builtin_sqrtf((float) x[i]); break;
    }
    out[i] = r;
}

int
math_probe(int fn, const double* x, const double* y, double* out, int64_t n, hipStream_t s) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_math_probe, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, s, fn, x, y, out, n);
    return (int) hipGetLastError();
}

// ------------------------------------------------------------------ atomic probe
// n_updates FP64 atomic adds of 1.0 to slots picked by a 64-bit hash of the
// update index over n_slots: the scoring pattern of the transport (random
// voxels, device-scope adds whose results are not read) without the rest of
// the step, to measure the part's throughput for such atomics.
MQI_DI uint32_t fmix32(uint32_t h) {   // the murmur3 32-bit finaliser
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}

// slot = a hash of i scaled to [0, n_slots) by a high multiply (n_slots <=
// 2^32): a few integer operations per update, so the probe measures the
// atomic stream, not a 64-bit modulo.  window > 0: the wave's 64 updates
// (i with the same i / 64) share one window [base, base + window).
__global__ void k_atomic_probe(double* a, uint64_t n_slots, uint64_t n_updates, uint64_t window) {
    for (uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n_updates;
         i += (uint64_t) gridDim.x * blockDim.x) {
        uint64_t slot;
        if (window == 0) {
            slot = ((uint64_t) fmix32((uint32_t) i) * n_slots) >> 32;
        } else {
            const uint64_t base = ((uint64_t) fmix32((uint32_t) (i >> 6) ^ 0x9E3779B9u) * (n_slots - window)) >> 32;
            slot                = base + (fmix32((uint32_t) i) & (uint32_t) (window - 1));
        }
        atomicAdd(a + slot, 1.0);
    }
}

int
atomic_probe(double* a, uint64_t n_slots, uint64_t n_updates, uint64_t window, hipStream_t s) {
    if (n_slots == 0 || n_updates == 0) return 0;
    if (n_slots > (1ull << 32)) return (int) hipErrorInvalidValue;
    if (window != 0 && (window >= n_slots || (window & (window - 1)) != 0)) return (int) hipErrorInvalidValue;
    hipLaunchKernelGGL(k_atomic_probe, dim3(65536), dim3(256), 0, s, a, n_slots, n_updates, window);
    return (int) hipGetLastError();
}

// ------------------------------------------------------------------ statistics
// calculate_standard_deviation (mqi_variables.hpp:21-48) fused with the
// reduction of calculate_stat (mqi_tps_env.hpp:1396-1425):
//   mean_i = S1_i / N, sd_i = sqrtf((S2_i/N - mean_i^2) / (N - 1)),
//   stat = mean over {i : mean_i > thr * max mean} of sd_i / mean_i.
// Pass 1: max of the float means.  Pass 2: sum and count above threshold.
__global__ void k_stat_max(const double* s1, uint64_t n, double N, float* out_max) {
    float m = 0.0f;
    for (uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t) gridDim.x * blockDim.x) {
        float mu = (float) (s1[i] / N);
        m        = mu > m ? mu : m;
    }
    typedef hipcub::BlockReduce<float, 256> BR;
    __shared__ typename BR::TempStorage tmp;
    float bm = BR(tmp).Reduce(m, hipcub::Max());
    if (threadIdx.x == 0) atomicMax((int*) out_max, __float_as_int(bm));   // non-negative floats order as ints
}

__global__ void k_stat_sum(const double* s1, const double* s2, uint64_t n, double N, const float* max_mu, float thr,
                           double* out_sum, unsigned long long* out_cnt) {
    // dose_mean[i] > dose_max * stat_threshold with dose_max a double
    // (mqi_tps_env.hpp:1409-1425)
    const double lim = (double) *max_mu * (double) thr;
    double      acc = 0.0;
    unsigned long long cnt = 0;
    for (uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t) gridDim.x * blockDim.x) {
        if (s1[i] == 0.0 && s2[i] == 0.0) continue;   // untouched slot (key1 == empty_pair)
        float mu = (float) (s1[i] / N);
        float sd = sqrtf((float) (((s2[i] / N) - ((s1[i] / N) * (s1[i] / N))) / (N - 1)));
        if ((double) mu > (double) lim) {
            acc += (double) (sd / mu);
            cnt += 1;
        }
    }
    typedef hipcub::BlockReduce<double, 256>             BRd;
    typedef hipcub::BlockReduce<unsigned long long, 256> BRu;
    __shared__ typename BRd::TempStorage t1;
    __shared__ typename BRu::TempStorage t2;
    double                               bs = BRd(t1).Sum(acc);
    __syncthreads();
    unsigned long long bc = BRu(t2).Sum(cnt);
    if (threadIdx.x == 0) {
        atomicAdd(out_sum, bs);
        atomicAdd(out_cnt, bc);
    }
}

int
stat_reduce(const double* s1, const double* s2, uint64_t n, double N, float thr, float* d_max, double* d_sum,
            unsigned long long* d_cnt, hipStream_t s) {
    fill_u32(d_max, 0u, 1, s);
    fill_u64(d_sum, 0ull, 1, s);
    fill_u64(d_cnt, 0ull, 1, s);
    int g = grid_for(n);
    if (g > 2048) g = 2048;
    hipLaunchKernelGGL(k_stat_max, dim3(g), dim3(256), 0, s, s1, n, N, d_max);
    hipLaunchKernelGGL(k_stat_sum, dim3(g), dim3(256), 0, s, s1, s2, n, N, d_max, thr, d_sum, d_cnt);
    return (int) hipGetLastError();
}

// rescale of scorers by N_target / N_tracked (calculate_average, mqi_variables.hpp:51-66)
// dose[i] += w * (stat[i] - snap[i]) outside voxel 0 of every scenario block
__global__ void k_dose_from_stat(double* dose, const double* stat, const double* snap, uint64_t n, uint64_t nvox,
                                 double w) {
    for (uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t) gridDim.x * blockDim.x)
        if (i % nvox != 0) dose[i] += w * (stat[i] - snap[i]);
}

int
dose_from_stat(double* dose, const double* stat, const double* snap, uint64_t n, uint64_t nvox, double w,
               hipStream_t s) {
    if (n == 0 || nvox == 0) return 0;
    hipLaunchKernelGGL(k_dose_from_stat, dim3(grid_for(n)), dim3(256), 0, s, dose, stat, snap, n, nvox, w);
    return (int) hipGetLastError();
}

__global__ void k_scale(double* p, uint64_t n, double w) {
    for (uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t) gridDim.x * blockDim.x)
        p[i] *= w;
}
int
scale(double* p, uint64_t n, double w, hipStream_t s) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(256), 0, s, p, n, w);
    return (int) hipGetLastError();
}

// ------------------------------------------------------------------ Dij compaction
// One workgroup per item (spot of a scenario) scans its hash table, slots
// [off[item], off[item + 1]) of keys/vals.
// Pass A: per-item maximum.  Pass B: count entries >= rel_thr * max.
// Pass C: write (voxel, value) in slot order into the CSR row; rows are then
// sorted by voxel with a segmented radix sort.
__global__ void k_dij_max(const uint32_t* keys, const double* vals, const uint64_t* off, double* item_max) {
    const uint32_t  item = blockIdx.x;
    const uint64_t  cap  = off[item + 1] - off[item];
    const uint32_t* k    = keys + off[item];
    const double*   v    = vals + off[item];
    double          m   = 0.0;
    for (uint64_t i = threadIdx.x; i < cap; i += blockDim.x)
        if (k[i] != 0xffffffffu && v[i] > m) m = v[i];
    typedef hipcub::BlockReduce<double, 256> BR;
    __shared__ typename BR::TempStorage tmp;
    double                              bm = BR(tmp).Reduce(m, hipcub::Max());
    if (threadIdx.x == 0) item_max[item] = bm;
}

__global__ void k_dij_count(const uint32_t* keys, const double* vals, const uint64_t* off, const double* item_max,
                            double rel_thr, uint32_t* counts) {
    const uint32_t  item = blockIdx.x;
    const uint64_t  cap  = off[item + 1] - off[item];
    const uint32_t* k    = keys + off[item];
    const double*   v    = vals + off[item];
    const double    lim = item_max[item] * rel_thr;
    uint32_t        c   = 0;
    for (uint64_t i = threadIdx.x; i < cap; i += blockDim.x)
        if (k[i] != 0xffffffffu && v[i] > 0.0 && v[i] >= lim) ++c;
    typedef hipcub::BlockReduce<uint32_t, 256> BR;
    __shared__ typename BR::TempStorage tmp;
    uint32_t                            bc = BR(tmp).Sum(c);
    if (threadIdx.x == 0) counts[item] = bc;
}

__global__ void k_dij_write(const uint32_t* keys, const double* vals, const uint64_t* off, const double* item_max,
                            double rel_thr, const int64_t* row_start, int32_t* cols, float* out_vals) {
    const uint32_t  item = blockIdx.x;
    const uint64_t  cap  = off[item + 1] - off[item];
    const uint32_t* k    = keys + off[item];
    const double*   v    = vals + off[item];
    const double    lim = item_max[item] * rel_thr;
    int64_t         base = row_start[item];
    typedef hipcub::BlockScan<uint32_t, 256> BS;
    __shared__ typename BS::TempStorage tmp;
    __shared__ uint32_t                 chunk_total;
    for (uint64_t i0 = 0; i0 < cap; i0 += blockDim.x) {
        const uint64_t i    = i0 + threadIdx.x;
        const bool     keep = i < cap && k[i] != 0xffffffffu && v[i] > 0.0 && v[i] >= lim;
        uint32_t       pos, tot;
        BS(tmp).ExclusiveSum(keep ? 1u : 0u, pos, tot);
        if (keep) {
            cols[base + pos]     = (int32_t) k[i];
            out_vals[base + pos] = (float) v[i];
        }
        if (threadIdx.x == 0) chunk_total = tot;
        __syncthreads();
        base += chunk_total;
        __syncthreads();
    }
}

int
dij_compact(const uint32_t* keys, const double* vals, const uint64_t* off, uint32_t n_items, double rel_thr,
            double* item_max, uint32_t* counts, hipStream_t s) {
    hipLaunchKernelGGL(k_dij_max, dim3(n_items), dim3(256), 0, s, keys, vals, off, item_max);
    hipLaunchKernelGGL(k_dij_count, dim3(n_items), dim3(256), 0, s, keys, vals, off, item_max, rel_thr, counts);
    return (int) hipGetLastError();
}

int
dij_write(const uint32_t* keys, const double* vals, const uint64_t* off, uint32_t n_items, double rel_thr,
          const double* item_max, const int64_t* row_start, int32_t* cols, float* out_vals, hipStream_t s) {
    hipLaunchKernelGGL(k_dij_write, dim3(n_items), dim3(256), 0, s, keys, vals, off, item_max, rel_thr, row_start,
                       cols, out_vals);
    return (int) hipGetLastError();
}

int
segmented_sort(int32_t* keys, float* vals, int32_t* keys_tmp, float* vals_tmp, int64_t nnz, int n_seg,
               const int64_t* seg_begin, const int64_t* seg_end, void* scratch, size_t* scratch_bytes,
               hipStream_t s) {
    if (nnz > (int64_t) INT32_MAX) return (int) hipErrorInvalidValue;   // hipcub counts items in an int
    hipcub::DoubleBuffer<int32_t> k(keys, keys_tmp);
    hipcub::DoubleBuffer<float>   v(vals, vals_tmp);
    hipError_t e = hipcub::DeviceSegmentedRadixSort::SortPairs(scratch, *scratch_bytes, k, v, (int) nnz, n_seg,
                                                               seg_begin, seg_end, 0, 32, s);
    if (e != hipSuccess) return (int) e;
    if (scratch && k.Current() != keys) {
        hipMemcpyAsync(keys, k.Current(), sizeof(int32_t) * nnz, hipMemcpyDeviceToDevice, s);
        hipMemcpyAsync(vals, v.Current(), sizeof(float) * nnz, hipMemcpyDeviceToDevice, s);
    }
    return (int) hipGetLastError();
}

// ------------------------------------------------------------------ sparse products
// y = alpha * A x (+ y) for a CSR matrix A (float values, int32 columns) with
// FP64 accumulation, in the CSR-Adaptive form (J. L. Greathouse and M. Daga,
// "Efficient sparse matrix-vector multiplication on GPUs using the CSR storage
// format", SC'14): the nonzeros are tiled into segments at build time.
//   partial segment  a piece (<= MQI_SEG_LONG entries) of one long row; the
//                    wave streams it with 16-byte loads and leaves its sum in
//                    part[s]; a second pass adds the pieces of each split row
//                    in order.
//   packed segment   whole rows, <= MQI_SEG_SHORT entries and <= MQI_SEG_ROWS
//                    rows: the wave forms the products into LDS with
//                    coalesced loads, then each row is summed by L lanes
//                    (the largest power of two with L * rows <= 64; strided
//                    parts, then shuffles within the group).
// Every sum runs in a fixed order, so results do not depend on scheduling.
// seg_info[s] = first row | (partial << 31); seg_k[s] = first nonzero;
// entry n_seg of both is the sentinel (rows, nnz).
typedef float    f32x4 __attribute__((ext_vector_type(4)));
typedef int32_t  i32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// four columns in one load: int32 columns, or 16-bit ones (D w over at most
// 65536 spots)
template<typename TC> struct col4;
template<> struct col4<int32_t> { typedef i32x4 type; };
template<> struct col4<uint16_t> { typedef u16x4 type; };

// a packed segment is loaded as at most 4 x 64 vectors of four entries
static_assert(MQI_SEG_SHORT <= 1024, "packed segments are loaded as at most 256 vectors");

// OBJ: instead of y, the quadratic objective of y (quad_obj): per row
// c_r (y_r - p_r)^2 summed over each packed segment into o.part[s] (0 for a
// long-row segment, whose rows k_csr_split_rows finishes), and
// o.g[r] = two_inv_n c_r (y_r - p_r) when o.g is set
template<typename TY, typename TC, bool OBJ>
__global__ void __launch_bounds__(64) k_csr_adaptive(const int64_t* __restrict__ indptr,
                                                     const TC* __restrict__ col,
                                                     const float* __restrict__ val, const float* __restrict__ x,
                                                     const int64_t* __restrict__ seg_k,
                                                     const uint32_t* __restrict__ seg_info, int64_t n_seg,
                                                     double* __restrict__ part, TY* __restrict__ y, double alpha,
                                                     int accumulate, quad_obj o,
                                                     const int32_t* __restrict__ rowmap,
                                                     const uint32_t* __restrict__ nonempty, int64_t n_y,
                                                     int zero_blocks) {
    __shared__ double buf[MQI_SEG_SHORT];
    const int lane = threadIdx.x;
    // the first zero_blocks workgroups write y = 0 at the elements no row
    // reaches (bit clear in nonempty), alongside the segments' workgroups
    if ((int) blockIdx.x < zero_blocks) {
        // eight elements per lane and pass, their mask words loaded together
        const int64_t zs = (int64_t) zero_blocks * 64;
        for (int64_t v0 = (int64_t) blockIdx.x * 64 + lane; v0 < n_y; v0 += 8 * zs) {
            uint32_t bits[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) bits[j] = (v0 + j * zs < n_y) ? nonempty[(v0 + j * zs) >> 5] : 0xffffffffu;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int64_t v = v0 + j * zs;
                if (v < n_y && !((bits[j] >> (v & 31)) & 1u)) y[v] = (TY) 0;
            }
        }
        return;
    }
    // plain grid-stride walk: a remap giving each XCD a contiguous run of
    // segments measured ~7 % slower on the 5000-spot Dij
    const int64_t stride = (int64_t) gridDim.x - zero_blocks;
    for (int64_t s = (int64_t) blockIdx.x - zero_blocks; s < n_seg; s += stride) {
        const uint32_t info = seg_info[s];
        const int64_t  kb = seg_k[s], ke = seg_k[s + 1];
        if (info >> 31) {
            double  acc = 0.0;
            int64_t ka  = (kb + 3) & ~(int64_t) 3;
            if (ka > ke) ka = ke;
            if (lane < ka - kb) acc += (double) val[kb + lane] * (double) x[col[kb + lane]];
            const int64_t n4 = (ke - ka) >> 2;
            typedef typename col4<TC>::type c4t;
            const f32x4*  v4 = reinterpret_cast<const f32x4*>(val + ka);
            const c4t*    c4 = reinterpret_cast<const c4t*>(col + ka);
            int64_t       i  = lane;
            for (; i + 192 < n4; i += 256) {
                f32x4 v[4];
                c4t   c[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    v[u] = __builtin_nontemporal_load(v4 + i + 64 * u);
                    c[u] = __builtin_nontemporal_load(c4 + i + 64 * u);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    acc += (double) v[u].x * (double) x[c[u].x];
                    acc += (double) v[u].y * (double) x[c[u].y];
                    acc += (double) v[u].z * (double) x[c[u].z];
                    acc += (double) v[u].w * (double) x[c[u].w];
                }
            }
            for (; i < n4; i += 64) {
                const f32x4 v = __builtin_nontemporal_load(v4 + i);
                const c4t   c = __builtin_nontemporal_load(c4 + i);
                acc += (double) v.x * (double) x[c.x];
                acc += (double) v.y * (double) x[c.y];
                acc += (double) v.z * (double) x[c.z];
                acc += (double) v.w * (double) x[c.w];
            }
            const int64_t kt = ka + (n4 << 2);
            if (lane < ke - kt) acc += (double) val[kt + lane] * (double) x[col[kt + lane]];
            for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
            if (lane == 0) {
                part[s] = acc;
                if (OBJ) o.part[s] = 0.0;
            }
        } else {
            const int64_t r0 = info, r1 = seg_info[s + 1] & 0x7fffffffu;
            // products into LDS: the segment's aligned body as 16-byte loads
            // (four consecutive entries per lane, the whole segment in flight
            // at once: at most 1024 entries, 256 vectors), head and tail of
            // up to three entries one per lane
            typedef typename col4<TC>::type c4t;
            int64_t ka = (kb + 3) & ~(int64_t) 3;
            if (ka > ke) ka = ke;
            const int64_t n4 = (ke - ka) >> 2;
            const int64_t kt = ka + (n4 << 2);
            const f32x4*  v4 = reinterpret_cast<const f32x4*>(val + ka);
            const c4t*    c4 = reinterpret_cast<const c4t*>(col + ka);
            f32x4         v[4];
            c4t           c[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int64_t i = lane + 64 * u;
                if (i < n4) {
                    v[u] = __builtin_nontemporal_load(v4 + i);
                    c[u] = __builtin_nontemporal_load(c4 + i);
                }
            }
            if (lane < ka - kb) buf[lane] = (double) val[kb + lane] * (double) x[col[kb + lane]];
            if (lane < ke - kt) buf[kt - kb + lane] = (double) val[kt + lane] * (double) x[col[kt + lane]];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int64_t i = lane + 64 * u;
                if (i < n4) {
                    double* b = buf + (ka - kb) + 4 * i;
                    b[0]      = (double) v[u].x * (double) x[c[u].x];
                    b[1]      = (double) v[u].y * (double) x[c[u].y];
                    b[2]      = (double) v[u].z * (double) x[c[u].z];
                    b[3]      = (double) v[u].w * (double) x[c[u].w];
                }
            }
            __syncthreads();
            // L lanes per row (a power of two, L * rows <= 64): each sums a
            // strided part, then the group adds its parts with shuffles
            const int64_t nr = r1 - r0;
            int           L  = 1;
            while (L < 64 && 2 * L * nr <= 64) L *= 2;
            const int sub = lane & (L - 1);
            double    obj = 0.0;   // this lane's rows' objective terms, in row order
            for (int64_t rb = r0; rb < r1; rb += 64 / L) {
                const int64_t r   = rb + lane / L;
                double        acc = 0.0;
                // the row's prescription and weight are loaded with its
                // pointers, not after its sum (one round trip per row)
                // row r of the matrix is element v of y (rowmap: the voxel of
                // each row of a matrix with its empty rows left out)
                const int64_t v  = (rowmap && r < r1) ? (int64_t) rowmap[r] : r;
                double        pr = 0.0, cw = 1.0;
                if (OBJ && r < r1 && sub == 0) {
                    pr = o.p[v];
                    if (o.c) cw = (double) o.c[v];
                }
                if (r < r1) {
                    const int64_t e = indptr[r + 1] - kb;
                    for (int64_t k = indptr[r] - kb + sub; k < e; k += L) acc += buf[k];
                }
                for (int off = L >> 1; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
                if (r < r1 && sub == 0) {
                    const double out = alpha * acc;
                    if (OBJ) {
                        const double res = out - pr;
                        obj += cw * res * res;
                        if (o.g) o.g[v] = (float) (o.two_inv_n * cw * res);
                    } else {
                        y[v] = accumulate ? (TY) ((double) y[v] + out) : (TY) out;
                    }
                }
            }
            if (OBJ) {
                for (int off = 32; off > 0; off >>= 1) obj += __shfl_xor(obj, off);
                if (lane == 0) o.part[s] = obj;
            }
            __syncthreads();
        }
    }
}

// rows split over partial segments [split_seg[2i], split_seg[2i+1]):
// y[r] = alpha * sum of the pieces in order (+ y[r])
template<typename TY, bool OBJ>
__global__ void k_csr_split_rows(const uint32_t* __restrict__ seg_info, const int32_t* __restrict__ split_seg,
                                 int64_t n_split, const double* __restrict__ part, TY* __restrict__ y, double alpha,
                                 int accumulate, quad_obj o, const int32_t* __restrict__ rowmap) {
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_split) return;
    const int32_t s0 = split_seg[2 * i], s1 = split_seg[2 * i + 1];
    const int64_t rr = seg_info[s0] & 0x7fffffffu;
    const int64_t r  = rowmap ? (int64_t) rowmap[rr] : rr;   // the element of y
    double        acc = 0.0;
    for (int32_t s = s0; s < s1; ++s) acc += part[s];
    const double out = alpha * acc;
    if (OBJ) {   // the row's objective term after the packed segments' ones
        const double res = out - o.p[r];
        const double cw  = o.c ? (double) o.c[r] : 1.0;
        o.part[o.n_seg + i] = cw * res * res;
        if (o.g) o.g[r] = (float) (o.two_inv_n * cw * res);
    } else {
        y[r] = accumulate ? (TY) ((double) y[r] + out) : (TY) out;
    }
}

// f = inv_n * (o.part[0] + ... + o.part[n - 1]) in a fixed order (one
// workgroup: strided partial sums, then a block reduction)
// Fixed-order sum of the objective's terms in two launches: workgroup b
// sums chunk b of part (each thread its elements in index order, four loads
// in flight) into out[b]; a single workgroup then sums the
// MQI_QUAD_SUM_BLOCKS chunk sums and scales by inv_n.
__global__ void k_quad_chunks(const double* __restrict__ part, int64_t n, double* __restrict__ out) {
    const int64_t ch = (n + gridDim.x - 1) / gridDim.x;
    const int64_t b0 = (int64_t) blockIdx.x * ch;
    const int64_t b1 = b0 + ch < n ? b0 + ch : n;
    double        acc = 0.0;
    int64_t       i   = b0 + threadIdx.x;
    for (; i + 768 < b1; i += 1024) {
        const double a0 = part[i], a1 = part[i + 256], a2 = part[i + 512], a3 = part[i + 768];
        acc += a0;
        acc += a1;
        acc += a2;
        acc += a3;
    }
    for (; i < b1; i += 256) acc += part[i];
    typedef hipcub::BlockReduce<double, 256> BR;
    __shared__ typename BR::TempStorage tmp;
    const double t = BR(tmp).Sum(acc);
    if (threadIdx.x == 0) out[blockIdx.x] = t;
}

__global__ void k_quad_sum(const double* __restrict__ part, int64_t n, double inv_n, double* __restrict__ f) {
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) acc += part[i];
    typedef hipcub::BlockReduce<double, 256> BR;
    __shared__ typename BR::TempStorage tmp;
    const double t = BR(tmp).Sum(acc);
    if (threadIdx.x == 0) *f = inv_n * t;
}

template<typename TC>
int
csr_adaptive_t(const int64_t* indptr, const TC* col, const float* val, const float* x, const int64_t* seg_k,
               const uint32_t* seg_info, int64_t n_seg, const int32_t* split_seg, int64_t n_split, double* part,
               void* y, int y_is_double, double alpha, int accumulate, const int32_t* rowmap,
               const uint32_t* nonempty, int64_t n_y, hipStream_t s) {
    if (rowmap && !accumulate && (!nonempty || n_y <= 0)) return (int) hipErrorInvalidValue;
    if (n_seg <= 0) return 0;
    const int blocks = (int) (n_seg < 262144 ? n_seg : 262144);
    const int zb     = (rowmap && !accumulate) ? MQI_ZERO_BLOCKS : 0;
    const quad_obj o{};
    if (y_is_double)
        hipLaunchKernelGGL((k_csr_adaptive<double, TC, false>), dim3(blocks + zb), dim3(64), 0, s, indptr, col, val,
                           x, seg_k, seg_info, n_seg, part, (double*) y, alpha, accumulate, o, rowmap, nonempty, n_y,
                           zb);
    else
        hipLaunchKernelGGL((k_csr_adaptive<float, TC, false>), dim3(blocks + zb), dim3(64), 0, s, indptr, col, val,
                           x, seg_k, seg_info, n_seg, part, (float*) y, alpha, accumulate, o, rowmap, nonempty, n_y,
                           zb);
    if (n_split > 0) {
        const int b = (int) ((n_split + 255) / 256);
        if (y_is_double)
            hipLaunchKernelGGL((k_csr_split_rows<double, false>), dim3(b), dim3(256), 0, s, seg_info, split_seg,
                               n_split, part, (double*) y, alpha, accumulate, o, rowmap);
        else
            hipLaunchKernelGGL((k_csr_split_rows<float, false>), dim3(b), dim3(256), 0, s, seg_info, split_seg,
                               n_split, part, (float*) y, alpha, accumulate, o, rowmap);
    }
    return (int) hipGetLastError();
}

// the objective's terms of the elements of y no row reaches (bit v of
// nonempty clear): c p^2 per element, summed per workgroup of a fixed grid
// into part[0 .. MQI_QUAD_EMPTY_BLOCKS), and g = -two_inv_n c p there
__global__ void k_quad_empty(const uint32_t* __restrict__ nonempty, int64_t n, quad_obj o, double* __restrict__ part) {
    // eight elements per thread and pass, their loads issued together (one
    // load in flight per thread left the pass latency-bound); the thread's
    // sum runs over its elements in index order
    constexpr int U      = 8;
    const int64_t stride = (int64_t) gridDim.x * blockDim.x;
    double        acc    = 0.0;
    for (int64_t v0 = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; v0 < n; v0 += U * stride) {
        uint32_t bits[U];
#pragma unroll
        for (int j = 0; j < U; ++j) {
            const int64_t v = v0 + j * stride;
            bits[j]         = v < n ? nonempty[v >> 5] : 0xffffffffu;
        }
        double p[U], cw[U];
#pragma unroll
        for (int j = 0; j < U; ++j) {
            const int64_t v     = v0 + j * stride;
            const bool    empty = v < n && !((bits[j] >> (v & 31)) & 1u);
            p[j]                = empty ? o.p[v] : 0.0;
            cw[j]               = (empty && o.c) ? (double) o.c[v] : 1.0;
        }
#pragma unroll
        for (int j = 0; j < U; ++j) {
            const int64_t v = v0 + j * stride;
            if (v >= n || ((bits[j] >> (v & 31)) & 1u)) continue;
            acc += cw[j] * p[j] * p[j];
            if (o.g) o.g[v] = (float) (-o.two_inv_n * cw[j] * p[j]);
        }
    }
    typedef hipcub::BlockReduce<double, 256> BR;
    __shared__ typename BR::TempStorage tmp;
    const double t = BR(tmp).Sum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

template<typename TC>
int
csr_quadratic_t(const int64_t* indptr, const TC* col, const float* val, const float* x, const int64_t* seg_k,
                const uint32_t* seg_info, int64_t n_seg, const int32_t* split_seg, int64_t n_split, double* part,
                double alpha, quad_obj o, double inv_n, double* f, const int32_t* rowmap, const uint32_t* nonempty,
                int64_t n_y, hipStream_t s) {
    if (n_seg <= 0 || !o.part || !o.p || !f) return (int) hipErrorInvalidValue;
    if (rowmap && !nonempty) return (int) hipErrorInvalidValue;
    o.n_seg          = n_seg;
    const int blocks = (int) (n_seg < 262144 ? n_seg : 262144);
    int64_t   n_part = n_seg + n_split;
    if (rowmap) {   // the elements without a row first (their g is not written again)
        hipLaunchKernelGGL(k_quad_empty, dim3(MQI_QUAD_EMPTY_BLOCKS), dim3(256), 0, s, nonempty, n_y, o,
                           o.part + n_part);
        n_part += MQI_QUAD_EMPTY_BLOCKS;
    }
    hipLaunchKernelGGL((k_csr_adaptive<double, TC, true>), dim3(blocks), dim3(64), 0, s, indptr, col, val, x, seg_k,
                       seg_info, n_seg, part, (double*) nullptr, alpha, 0, o, rowmap, (const uint32_t*) nullptr,
                       (int64_t) 0, 0);
    if (n_split > 0)
        hipLaunchKernelGGL((k_csr_split_rows<double, true>), dim3((unsigned) ((n_split + 255) / 256)), dim3(256), 0, s,
                           seg_info, split_seg, n_split, part, (double*) nullptr, alpha, 0, o, rowmap);
    double* chunks = o.part + n_seg + n_split + MQI_QUAD_EMPTY_BLOCKS;
    hipLaunchKernelGGL(k_quad_chunks, dim3(MQI_QUAD_SUM_BLOCKS), dim3(256), 0, s, o.part, n_part, chunks);
    hipLaunchKernelGGL(k_quad_sum, dim3(1), dim3(256), 0, s, chunks, (int64_t) MQI_QUAD_SUM_BLOCKS, inv_n, f);
    return (int) hipGetLastError();
}

int
csr_adaptive(const int64_t* indptr, const int32_t* col, const float* val, const float* x, const int64_t* seg_k,
             const uint32_t* seg_info, int64_t n_seg, const int32_t* split_seg, int64_t n_split, double* part,
             void* y, int y_is_double, double alpha, int accumulate, const int32_t* rowmap, const uint32_t* nonempty,
             int64_t n_y, hipStream_t s) {
    return csr_adaptive_t(indptr, col, val, x, seg_k, seg_info, n_seg, split_seg, n_split, part, y, y_is_double,
                          alpha, accumulate, rowmap, nonempty, n_y, s);
}
int
csr_adaptive16(const int64_t* indptr, const uint16_t* col, const float* val, const float* x, const int64_t* seg_k,
               const uint32_t* seg_info, int64_t n_seg, const int32_t* split_seg, int64_t n_split, double* part,
               void* y, int y_is_double, double alpha, int accumulate, const int32_t* rowmap, const uint32_t* nonempty,
             int64_t n_y, hipStream_t s) {
    return csr_adaptive_t(indptr, col, val, x, seg_k, seg_info, n_seg, split_seg, n_split, part, y, y_is_double,
                          alpha, accumulate, rowmap, nonempty, n_y, s);
}

int
csr_quadratic(const int64_t* indptr, const int32_t* col, const float* val, const float* x, const int64_t* seg_k,
              const uint32_t* seg_info, int64_t n_seg, const int32_t* split_seg, int64_t n_split, double* part,
              double alpha, quad_obj o, double inv_n, double* f, const int32_t* rowmap, const uint32_t* nonempty,
              int64_t n_y, hipStream_t s) {
    return csr_quadratic_t(indptr, col, val, x, seg_k, seg_info, n_seg, split_seg, n_split, part, alpha, o, inv_n, f,
                           rowmap, nonempty, n_y, s);
}
int
csr_quadratic16(const int64_t* indptr, const uint16_t* col, const float* val, const float* x, const int64_t* seg_k,
                const uint32_t* seg_info, int64_t n_seg, const int32_t* split_seg, int64_t n_split, double* part,
                double alpha, quad_obj o, double inv_n, double* f, const int32_t* rowmap, const uint32_t* nonempty,
                int64_t n_y, hipStream_t s) {
    return csr_quadratic_t(indptr, col, val, x, seg_k, seg_info, n_seg, split_seg, n_split, part, alpha, o, inv_n, f,
                           rowmap, nonempty, n_y, s);
}

// ------------------------------------------------------------------ D^T g over column deltas
// Index compression of the CSR stream (delta coding of the column indices,
// after K. Kourtis, G. Goumas and N. Koziris, "Optimizing sparse
// matrix-vector multiplication using index and value compression", CF'08):
// a nonzero keeps its value and the difference of its column to the
// previous nonzero of its row.  The nonzeros are cut into pieces of <=
// MQI_PIECE entries of one row, also where a row skips more than 65535
// columns; a piece keeps the column its first delta adds to (pbase).  For
// the product the pieces are restaged: padded with zero values and zero
// deltas to whole batches of 256 nonzeros, each batch lane-major (the
// 16-byte value load of lane l holds the batch's nonzeros l, l + 64,
// l + 128 and l + 192, its 4-byte delta load their deltas), so that each of
// the four gathers of a batch reads the columns of 64 consecutive nonzeros:
// few distinct cache lines per gather instruction instead of every fourth
// column of 256.  A delta takes one byte; the 1.2 % of deltas of 255 and
// more (the steps between the rows and slices of a spot's footprint on the
// 5000-spot Dij) are escapes, kept in order in the piece's escape list:
// 5.02 bytes per nonzero instead of 8.
// One wave per piece rebuilds the columns with wave prefix sums of the
// deltas and forms the FP64 sum of the piece into part[piece]; a second pass
// adds each row's pieces in a fixed order, so results do not depend on
// scheduling or on the order the pieces run in.  (Offsets from each batch's
// first column instead of deltas need no prefix sums, but the sparse halo of
// a spot spans more than 65535 columns in 256 nonzeros so often that the
// cuts it needs made the pieces tiny: 13x slower on the 5000-spot Dij.)

// dcol[k] = col[k] - col[k - 1] where that is in [1, 65535], else 0; the
// positions with a larger step are listed in gaps[] (count in *n_gaps, at
// most cap written).  Row starts are cleared afterwards (k_row_starts).
__global__ void k_deltas(const int32_t* __restrict__ col, int64_t nnz, uint16_t* __restrict__ dcol,
                         unsigned long long* __restrict__ n_gaps, int64_t* __restrict__ gaps, int64_t cap) {
    for (int64_t k = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; k < nnz; k += (int64_t) gridDim.x * blockDim.x) {
        const int64_t d = k > 0 ? (int64_t) col[k] - (int64_t) col[k - 1] : 0;
        const bool    ok = d >= 1 && d <= 65535;
        dcol[k]          = ok ? (uint16_t) d : (uint16_t) 0;
        if (!ok && d > 65535) {
            const unsigned long long at = atomicAdd(n_gaps, 1ull);
            if ((int64_t) at < cap) gaps[at] = k;
        }
    }
}

__global__ void k_row_starts(const int64_t* __restrict__ indptr, int64_t rows, uint16_t* __restrict__ dcol) {
    const int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (r < rows && indptr[r + 1] > indptr[r]) dcol[indptr[r]] = 0;
}

// pbase[p] = col[pk[p]] - dcol[pk[p]]: the column the piece's first delta
// adds to (0 for the empty piece of an empty row)
__global__ void k_piece_base(const int32_t* __restrict__ col, const uint16_t* __restrict__ dcol,
                             const int64_t* __restrict__ pk, int64_t n_p, int32_t* __restrict__ pbase) {
    const int64_t p = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n_p) pbase[p] = pk[p] < pk[p + 1] ? col[pk[p]] - (int32_t) dcol[pk[p]] : 0;
}

// deltas of 255 and more are escapes: the byte holds 255 and the piece's
// escape list (in nonzero order) the delta
#define MQI_ESC 255

// ecount[p]: the escapes of piece p; one workgroup of 256 per piece
__global__ void k_piece_escapes(const uint16_t* __restrict__ dcol, const int64_t* __restrict__ pk, int64_t n_p,
                                uint32_t* __restrict__ ecount) {
    typedef hipcub::BlockReduce<uint32_t, 256> BR;
    __shared__ typename BR::TempStorage tmp;
    for (int64_t p = blockIdx.x; p < n_p; p += gridDim.x) {
        const int64_t kb = pk[p], len = pk[p + 1] - kb;
        uint32_t      n  = 0;
        for (int64_t i = threadIdx.x; i < len; i += 256) n += dcol[kb + i] >= MQI_ESC ? 1u : 0u;
        const uint32_t t = BR(tmp).Sum(n);
        if (threadIdx.x == 0) ecount[p] = t;
        __syncthreads();
    }
}

// pb[p]: the piece's first staged slot (a multiple of 256); pb[n_p]: total.
// Nonzero i of a piece goes to batch i / 256, lane i % 64, byte (i / 64) % 4
// of the lane's word; its escape (if any) to esc[ebase[p] + the escapes
// before it in the piece].  One workgroup of 256 per piece.
__global__ void k_stage_pieces(const float* __restrict__ val, const uint16_t* __restrict__ dcol,
                               const int64_t* __restrict__ pk, const int64_t* __restrict__ pb,
                               const uint32_t* __restrict__ ebase, int64_t n_p, float* __restrict__ sval,
                               uint8_t* __restrict__ sd8, uint16_t* __restrict__ esc) {
    typedef hipcub::BlockScan<uint32_t, 256> BS;
    __shared__ typename BS::TempStorage tmp;
    for (int64_t p = blockIdx.x; p < n_p; p += gridDim.x) {
        const int64_t kb = pk[p], len = pk[p + 1] - kb, base = pb[p], padded = pb[p + 1] - base;
        uint32_t      e  = ebase[p];
        for (int64_t i0 = 0; i0 < padded; i0 += 256) {   // padded is a multiple of 256
            const int64_t  i   = i0 + threadIdx.x;
            const int64_t  j   = i & 255;
            const int64_t  dst = base + i0 + ((j & 63) << 2) + (j >> 6);
            const uint16_t d   = i < len ? dcol[kb + i] : (uint16_t) 0;
            const uint32_t f   = d >= MQI_ESC ? 1u : 0u;
            uint32_t       off, tot;
            BS(tmp).ExclusiveSum(f, off, tot);
            sval[dst] = i < len ? val[kb + i] : 0.0f;
            sd8[dst]  = (uint8_t) (f ? MQI_ESC : d);
            if (f) esc[e + off] = d;
            e += tot;
            __syncthreads();
        }
    }
}

// inclusive prefix sum over the wave: DPP row shifts within the four rows
// of 16 lanes, then the row broadcasts of lanes 15 and 31
MQI_DI uint32_t wave_scan_dpp(uint32_t x) {
    x += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) x, 0x111, 0xf, 0xf, false);   // row_shr:1
    x += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) x, 0x112, 0xf, 0xf, false);   // row_shr:2
    x += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) x, 0x114, 0xf, 0xf, false);   // row_shr:4
    x += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) x, 0x118, 0xf, 0xf, false);   // row_shr:8
    x += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) x, 0x142, 0xa, 0xf, false);   // row_bcast:15
    x += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) x, 0x143, 0xc, 0xf, false);   // row_bcast:31
    return x;
}

// One wave per piece, MQI_PIECE_BATCHES batches (2048 nonzeros) in flight;
// the pieces run in the order of their headers (first columns ascending:
// pieces that run at the same time gather from the same stretch of x).
// The wave keeps the next 64 escapes of its piece in a register (one per
// lane) and hands them to the lanes whose byte is an escape.
__global__ void __launch_bounds__(64 * MQI_PIECE_WAVES)
    k_pieces_staged(const float* __restrict__ sval, const uint8_t* __restrict__ sd8, const uint16_t* __restrict__ esc,
                    const float* __restrict__ x, const piece_hdr* __restrict__ hdr, int64_t n_p,
                    double* __restrict__ part) {
    constexpr int U    = MQI_PIECE_BATCHES;
    const int     lane = threadIdx.x & 63;
    const int64_t w    = (int64_t) blockIdx.x * MQI_PIECE_WAVES + (threadIdx.x >> 6);
    const int64_t nw   = (int64_t) gridDim.x * MQI_PIECE_WAVES;
    for (int64_t i = w; i < n_p; i += nw) {
        // the whole header in two 16-byte loads
        const u32x4     h  = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(hdr + i));
        const u32x4     h2 = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(hdr + i) + 1);
        const int64_t   p  = (int32_t) h.w;
        const int64_t   b0 = h.x, nb = h.y;
        const f32x4*    v4 = reinterpret_cast<const f32x4*>(sval) + b0 * 64;
        const uint32_t* d4 = reinterpret_cast<const uint32_t*>(sd8) + b0 * 64;
        uint32_t        run = h.z;
        // escapes: esc[e_ptr + lane] in e_reg, e_off of them used
        uint32_t        e_ptr = h2.x, e_off = 0;
        uint32_t        e_reg = esc[e_ptr + lane];   // the list is padded by 64
        double          acc   = 0.0;
        for (int64_t b = 0; b < nb; b += U) {
            f32x4    v[U];
            uint32_t d[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (b + u < nb) {
                    v[u] = __builtin_nontemporal_load(v4 + (b + u) * 64 + lane);
                    d[u] = __builtin_nontemporal_load(d4 + (b + u) * 64 + lane);
                } else {
                    v[u] = f32x4{ 0.0f, 0.0f, 0.0f, 0.0f };
                    d[u] = 0u;
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                // quarter q of the batch (nonzeros 64 q .. 64 q + 63, one per
                // lane): byte q of the lane's word, its columns a prefix sum
                // over the wave
                uint32_t c[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    uint32_t                 dq = (d[u] >> (8 * q)) & 0xffu;
                    const unsigned long long m  = __ballot(dq == MQI_ESC);
                    if (m) {
                        const uint32_t n = (uint32_t) __popcll(m);
                        if (e_off + n > 64) {
                            e_ptr += e_off;
                            e_off = 0;
                            e_reg = esc[e_ptr + lane];
                        }
                        const uint32_t k  = e_off + (uint32_t) __builtin_amdgcn_mbcnt_hi(
                                                       (uint32_t) (m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t) m, 0u));
                        const uint32_t ev = (uint32_t) __shfl((int) e_reg, (int) (k & 63u));
                        if (dq == MQI_ESC) dq = ev;
                        e_off += n;
                    }
                    c[q] = dq;
                }
                const uint32_t s0 = wave_scan_dpp(c[0]);
                const uint32_t t0 = (uint32_t) __builtin_amdgcn_readlane((int) s0, 63);
                const uint32_t s1 = wave_scan_dpp(c[1]);
                const uint32_t t1 = (uint32_t) __builtin_amdgcn_readlane((int) s1, 63);
                const uint32_t s2 = wave_scan_dpp(c[2]);
                const uint32_t t2 = (uint32_t) __builtin_amdgcn_readlane((int) s2, 63);
                const uint32_t s3 = wave_scan_dpp(c[3]);
                const uint32_t t3 = (uint32_t) __builtin_amdgcn_readlane((int) s3, 63);
                acc += (double) v[u].x * (double) x[run + s0];
                acc += (double) v[u].y * (double) x[run + t0 + s1];
                acc += (double) v[u].z * (double) x[run + t0 + t1 + s2];
                acc += (double) v[u].w * (double) x[run + t0 + t1 + t2 + s3];
                run += t0 + t1 + t2 + t3;
            }
        }
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
        if (lane == 0) part[p] = acc;
    }
}

// y[r] = alpha * (the pieces row_piece[r] .. row_piece[r + 1] - 1) (+ y[r]):
// one wave per row, lane l takes pieces q0 + l + 64 k, the wave adds them in
// a fixed tree order
template<typename TY>
__global__ void k_piece_rows(const int32_t* __restrict__ row_piece, int64_t rows, const double* __restrict__ part,
                             TY* __restrict__ y, double alpha, int accumulate) {
    const int64_t r    = (int64_t) blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int     lane = threadIdx.x & 63;
    if (r >= rows) return;
    const int32_t q0 = row_piece[r], q1 = row_piece[r + 1];
    double        acc = 0.0;
    for (int32_t q = q0 + lane; q < q1; q += 64) acc += part[q];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) {
        const double out = alpha * acc;
        y[r]             = accumulate ? (TY) ((double) y[r] + out) : (TY) out;
    }
}

int
deltas16(const int64_t* indptr, const int32_t* col, int64_t rows, int64_t nnz, uint16_t* dcol,
         unsigned long long* n_gaps, int64_t* gaps, int64_t cap, hipStream_t s) {
    if (nnz <= 0) return 0;
    int e = fill_u64(n_gaps, 0ull, 1, s);
    if (e) return e;
    hipLaunchKernelGGL(k_deltas, dim3(grid_for((uint64_t) nnz)), dim3(256), 0, s, col, nnz, dcol, n_gaps, gaps, cap);
    if (rows > 0)
        hipLaunchKernelGGL(k_row_starts, dim3((unsigned) ((rows + 255) / 256)), dim3(256), 0, s, indptr, rows, dcol);
    return (int) hipGetLastError();
}

int
piece_base(const int32_t* col, const uint16_t* dcol, const int64_t* pk, int64_t n_p, int32_t* pbase, hipStream_t s) {
    if (n_p <= 0) return 0;
    hipLaunchKernelGGL(k_piece_base, dim3((unsigned) ((n_p + 255) / 256)), dim3(256), 0, s, col, dcol, pk, n_p, pbase);
    return (int) hipGetLastError();
}

int
piece_escapes(const uint16_t* dcol, const int64_t* pk, int64_t n_p, uint32_t* ecount, hipStream_t s) {
    if (n_p <= 0) return 0;
    hipLaunchKernelGGL(k_piece_escapes, dim3((unsigned) (n_p < 65536 ? n_p : 65536)), dim3(256), 0, s, dcol, pk, n_p,
                       ecount);
    return (int) hipGetLastError();
}

int
stage_pieces(const float* val, const uint16_t* dcol, const int64_t* pk, const int64_t* pb, const uint32_t* ebase,
             int64_t n_p, float* sval, uint8_t* sd8, uint16_t* esc, hipStream_t s) {
    if (n_p <= 0) return 0;
    hipLaunchKernelGGL(k_stage_pieces, dim3((unsigned) (n_p < 65536 ? n_p : 65536)), dim3(256), 0, s, val, dcol, pk, pb,
                       ebase, n_p, sval, sd8, esc);
    return (int) hipGetLastError();
}

int
pieces_staged(const float* sval, const uint8_t* sd8, const uint16_t* esc, const float* x, const piece_hdr* hdr,
              int64_t n_p, const int32_t* row_piece, int64_t rows, double* part, void* y, int y_is_double,
              double alpha, int accumulate, hipStream_t s) {
    if (n_p > 0) {
        const int64_t b = (n_p + MQI_PIECE_WAVES - 1) / MQI_PIECE_WAVES;
        hipLaunchKernelGGL(k_pieces_staged, dim3((unsigned) (b < 65536 ? b : 65536)), dim3(64 * MQI_PIECE_WAVES), 0,
                           s, sval, sd8, esc, x, hdr, n_p, part);
    }
    if (rows > 0) {
        const unsigned b = (unsigned) ((rows + 3) / 4);
        if (y_is_double)
            hipLaunchKernelGGL(k_piece_rows<double>, dim3(b), dim3(256), 0, s, row_piece, rows, part, (double*) y,
                               alpha, accumulate);
        else
            hipLaunchKernelGGL(k_piece_rows<float>, dim3(b), dim3(256), 0, s, row_piece, rows, part, (float*) y,
                               alpha, accumulate);
    }
    return (int) hipGetLastError();
}

// low 16 bits of int32 columns (D w over at most 65536 spots)
__global__ void k_low16(const int32_t* __restrict__ col, int64_t nnz, uint16_t* __restrict__ out) {
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += (int64_t) gridDim.x * blockDim.x)
        out[i] = (uint16_t) (col[i] & 0xffff);
}

int
low16(const int32_t* col, int64_t nnz, uint16_t* out, hipStream_t s) {
    if (nnz <= 0) return 0;
    hipLaunchKernelGGL(k_low16, dim3(grid_for((uint64_t) nnz)), dim3(256), 0, s, col, nnz, out);
    return (int) hipGetLastError();
}

// Compacted x for D^T g: xc[i] = x[nzvox[i]] (nzvox: the voxels that appear
// in some row, ascending); the rows' columns renumbered into that list.
// four entries per thread: 16-byte index loads and stores
__global__ void k_gather_x(const float* __restrict__ x, const int32_t* __restrict__ nzvox, int64_t n,
                           float* __restrict__ xc) {
    const int64_t n4 = n >> 2;
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (int64_t) gridDim.x * blockDim.x) {
        const i32x4 v = __builtin_nontemporal_load(reinterpret_cast<const i32x4*>(nzvox) + i);
        f32x4       o;
        o.x = x[v.x];
        o.y = x[v.y];
        o.z = x[v.z];
        o.w = x[v.w];
        __builtin_nontemporal_store(o, reinterpret_cast<f32x4*>(xc) + i);
    }
    const int64_t t = (n4 << 2) + (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x == 0 && t < n) xc[t] = x[nzvox[t]];
}
__global__ void k_renumber(const int32_t* __restrict__ col, int64_t nnz, const int32_t* __restrict__ rank,
                           int32_t* __restrict__ ccol) {
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += (int64_t) gridDim.x * blockDim.x)
        ccol[i] = rank[col[i]];
}

int
gather_x(const float* x, const int32_t* nzvox, int64_t n, float* xc, hipStream_t s) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_gather_x, dim3(grid_for((uint64_t) (n + 3) / 4)), dim3(256), 0, s, x, nzvox, n, xc);
    return (int) hipGetLastError();
}
int
renumber(const int32_t* col, int64_t nnz, const int32_t* rank, int32_t* ccol, hipStream_t s) {
    if (nnz <= 0) return 0;
    hipLaunchKernelGGL(k_renumber, dim3(grid_for((uint64_t) nnz)), dim3(256), 0, s, col, nnz, rank, ccol);
    return (int) hipGetLastError();
}

// Transpose of a CSR matrix: voxel-major copy for D w (row lengths by
// counting, then the entries in a fixed order).
__global__ void k_count_cols(const int32_t* col, int64_t nnz, uint32_t* cnt) {
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += (int64_t) gridDim.x * blockDim.x)
        atomicAdd(cnt + col[i], 1u);
}
// The voxel-major copy lists each voxel's spots in ascending order whatever
// the scheduling: every nonzero becomes (voxel; spot, value) with the spot and
// the value's bits packed in 64 bits, and a stable LSD radix sort on the voxel
// keeps the spot-major order (spots ascending) inside each voxel.
__global__ void k_pack_spot_value(const int64_t* indptr, const float* val, int64_t n_rows, uint64_t* packed) {
    const int     lane  = threadIdx.x & 63;
    const int64_t wave  = ((int64_t) blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t waves = ((int64_t) gridDim.x * blockDim.x) >> 6;
    for (int64_t r = wave; r < n_rows; r += waves)
        for (int64_t k = indptr[r] + lane; k < indptr[r + 1]; k += 64)
            packed[k] = ((uint64_t) (uint32_t) r << 32) | (uint64_t) __float_as_uint(val[k]);
}
__global__ void k_unpack_spot_value(const uint64_t* packed, int64_t nnz, int32_t* tcol, float* tval) {
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += (int64_t) gridDim.x * blockDim.x) {
        const uint64_t p = packed[i];
        tcol[i]          = (int32_t) (uint32_t) (p >> 32);
        tval[i]          = __uint_as_float((uint32_t) p);
    }
}

int
csr_transpose_sorted(const int64_t* indptr, const int32_t* col, const float* val, int64_t n_rows, int64_t nnz,
                     int key_bits, int32_t* keys_tmp, uint64_t* packed, uint64_t* packed_tmp, int32_t* tcol,
                     float* tval, void* scratch, size_t* scratch_bytes, hipStream_t s) {
    if (!scratch) {
        return (int) hipcub::DeviceRadixSort::SortPairs(nullptr, *scratch_bytes, (const uint32_t*) col,
                                                        (uint32_t*) keys_tmp, packed, packed_tmp, (size_t) nnz, 0,
                                                        key_bits, s);
    }
    if (nnz <= 0) return 0;
    hipLaunchKernelGGL(k_pack_spot_value, dim3(4096), dim3(256), 0, s, indptr, val, n_rows, packed);
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(scratch, *scratch_bytes, (const uint32_t*) col,
                                                      (uint32_t*) keys_tmp, (const uint64_t*) packed, packed_tmp,
                                                      (size_t) nnz, 0, key_bits, s);
    if (e != hipSuccess) return (int) e;
    hipLaunchKernelGGL(k_unpack_spot_value, dim3(grid_for((uint64_t) nnz)), dim3(256), 0, s, packed_tmp, nnz, tcol,
                       tval);
    return (int) hipGetLastError();
}

int
csr_transpose_count(const int32_t* col, int64_t nnz, uint32_t* cnt, hipStream_t s) {
    if (nnz <= 0) return 0;
    hipLaunchKernelGGL(k_count_cols, dim3(grid_for((uint64_t) nnz)), dim3(256), 0, s, col, nnz, cnt);
    return (int) hipGetLastError();
}

}   // namespace mqi_aux

namespace mqi_aux
{
// out[s][i] += pair[stride i + s] for s < ns over n elements (stride 2 or 4);
// the pair array is cleared where it held a sum
struct fold_outs {
    double* o[4];
};
__global__ void k_pair_fold(fold_outs out, int ns, int stride, double* __restrict__ pair, uint64_t n) {
    for (uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t) gridDim.x * blockDim.x) {
        double* e = pair + (uint64_t) stride * i;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            if (s >= ns) break;
            const double x = e[s];
            if (x != 0.0) {
                out.o[s][i] += x;
                e[s] = 0.0;
            }
        }
    }
}
int
pair_fold(double* const* out, int ns, int stride, double* pair, uint64_t n, hipStream_t s) {
    if (n == 0) return 0;
    if (ns < 1 || ns > 4 || ns > stride) return (int) hipErrorInvalidValue;
    fold_outs o{};
    for (int k = 0; k < ns; ++k) o.o[k] = out[k];
    hipLaunchKernelGGL(k_pair_fold, dim3(grid_for(n)), dim3(256), 0, s, o, ns, stride, pair, n);
    return (int) hipGetLastError();
}
}   // namespace mqi_aux
