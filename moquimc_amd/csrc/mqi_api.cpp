// Host runtime behind the C ABI of include/mqi_hip.h.
//
// A context holds the device copies of one voxel child (edges, density), the
// spot table, the scenario table, the scorer arrays (FP64, dense, one slot
// per voxel instead of the reference's key_value tables), the per-lane
// secondary stacks, the Dij hash tables and the CSR matrices built from them.
// All device work is enqueued on the context's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/mqi_hip.h"
#include "mqi_aux.h"
#include "mqi_kernels.h"
#include "mqi_limits.h"
#include "mqi_plan.h"
#include "../host/mqi_hip_io.hpp"

namespace
{

// CSR-Adaptive segmentation of one CSR matrix (see mqi_aux::csr_adaptive)
struct csr_plan {
    int64_t   n_seg = 0, n_split = 0;
    int64_t*  seg_k     = nullptr;   // n_seg + 1
    uint32_t* seg_info  = nullptr;   // n_seg + 1
    int32_t*  split_seg = nullptr;   // 2 per split row
    double*   part      = nullptr;   // n_seg
    double*   qpart     = nullptr;   // n_seg + n_split + MQI_QUAD_EMPTY_BLOCKS + MQI_QUAD_SUM_BLOCKS: objective
                                     // terms and chunk sums (on first use)
};

struct csr_dev {
    int64_t  rows = 0, cols = 0, nnz = 0;
    int64_t* indptr  = nullptr;   // rows + 1
    int32_t* col     = nullptr;
    float*   val     = nullptr;
    int64_t* tindptr = nullptr;   // transposed (voxel-major) copy without its empty rows: n_nz + 1
    uint32_t* tnonempty = nullptr;   // bit v: voxel v has a nonzero ((cols + 31) / 32 words)
    int32_t* tcol    = nullptr;
    float*   tval    = nullptr;
    csr_plan plan, tplan;         // segmentations of the spot-major and voxel-major forms
    // D^T g gathers x only at the voxels that appear in some row: their list,
    // the rows' columns renumbered into it, and the compacted x
    int64_t  n_nz  = 0;
    int32_t* nzvox = nullptr;
    int32_t* ccol  = nullptr;
    float*   xc    = nullptr;
    // D^T g over byte deltas (with escapes) of the compacted columns, for
    // matrices whose rows are long (mqi_aux::pieces_staged); ccol is dropped
    // once they are built
    int64_t   n_piece   = 0;
    int32_t*  row_piece = nullptr;   // rows + 1: each row's first piece
    double*   ppart     = nullptr;   // n_piece: sums of the pieces
    float*    sval      = nullptr;   // staged values (lane-major batches, zero padding)
    uint8_t*  sd8       = nullptr;   // staged deltas, one byte each (255: escape)
    uint16_t* pesc      = nullptr;   // the escapes of all pieces (+ 64 padding entries)
    // n_piece headers in the order the product runs the pieces: first batch,
    // batches, the column the first delta adds to, piece index
    mqi_aux::piece_hdr* phdr = nullptr;
    // D w over 16-bit spot columns (at most 65536 spots; tcol is then dropped)
    uint16_t* tcol16 = nullptr;
};

}   // namespace

struct mqi_ctx {
    int         device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int         n_cu   = 0;
    size_t      memory = 0;

    // geometry
    int                nx = 0, ny = 0, nz = 0;
    float *            d_xe = nullptr, *d_ye = nullptr, *d_ze = nullptr, *d_density = nullptr;
    float              fwd[9], inv[9], trans[3];
    std::vector<float> h_xe, h_ye, h_ze;

    // spots and scenarios
    std::vector<mqi_spot> spots;
    mqi_spot*             d_spots = nullptr;
    int                   n_scen  = 1;
    // grouped scoring: the interleaved array of the leading scorers (2 or 4
    // doubles per element, zero between passes) and the arrays it folds into
    double*               d_pair     = nullptr;
    size_t                pair_cap   = 0;   // doubles
    size_t                pair_n     = 0;   // elements of the current pass (0: not grouped)
    int                   pair_ns = 0, pair_stride = 0;
    double*               pair_real[4] = {nullptr, nullptr, nullptr, nullptr};
    std::vector<double>   scen_budget;   // histories factor per scenario (empty: 1)
    bool                  scen_set = false;   // mqi_set_scenarios gave explicit shifts / scales
    float *               d_shift = nullptr, *d_scale = nullptr;

    // scorers
    int              mask = 0, per_scen = 0;
    std::vector<int> kinds;
    double*          d_out[MQI_MAX_SCORERS] = { nullptr };

    // counters: next, tracked, drops
    unsigned long long* d_counters = nullptr;   // [1] tracked, [2] dropped, [3] steps
    unsigned long long* d_queue    = nullptr;   // per-XCD work-range counters of dynamic launches
    // Dose derived from Dose_stat (mqi_run with each scorer once): the kernel
    // scores Dose_stat only; d_snap holds Dose_stat before the run
    bool    alias_dose = false;
    double* d_snap     = nullptr;
    size_t  snap_n     = 0;
    unsigned int*       d_status   = nullptr;
    uint64_t*           d_hoff     = nullptr;
    float*              d_statmax  = nullptr;
    double*             d_statsum  = nullptr;
    unsigned long long* d_statcnt  = nullptr;

    // per-lane secondary stacks
    float* d_stack     = nullptr;
    size_t stack_bytes = 0;

    // work list
    uint64_t* d_item_end  = nullptr;
    uint32_t* d_item_spot = nullptr;
    uint32_t* d_item_scen = nullptr;
    size_t    items_cap   = 0;

    // ROIs of the regular scorers ([0]) and the statistics pair ([1]):
    // CONTOUR runs (start, stride) when set, and their per-voxel expansion
    bool                  roi_set[2] = {false, false};
    std::vector<uint32_t> roi_start[2], roi_stride[2];
    uint8_t*              d_roi_flags = nullptr;

    // beamline children (device copies; see transport_params)
    int       bl_n       = 0;
    int       bl_n_edges = 0;
    int*      d_bl_meta  = nullptr;
    float*    d_bl_xf    = nullptr;
    float*    d_bl_edges = nullptr;
    float*    d_bl_dens  = nullptr;

    // Dij
    uint32_t*             d_dij_keys  = nullptr;
    double*               d_dij_vals  = nullptr;
    uint64_t*             d_dij_off   = nullptr;   // per-item slot offsets (n_items + 1)
    uint64_t              dij_alloc   = 0;         // slots allocated
    size_t                dij_items   = 0;
    std::vector<uint64_t> dij_off;                 // host copy of the offsets of the last run
    std::vector<csr_dev> dij;

    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev_mv0 = nullptr, ev_mv1 = nullptr;   // around the last sparse product
    // captured batched pass (cfg.use_graph): replayed while the launch it was
    // captured from (kernel arguments, variant, shape) stays the same
    hipGraphExec_t graph_exec = nullptr;
    std::vector<unsigned char> graph_key;
    bool       matvec_pending = false;

    // this device's copy of the no-allocation scratch array (see scratch below)
    unsigned char* scratch     = nullptr;
    std::string    scratch_err;
};

// ------------------------------------------------------------------ no-allocation scratch
// mqi_run_vertices_noalloc keeps its one-item work list, counters, work-range
// queue and per-lane secondary stacks in this array instead of buffers the
// library allocates.  Each GPU holds its own copy from module load to process
// end.  In host code the name stands for no device address: the per-device
// address comes from hipGetSymbolAddress (scratch_address) and only that
// address is passed to kernels and fills.
#define MQI_SCRATCH_HEADER 8192
static __device__ __align__(256) unsigned char scratch[MQI_SCRATCH_HEADER + MQI_NOALLOC_STACK_BYTES];

namespace
{

// byte offsets inside the header
constexpr size_t SCR_COUNTERS = 0;      // 4 x u64: -, tracked, dropped, steps
constexpr size_t SCR_STATUS   = 256;    // u32
constexpr size_t SCR_HOFF     = 512;    // u64
constexpr size_t SCR_ITEM_END = 768;    // u64
constexpr size_t SCR_ITEM_SPT = 1024;   // u32
constexpr size_t SCR_ITEM_SCN = 1280;   // u32
constexpr size_t SCR_QUEUE    = 4096;   // MQI_QUEUES * MQI_QUEUE_STRIDE u64
static_assert(SCR_QUEUE + sizeof(unsigned long long) * MQI_QUEUES * MQI_QUEUE_STRIDE <= MQI_SCRATCH_HEADER,
              "the work-range queue must fit the scratch header");

constexpr int MQI_SCRATCH_DEVICES = 64;
std::mutex    g_scratch_mu;
void*         g_scratch_addr[MQI_SCRATCH_DEVICES] = { nullptr };

// The address of `scratch` on `device` (which must be current), looked up
// once per device and kept.
hipError_t
scratch_address(int device, void** out) {
    *out = nullptr;
    if (device < 0 || device >= MQI_SCRATCH_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[device]) {
        void*      p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_addr[device] = p;
    }
    *out = g_scratch_addr[device];
    return hipSuccess;
}

}   // namespace

namespace
{

int
fail(mqi_ctx* c, int code, const char* fmt, ...) {
    char    buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPCHK(c, expr)                                                                                    \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess) return fail(c, MQI_E_HIP, "%s: %s", #expr, hipGetErrorString(e_));          \
    } while (0)

#define KCHK(c, expr)                                                                                      \
    do {                                                                                                   \
        int e_ = (expr);                                                                                   \
        if (e_ != 0) return fail(c, MQI_E_HIP, "%s: %s", #expr, hipGetErrorString((hipError_t) e_));      \
    } while (0)

template<typename T>
int
dalloc(mqi_ctx* c, T** p, size_t count) {
    if (*p) {
        hipFree(*p);
        *p = nullptr;
    }
    if (count == 0) return 0;
    hipError_t e = hipMalloc((void**) p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(c, MQI_E_MEMORY, "hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
    }
    return 0;
}

template<typename T>
void
dfree(T*& p) {
    if (p) hipFree(p);
    p = nullptr;
}

void
free_csr(csr_dev& m) {
    {
        dfree(m.indptr);
        dfree(m.col);
        dfree(m.val);
        dfree(m.tindptr);
        dfree(m.tnonempty);
        dfree(m.tcol);
        dfree(m.tval);
        dfree(m.nzvox);
        dfree(m.ccol);
        dfree(m.xc);
        dfree(m.sval);
        dfree(m.sd8);
        dfree(m.pesc);
        dfree(m.phdr);
        dfree(m.row_piece);
        dfree(m.ppart);
        dfree(m.tcol16);
        for (csr_plan* p : {&m.plan, &m.tplan}) {
            dfree(p->seg_k);
            dfree(p->seg_info);
            dfree(p->split_seg);
            dfree(p->part);
            dfree(p->qpart);
        }
    }
    m = csr_dev();
}

void
free_dij(mqi_ctx* c) {
    for (auto& m : c->dij) free_csr(m);
    c->dij.clear();
}

// CSR-Adaptive segmentation of a row pointer array: rows longer than
// MQI_SEG_SHORT are cut into MQI_SEG_LONG pieces, shorter rows are packed
// (MQI_SEG_SHORT entries, MQI_SEG_ROWS rows at most per segment).
int
build_plan(mqi_ctx* c, const std::vector<int64_t>& ip, csr_plan& P) {
    const int64_t         rows = (int64_t) ip.size() - 1;
    std::vector<int64_t>  sk;
    std::vector<uint32_t> si;
    std::vector<int32_t>  split;
    if (rows >= (1ll << 31)) return fail(c, MQI_E_ARG, "more than 2^31 - 1 rows");
    int64_t r = 0;
    while (r < rows) {
        const int64_t len = ip[r + 1] - ip[r];
        if (len > MQI_SEG_SHORT) {
            split.push_back((int32_t) sk.size());
            for (int64_t off = 0; off < len; off += MQI_SEG_LONG) {
                sk.push_back(ip[r] + off);
                si.push_back((uint32_t) r | 0x80000000u);
            }
            split.push_back((int32_t) sk.size());
            ++r;
        } else {
            const int64_t r0 = r, kb = ip[r];
            while (r < rows && r - r0 < MQI_SEG_ROWS && ip[r + 1] - kb <= MQI_SEG_SHORT) ++r;
            sk.push_back(kb);
            si.push_back((uint32_t) r0);
        }
        if (sk.size() >= (size_t) INT32_MAX) return fail(c, MQI_E_ARG, "too many CSR segments");
    }
    P.n_seg   = (int64_t) sk.size();
    P.n_split = (int64_t) split.size() / 2;
    sk.push_back(ip[rows]);
    si.push_back((uint32_t) rows);
    int e;
    if ((e = dalloc(c, &P.seg_k, sk.size())) || (e = dalloc(c, &P.seg_info, si.size())) ||
        (e = dalloc(c, &P.split_seg, std::max<size_t>(split.size(), 1))) ||
        (e = dalloc(c, &P.part, std::max<size_t>(P.n_seg, 1))))
        return e;
    HIPCHK(c, hipMemcpy(P.seg_k, sk.data(), sizeof(int64_t) * sk.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(P.seg_info, si.data(), sizeof(uint32_t) * si.size(), hipMemcpyHostToDevice));
    if (!split.empty())
        HIPCHK(c, hipMemcpy(P.split_seg, split.data(), sizeof(int32_t) * split.size(), hipMemcpyHostToDevice));
    return 0;
}

// The pieces of D^T g over column deltas (mqi_aux::pieces_staged): the
// 16-bit deltas of the compacted columns, and the positions where a row
// skips more than 65535 columns, come from the device; each row is cut there
// and every MQI_PIECE nonzeros (cut_pieces), then the pieces are restaged in
// whole lane-major batches with one byte per delta (the deltas of 255 and
// more in per-piece escape lists) and ordered by their first columns.
int
build_pieces(mqi_ctx* c, csr_dev& m, const std::vector<int64_t>& ip) {
    const int64_t       rows   = m.rows;
    unsigned long long* d_ngap = nullptr;
    int64_t *           d_gaps = nullptr, *d_pk = nullptr;
    uint16_t*           d_dcol = nullptr;
    int                 r      = 0;
    struct guard {
        unsigned long long*& n;
        int64_t *&           g, *&k;
        uint16_t*&           d;
        ~guard() {
            hipFree(n);
            hipFree(g);
            hipFree(k);
            hipFree(d);
        }
    } g{ d_ngap, d_gaps, d_pk, d_dcol };
    int64_t cap = 1 << 16;
    if ((r = dalloc(c, &d_dcol, m.nnz)) || (r = dalloc(c, &d_ngap, 1)) || (r = dalloc(c, &d_gaps, cap))) return r;
    unsigned long long n_gaps = 0;
    for (int pass = 0; pass < 2; ++pass) {
        KCHK(c, mqi_aux::deltas16(m.indptr, m.ccol, rows, m.nnz, d_dcol, d_ngap, d_gaps, cap, c->stream));
        HIPCHK(c, hipMemcpy(&n_gaps, d_ngap, sizeof n_gaps, hipMemcpyDeviceToHost));
        if ((int64_t) n_gaps <= cap) break;
        hipFree(d_gaps);
        d_gaps = nullptr;
        cap    = (int64_t) n_gaps;
        if ((r = dalloc(c, &d_gaps, cap))) return r;
    }
    std::vector<int64_t> gaps(n_gaps);
    if (n_gaps) HIPCHK(c, hipMemcpy(gaps.data(), d_gaps, sizeof(int64_t) * n_gaps, hipMemcpyDeviceToHost));
    std::sort(gaps.begin(), gaps.end());
    std::vector<int64_t> pk;
    std::vector<int32_t> row_piece;
    if (!cut_pieces(ip, gaps, MQI_PIECE, pk, row_piece)) return fail(c, MQI_E_ARG, "too many Dij pieces");
    m.n_piece = (int64_t) pk.size() - 1;
    std::vector<int64_t> pb(m.n_piece + 1, 0);
    for (int64_t i = 0; i < m.n_piece; ++i) pb[i + 1] = pb[i] + ((pk[i + 1] - pk[i] + 255) & ~(int64_t) 255);
    if ((pb.back() >> 8) > (int64_t) UINT32_MAX) return fail(c, MQI_E_CAPACITY, "too many Dij batches");
    int32_t*  d_base = nullptr;
    int64_t*  d_pb   = nullptr;
    uint32_t* d_ecnt = nullptr;
    struct guard2 {
        int32_t*&  b;
        int64_t*&  p;
        uint32_t*& e;
        ~guard2() {
            hipFree(b);
            hipFree(p);
            hipFree(e);
        }
    } g2{ d_base, d_pb, d_ecnt };
    const int64_t np1 = std::max<int64_t>(m.n_piece, 1);
    if ((r = dalloc(c, &d_pk, pk.size())) || (r = dalloc(c, &d_base, np1)) || (r = dalloc(c, &d_ecnt, np1)) ||
        (r = dalloc(c, &m.row_piece, rows + 1)) || (r = dalloc(c, &m.ppart, np1)) ||
        (r = dalloc(c, &d_pb, pb.size())) || (r = dalloc(c, &m.sval, std::max<int64_t>(pb.back(), 1))) ||
        (r = dalloc(c, &m.sd8, std::max<int64_t>(pb.back(), 1))) || (r = dalloc(c, &m.phdr, np1)))
        return r;
    HIPCHK(c, hipMemcpy(d_pk, pk.data(), sizeof(int64_t) * pk.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(m.row_piece, row_piece.data(), sizeof(int32_t) * (rows + 1), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(d_pb, pb.data(), sizeof(int64_t) * pb.size(), hipMemcpyHostToDevice));
    KCHK(c, mqi_aux::piece_base(m.ccol, d_dcol, d_pk, m.n_piece, d_base, c->stream));
    // escapes (deltas of 255 and more): counted per piece, their lists placed
    // by the exclusive prefix sum, padded by the 64 entries the product may
    // read past the last one
    std::vector<uint32_t> ecnt(m.n_piece), ebase(np1, 0);
    if (m.n_piece) {
        KCHK(c, mqi_aux::piece_escapes(d_dcol, d_pk, m.n_piece, d_ecnt, c->stream));
        HIPCHK(c, hipMemcpy(ecnt.data(), d_ecnt, sizeof(uint32_t) * m.n_piece, hipMemcpyDeviceToHost));
    }
    uint64_t n_esc = 0;
    for (int64_t i = 0; i < m.n_piece; ++i) {
        ebase[i] = (uint32_t) n_esc;
        n_esc += ecnt[i];
    }
    if (n_esc + 64 > (uint64_t) UINT32_MAX) return fail(c, MQI_E_CAPACITY, "too many Dij escapes");
    const uint64_t n_esc_alloc = (n_esc + 64 + 1) & ~(uint64_t) 1;
    if ((r = dalloc(c, &m.pesc, n_esc_alloc))) return r;
    KCHK(c, mqi_aux::fill_u32(m.pesc, 0u, n_esc_alloc / 2, c->stream));
    if (m.n_piece) HIPCHK(c, hipMemcpy(d_ecnt, ebase.data(), sizeof(uint32_t) * m.n_piece, hipMemcpyHostToDevice));
    KCHK(c, mqi_aux::stage_pieces(m.val, d_dcol, d_pk, d_pb, d_ecnt, m.n_piece, m.sval, m.sd8, m.pesc, c->stream));
    // processed in the order of their first columns: the pieces that run at
    // the same time gather from the same stretch of x (7 % faster on the
    // 5000-spot Dij than row order); one 32-byte header per piece in that
    // order, so a wave starts a piece after one round trip
    std::vector<int32_t> first(m.n_piece), order(m.n_piece);
    if (m.n_piece) HIPCHK(c, hipMemcpy(first.data(), d_base, sizeof(int32_t) * m.n_piece, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < m.n_piece; ++i) order[i] = (int32_t) i;
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return first[a] < first[b]; });
    std::vector<mqi_aux::piece_hdr> hdr(m.n_piece);
    for (int64_t i = 0; i < m.n_piece; ++i) {
        const int32_t q = order[i];
        hdr[i].batch0   = (uint32_t) (pb[q] >> 8);
        hdr[i].batches  = (uint32_t) ((pb[q + 1] - pb[q]) >> 8);
        hdr[i].base     = first[q];
        hdr[i].piece    = q;
        hdr[i].esc0     = ebase[q];
        hdr[i].pad[0] = hdr[i].pad[1] = hdr[i].pad[2] = 0;
    }
    if (m.n_piece)
        HIPCHK(c, hipMemcpy(m.phdr, hdr.data(), sizeof(mqi_aux::piece_hdr) * m.n_piece, hipMemcpyHostToDevice));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

uint64_t
nvox(const mqi_ctx* c) {
    return (uint64_t) c->nx * c->ny * c->nz;
}

int
check_edges(mqi_ctx* c, const float* e, int n, const char* axis) {
    if (n < 1) return fail(c, MQI_E_ARG, "%s: need at least one voxel", axis);
    for (int i = 0; i < n; ++i) {
        if (!(e[i + 1] > e[i])) return fail(c, MQI_E_ARG, "%s edges must ascend (index %d)", axis, i);
        // The device index search tests only the three edges around the
        // position; that reproduces the reference's linear scan when every
        // voxel is wider than twice the 1e-3 mm tolerance.
        if (!((double) e[i + 1] - (double) e[i] > 2.5e-3))
            return fail(c, MQI_E_ARG, "%s voxel %d narrower than 2.5e-3 mm", axis, i);
    }
    return 0;
}

int
upload_grid_common(mqi_ctx* c, int nx, int ny, int nz, const float* xe, const float* ye, const float* ze,
                   const float* rot, const float* tr) {
    if (!xe || !ye || !ze) return fail(c, MQI_E_ARG, "edges are NULL");
    int r;
    if ((r = check_edges(c, xe, nx, "x")) || (r = check_edges(c, ye, ny, "y")) || (r = check_edges(c, ze, nz, "z")))
        return r;
    if ((uint64_t) nx * ny * nz >= (1ull << 31)) return fail(c, MQI_E_ARG, "more than 2^31 voxels");
    c->nx = nx;
    c->ny = ny;
    c->nz = nz;
    c->roi_set[0] = c->roi_set[1] = false;   // masks are per grid
    c->h_xe.assign(xe, xe + nx + 1);
    c->h_ye.assign(ye, ye + ny + 1);
    c->h_ze.assign(ze, ze + nz + 1);
    if ((r = dalloc(c, &c->d_xe, nx + 1)) || (r = dalloc(c, &c->d_ye, ny + 1)) || (r = dalloc(c, &c->d_ze, nz + 1)))
        return r;
    HIPCHK(c, hipMemcpyAsync(c->d_xe, xe, sizeof(float) * (nx + 1), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_ye, ye, sizeof(float) * (ny + 1), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_ze, ze, sizeof(float) * (nz + 1), hipMemcpyHostToDevice, c->stream));
    const float I[9] = { 1, 0, 0, 0, 1, 0, 0, 0, 1 };
    for (int k = 0; k < 9; ++k) c->fwd[k] = rot ? rot[k] : I[k];
    // grid3d::rotation_matrix_inv = fwd.inverse() (the transpose, mqi_matrix.hpp:290-293)
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) c->inv[3 * i + j] = c->fwd[3 * j + i];
    for (int k = 0; k < 3; ++k) c->trans[k] = tr ? tr[k] : 0.0f;
    return dalloc(c, &c->d_density, nvox(c));
}

size_t
shmem_bytes(const mqi_ctx* c, int threads) {
    return mqi_shmem_bytes(c->nx, c->ny, c->nz, c->bl_n_edges, threads);
}

int
ensure_items(mqi_ctx* c, size_t n) {
    if (n <= c->items_cap) return 0;
    int r;
    if ((r = dalloc(c, &c->d_item_end, n)) || (r = dalloc(c, &c->d_item_spot, n)) || (r = dalloc(c, &c->d_item_scen, n)))
        return r;
    c->items_cap = n;
    return 0;
}

int
ensure_stack(mqi_ctx* c, size_t bytes) {
    if (bytes <= c->stack_bytes) return 0;
    dfree(c->d_stack);
    int r = dalloc(c, (char**) &c->d_stack, bytes);
    if (r) return r;
    c->stack_bytes = bytes;
    return 0;
}

struct run_plan {
    std::vector<uint64_t> item_end;
    std::vector<uint32_t> item_spot, item_scen;
    uint64_t              total    = 0;
    uint64_t              max_item = 0;
};

// items are (scenario, spot), scenario-major; empty spots are skipped
run_plan
make_plan(const mqi_ctx* c, int only_spot) {
    run_plan p;
    for (int s = 0; s < c->n_scen; ++s)
        for (size_t j = 0; j < c->spots.size(); ++j) {
            if (only_spot >= 0 && (int) j != only_spot) continue;
            uint64_t h = c->spots[j].histories;
            if (!c->scen_budget.empty()) h = (uint64_t) std::llround((double) h * c->scen_budget[s]);
            if (h == 0) continue;
            p.total += h;
            p.max_item = std::max(p.max_item, h);
            p.item_end.push_back(p.total);
            p.item_spot.push_back((uint32_t) j);
            p.item_scen.push_back((uint32_t) s);
        }
    return p;
}

void
fill_params(mqi_ctx* c, transport_params& P, const mqi_run_config& cfg) {
    std::memset(&P, 0, sizeof P);
    P.nx      = c->nx;
    P.ny      = c->ny;
    P.nz      = c->nz;
    P.xe      = c->d_xe;
    P.ye      = c->d_ye;
    P.ze      = c->d_ze;
    P.density = c->d_density;
    std::memcpy(P.fwd, c->fwd, sizeof P.fwd);
    std::memcpy(P.inv, c->inv, sizeof P.inv);
    std::memcpy(P.trans, c->trans, sizeof P.trans);
    P.scen_shift    = c->scen_set ? c->d_shift : nullptr;
    P.scen_scale    = c->scen_set ? c->d_scale : nullptr;
    P.spots         = c->d_spots;
    P.n_spots       = (int) c->spots.size();
    P.item_end      = c->d_item_end;
    P.item_spot     = c->d_item_spot;
    P.item_scenario = c->d_item_scen;
    P.queue         = c->d_queue;
    P.fetch_chunk   = cfg.fetch_chunk > 0 ? cfg.fetch_chunk : 1;
    // 0: 4 (1 refills at once), the CT plan's optimum with paired scoring;
    // whole-wave refills (64) cluster the atomics of unpaired passes
    P.refill_lanes  = cfg.refill_lanes > 0 ? cfg.refill_lanes : 4;
    P.tracked       = c->d_counters + 1;
    P.drops         = c->d_counters + 2;
    P.steps         = c->d_counters + 3;
    P.status        = c->d_status;
    P.hist_offset   = c->d_hoff;
    P.seed          = cfg.seed;
    P.roi_flags     = (c->roi_set[0] || c->roi_set[1]) ? c->d_roi_flags : nullptr;
    P.n_bl          = c->bl_n;
    P.bl_meta       = c->d_bl_meta;
    P.bl_xf         = c->d_bl_xf;
    P.bl_edges      = c->d_bl_edges;
    P.bl_n_edges    = c->bl_n_edges;
    P.bl_density    = c->d_bl_dens;
    // alias_dose: the Dose scorer is derived from Dose_stat after the run
    // (the kernel then does one atomic less per step)
    P.n_scorers = 0;
    for (size_t k = 0; k < c->kinds.size(); ++k) {
        if (c->alias_dose && k == 0) continue;
        P.kinds[P.n_scorers] = c->kinds[k];
        P.out[P.n_scorers]   = c->d_out[k];
        P.n_scorers++;
    }
    if (cfg.dij) {
        P.kinds[P.n_scorers] = MQI_KIND_DIJ;
        P.out[P.n_scorers]   = nullptr;
        P.n_scorers++;
        P.dij_keys     = c->d_dij_keys;
        P.dij_vals     = c->d_dij_vals;
        P.dij_off      = c->d_dij_off;
    }
    P.out_scenario_stride = c->per_scen ? nvox(c) : 0;
    P.loops               = cfg.scoring_loops;
    P.stack               = c->d_stack;
    P.stack_cap           = cfg.stack_depth > 0 ? cfg.stack_depth : (cfg.physics_debug ? 200 : 10);
}

// Every device array a launch reads must be live: the work list and the
// stack can be reallocated between fill_params and the launch, so the
// pointers are bound here, and the launch refuses NULL ones.
void
bind_buffers(mqi_ctx* c, transport_params& P) {
    P.item_end      = c->d_item_end;
    P.item_spot     = c->d_item_spot;
    P.item_scenario = c->d_item_scen;
    P.stack         = c->d_stack;
}

// A launch over the buffers P already names (work list, queue, counters,
// stacks); stack_bytes is the size of the array P.stack points into.
int
launch_bound(mqi_ctx* c, transport_params& P, const mqi_run_config& cfg, int dynamic, int trace, int blocks,
             int threads, size_t stack_bytes) {
    if (P.fetch_chunk < 1) return (int) hipErrorInvalidValue;
    if (!P.density || !P.xe || !P.ye || !P.ze || !P.stack || !P.queue || !P.tracked || !P.drops || !P.steps || !P.status ||
        !P.hist_offset || (!P.vertices && (!P.spots || !P.item_end || !P.item_spot || !P.item_scenario)))
        return (int) hipErrorInvalidValue;
    for (int s = 0; s < P.n_scorers; ++s)
        if (P.kinds[s] == MQI_KIND_DIJ ? (!P.dij_keys || !P.dij_vals) : !P.out[s]) return (int) hipErrorInvalidValue;
    if (trace && (!P.trace_cnb || !P.trace_rec || !P.trace_n)) return (int) hipErrorInvalidValue;
    if ((uint64_t) blocks * threads * (uint64_t) P.stack_cap * MQI_STACK_FLOATS * sizeof(float) > stack_bytes)
        return (int) hipErrorInvalidValue;
    const size_t sh = shmem_bytes(c, threads);
    if (dynamic) {
        const int e = mqi_aux::queue_init(P.queue, P.total, c->stream);
        if (e) return e;
    }
    if (cfg.exact)
        return mqi_exact::launch_transport(P, cfg.physics_debug, dynamic, trace, 0, blocks, threads, sh, c->stream);
    return mqi_fast::launch_transport(P, cfg.physics_debug, dynamic, trace, cfg.waves_per_simd, blocks, threads, sh,
                                      c->stream);
}

// Grouped scoring for a batched pass of the production kernels with each
// scorer scored once, no Dij and at least two scorers: the leading deferred
// scorers (two: pairs of lanes; three or four: groups of four lanes) add
// into one interleaved array, all sums of a voxel in one line and added by
// one atomic instruction from the group, folded into their arrays after the
// pass.
int
pair_setup(mqi_ctx* c, transport_params& P, const mqi_run_config& cfg) {
    c->pair_n = 0;
    if (cfg.exact || cfg.dij || cfg.scoring_loops != MQI_LOOPS_ONCE) return 0;
    int ns = 0;
    while (ns < P.n_scorers && ns < MQI_DEFER && P.kinds[ns] != MQI_KIND_DIJ) ++ns;
    if (ns < 2) return 0;
    const int    stride = ns == 2 ? 2 : 4;
    const size_t n      = (size_t) nvox(c) * (c->per_scen ? (size_t) c->n_scen : 1u);
    if (stride * n > c->pair_cap) {
        dfree(c->d_pair);
        c->pair_cap = 0;
        int r = dalloc(c, &c->d_pair, stride * n);
        if (r) return r;
        KCHK(c, mqi_aux::fill_u64(c->d_pair, 0ull, stride * n, c->stream));
        c->pair_cap = stride * n;
    }
    for (int s = 0; s < ns; ++s) c->pair_real[s] = P.out[s];
    c->pair_n      = n;
    c->pair_ns     = ns;
    c->pair_stride = stride;
    P.pair         = stride;
    P.group_ns     = ns;
    P.pair_out     = c->d_pair;
    return 0;
}

int
pair_fold(mqi_ctx* c) {
    if (c->pair_n)
        KCHK(c, mqi_aux::pair_fold(c->pair_real, c->pair_ns, c->pair_stride, c->d_pair, c->pair_n, c->stream));
    return 0;
}

int
launch(mqi_ctx* c, transport_params& P, const mqi_run_config& cfg, int dynamic, int trace, int blocks,
       int threads) {
    bind_buffers(c, P);
    return launch_bound(c, P, cfg, dynamic, trace, blocks, threads, c->stack_bytes);
}

int
check_tuning(mqi_ctx* c, const mqi_run_config& cfg) {
    const int w = cfg.waves_per_simd;
    if (w != 0 && (w < 2 || w > 6) && w != 8) return fail(c, MQI_E_ARG, "waves_per_simd must be 0, 2-6 or 8");
    if (cfg.fetch_chunk < 0 || cfg.fetch_chunk > 4096) return fail(c, MQI_E_ARG, "fetch_chunk must be in [0, 4096]");
    if (cfg.refill_lanes < 0 || cfg.refill_lanes > 64) return fail(c, MQI_E_ARG, "refill_lanes must be in [0, 64]");
    if (cfg.block_size != 0 && (cfg.block_size % 64 != 0 || cfg.block_size > 256 || cfg.block_size < 0))
        return fail(c, MQI_E_ARG, "block_size must be 64, 128, 192 or 256");
    if (cfg.blocks < 0) return fail(c, MQI_E_ARG, "blocks must be >= 0");
    if (cfg.stack_depth < 0 || cfg.stack_depth > 4096) return fail(c, MQI_E_ARG, "stack_depth must be in [0, 4096]");
    if (cfg.dij_slots_sqrt < 0 || cfg.dij_slots_sqrt > (1 << 20))
        return fail(c, MQI_E_ARG, "dij_slots_sqrt must be in [0, 2^20]");
    if (cfg.dij_capacity > (1ull << 32)) return fail(c, MQI_E_ARG, "dij_capacity must be at most 2^32");
    return 0;
}

int
persistent_blocks(mqi_ctx* c, const mqi_run_config& cfg, int threads, int trace) {
    if (cfg.blocks > 0) return cfg.blocks;
    const size_t sh  = shmem_bytes(c, threads);
    int          occ = cfg.exact ? mqi_exact::occupancy(cfg.physics_debug, 1, trace, 0, threads, sh)
                                 : mqi_fast::occupancy(cfg.physics_debug, 1, trace, cfg.waves_per_simd, threads, sh);
    return c->n_cu * occ;
}

int
upload_plan(mqi_ctx* c, const run_plan& p) {
    int r = ensure_items(c, std::max<size_t>(1, p.item_end.size()));
    if (r) return r;
    if (p.item_end.empty()) return 0;
    HIPCHK(c, hipMemcpyAsync(c->d_item_end, p.item_end.data(), sizeof(uint64_t) * p.item_end.size(),
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_item_spot, p.item_spot.data(), sizeof(uint32_t) * p.item_spot.size(),
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_item_scen, p.item_scen.data(), sizeof(uint32_t) * p.item_scen.size(),
                             hipMemcpyHostToDevice, c->stream));
    return 0;
}

// The batched pass as a captured graph (work-range initialisation + the
// persistent transport launch), instantiated once and replayed while the
// kernel arguments, variant and launch shape stay the same.  The history
// offset lives in device memory (set before each replay), so replays draw
// new histories; nothing is allocated and no memset node is recorded.
int
graph_pass(mqi_ctx* c, transport_params& P, const mqi_run_config& cfg, int blocks, int threads, double* ms) {
    bind_buffers(c, P);
    const int variant[6] = {cfg.physics_debug, cfg.exact, cfg.exact ? 0 : cfg.waves_per_simd, blocks, threads,
                            (int) shmem_bytes(c, threads)};
    std::vector<unsigned char> key(sizeof P + sizeof variant);
    std::memcpy(key.data(), &P, sizeof P);
    std::memcpy(key.data() + sizeof P, variant, sizeof variant);
    if (!c->graph_exec || key != c->graph_key) {
        if (c->graph_exec) {
            hipGraphExecDestroy(c->graph_exec);
            c->graph_exec = nullptr;
        }
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int  e = launch(c, P, cfg, 1, 0, blocks, threads);
        hipGraph_t g = nullptr;
        hipError_t ee = hipStreamEndCapture(c->stream, &g);
        if (e || ee != hipSuccess) {
            if (g) hipGraphDestroy(g);
            return fail(c, MQI_E_HIP, "graph capture failed: %s", hipGetErrorString(e ? (hipError_t) e : ee));
        }
        hipError_t ie = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (ie != hipSuccess) {
            c->graph_exec = nullptr;
            return fail(c, MQI_E_HIP, "graph instantiation failed: %s", hipGetErrorString(ie));
        }
        c->graph_key = key;
    }
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHK(c, hipGraphLaunch(c->graph_exec, c->stream));
    if (int e = pair_fold(c)) return e;
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float t = 0;
    HIPCHK(c, hipEventElapsedTime(&t, c->ev0, c->ev1));
    *ms += t;
    return 0;
}

// One pass over a plan: the persistent dynamic launch, or (plain) one static
// launch per spot of the plan.  Adds the device time of the transport
// kernels to *ms.
int
transport_pass(mqi_ctx* c, const mqi_run_config& cfg, uint64_t hoff, double* ms) {
    const int threads = cfg.block_size > 0 ? cfg.block_size : 256;
    if (threads % 64 != 0 || threads > 256) return fail(c, MQI_E_ARG, "block_size must be 64, 128, 192 or 256");
    transport_params P;
    std::memset(&P, 0, sizeof P);   // defined padding: the struct is compared byte-wise for graph reuse
    fill_params(c, P, cfg);
    KCHK(c, mqi_aux::fill_u64(c->d_hoff, hoff, 1, c->stream));
    const size_t lane_bytes = (size_t) P.stack_cap * MQI_STACK_FLOATS * sizeof(float);
    int          r;
    if (cfg.batched) {
        run_plan p = make_plan(c, -1);
        if ((r = upload_plan(c, p))) return r;
        if (p.total == 0) return 0;
        P.n_items      = (int) p.item_end.size();
        P.total        = p.total;
        const int blocks = persistent_blocks(c, cfg, threads, 0);
        if ((r = ensure_stack(c, lane_bytes * (size_t) blocks * threads))) return r;
        P.stack = c->d_stack;
        if ((r = pair_setup(c, P, cfg))) return r;
        if (cfg.use_graph) return graph_pass(c, P, cfg, blocks, threads, ms);
        HIPCHK(c, hipEventRecord(c->ev0, c->stream));
        KCHK(c, launch(c, P, cfg, 1, 0, blocks, threads));
        if ((r = pair_fold(c))) return r;
        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
        HIPCHK(c, hipEventSynchronize(c->ev1));
        float t = 0;
        HIPCHK(c, hipEventElapsedTime(&t, c->ev0, c->ev1));
        *ms += t;
        return 0;
    }
    // plain: per scenario, per spot
    const uint64_t max_lanes = 1ull << 18;
    const int      max_blocks = (int) ((max_lanes + threads - 1) / threads);
    if ((r = ensure_stack(c, lane_bytes * (size_t) max_blocks * threads))) return r;
    P.stack = c->d_stack;
    for (size_t j = 0; j < c->spots.size(); ++j) {
        run_plan p = make_plan(c, (int) j);
        if (p.total == 0) continue;
        if ((r = upload_plan(c, p))) return r;
        P.n_items    = (int) p.item_end.size();
        P.total      = p.total;
        uint64_t lanes = std::min<uint64_t>(p.total, max_lanes);
        int      blocks = (int) ((lanes + threads - 1) / threads);   // <= max_blocks
        HIPCHK(c, hipEventRecord(c->ev0, c->stream));
        KCHK(c, launch(c, P, cfg, 0, 0, blocks, threads));
        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
        HIPCHK(c, hipEventSynchronize(c->ev1));
        float t = 0;
        HIPCHK(c, hipEventElapsedTime(&t, c->ev0, c->ev1));
        *ms += t;
    }
    return 0;
}

int
stat_index(const mqi_ctx* c, int* s1, int* s2) {
    *s1 = *s2 = -1;
    for (size_t k = 0; k < c->kinds.size(); ++k) {
        if (c->kinds[k] == MQI_KIND_STAT_DOSE) *s1 = (int) k;
        if (c->kinds[k] == MQI_KIND_STAT_DOSE2) *s2 = (int) k;
    }
    return (*s1 >= 0 && *s2 >= 0) ? 0 : -1;
}

int
stat_value(mqi_ctx* c, int scenario, uint64_t n, float thr, double* out) {
    int s1, s2;
    if (stat_index(c, &s1, &s2)) return fail(c, MQI_E_STATE, "statistics need the MQI_SCORE_STAT scorers");
    if (n < 2) return fail(c, MQI_E_ARG, "statistics need at least two histories");
    const uint64_t off = c->per_scen ? (uint64_t) scenario * nvox(c) : 0;
    KCHK(c, mqi_aux::stat_reduce(c->d_out[s1] + off, c->d_out[s2] + off, nvox(c), (double) n, thr, c->d_statmax,
                                 c->d_statsum, c->d_statcnt, c->stream));
    double             sum = 0;
    unsigned long long cnt = 0;
    HIPCHK(c, hipMemcpyAsync(&sum, c->d_statsum, sizeof sum, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(&cnt, c->d_statcnt, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *out = cnt ? sum / (double) cnt : 0.0;
    return 0;
}

int
ready(mqi_ctx* c) {
    if (!c) return MQI_E_ARG;
    if (!c->d_density) return fail(c, MQI_E_STATE, "no grid: call mqi_set_grid first");
    return 0;
}

int
alloc_dij(mqi_ctx* c, const mqi_run_config& cfg) {
    // items are indexed scenario * n_spots + spot (every spot, including
    // empty ones); each owns a contiguous open-addressing table
    const size_t n_spots = c->spots.size();
    const size_t items   = (size_t) c->n_scen * n_spots;
    std::vector<uint64_t> off(items + 1, 0);
    for (int s = 0; s < c->n_scen; ++s)
        for (size_t j = 0; j < n_spots; ++j) {
            uint64_t cap;
            if (cfg.dij_slots_sqrt > 0) {
                uint64_t h = c->spots[j].histories;
                if (!c->scen_budget.empty()) h = (uint64_t) std::llround((double) h * c->scen_budget[s]);
                const uint64_t hi = cfg.dij_capacity ? cfg.dij_capacity : (1ull << 24);
                cap = (uint64_t) std::ceil((double) cfg.dij_slots_sqrt * std::sqrt((double) h));
                cap = std::min(hi, std::max<uint64_t>(cap, 1024));
            } else {
                cap = cfg.dij_capacity ? cfg.dij_capacity : 131072;
            }
            const size_t i = (size_t) s * n_spots + j;
            off[i + 1]     = off[i] + cap;
        }
    const uint64_t total = off[items];
    int            r;
    if (total > c->dij_alloc || !c->d_dij_keys) {
        if ((r = dalloc(c, &c->d_dij_keys, total)) || (r = dalloc(c, &c->d_dij_vals, total))) {
            c->dij_alloc = 0;
            return r;
        }
        c->dij_alloc = total;
    }
    if (items + 1 > c->dij_off.size() || !c->d_dij_off)
        if ((r = dalloc(c, &c->d_dij_off, items + 1))) return r;
    HIPCHK(c, hipMemcpy(c->d_dij_off, off.data(), sizeof(uint64_t) * (items + 1), hipMemcpyHostToDevice));
    c->dij_off   = std::move(off);
    c->dij_items = items;
    KCHK(c, mqi_aux::fill_u32(c->d_dij_keys, 0xffffffffu, total, c->stream));
    KCHK(c, mqi_aux::fill_u64(c->d_dij_vals, 0ull, total, c->stream));
    return 0;
}

// message of the last failed mqi_create on this thread (there is no context
// to hold it)
thread_local std::string g_create_error = "no error";

int
create_failed(int code, const std::string& msg) {
    g_create_error = msg;
    return code;
}

// The per-voxel ROI bytes the kernels read: bit 0 where the regular scorers
// score (inside a run of the scorers' ROI, or v > 0 under DIRECT), bit 1
// where the statistics pair does (inside a run of its ROI, or everywhere).
// One byte load per scoring step replaces roi_t's binary search over runs.
int
roi_upload(mqi_ctx* c) {
    const uint64_t       n = nvox(c);
    std::vector<uint8_t> fl(n, 0);
    if (!c->roi_set[0])
        for (uint64_t v = 1; v < n; ++v) fl[v] = 1;
    if (!c->roi_set[1])
        for (uint64_t v = 0; v < n; ++v) fl[v] |= 2;
    for (int w = 0; w < 2; ++w) {
        if (!c->roi_set[w]) continue;
        for (size_t r = 0; r < c->roi_start[w].size(); ++r)
            for (uint64_t v = c->roi_start[w][r]; v < (uint64_t) c->roi_start[w][r] + c->roi_stride[w][r]; ++v)
                fl[v] |= (uint8_t) (1u << w);
    }
    hipSetDevice(c->device);
    int r = dalloc(c, &c->d_roi_flags, n);
    if (r) return r;
    HIPCHK(c, hipMemcpy(c->d_roi_flags, fl.data(), n, hipMemcpyHostToDevice));
    return 0;
}

}   // namespace

extern "C" {

int
mqi_create(int device, mqi_ctx** out) {
    if (!out) return create_failed(MQI_E_ARG, "out is NULL");
    *out   = nullptr;
    int        nd = 0;
    hipError_t e  = hipGetDeviceCount(&nd);
    if (e != hipSuccess || nd <= 0)
        return create_failed(MQI_E_HIP, std::string("no HIP device: ") + (e != hipSuccess ? hipGetErrorString(e) : "0 devices"));
    if (device < 0 || device >= nd) return create_failed(MQI_E_ARG, "device index out of range");
    mqi_ctx* c = new (std::nothrow) mqi_ctx();
    if (!c) return create_failed(MQI_E_MEMORY, "out of host memory");
    c->device = device;
    if ((e = hipSetDevice(device)) != hipSuccess || (e = hipStreamCreate(&c->stream)) != hipSuccess) {
        delete c;
        return create_failed(MQI_E_HIP, std::string("device set-up failed: ") + hipGetErrorString(e));
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        c->n_cu   = prop.multiProcessorCount;
        c->memory = prop.totalGlobalMem;
    }
    if (c->n_cu <= 0) c->n_cu = 1;
    hipEventCreate(&c->ev0);
    hipEventCreate(&c->ev1);
    hipEventCreate(&c->ev_mv0);
    hipEventCreate(&c->ev_mv1);
    int r;
    if ((r = dalloc(c, &c->d_counters, 4)) || (r = dalloc(c, &c->d_queue, MQI_QUEUES * MQI_QUEUE_STRIDE)) || (r = dalloc(c, &c->d_status, 1)) || (r = dalloc(c, &c->d_hoff, 1)) ||
        (r = dalloc(c, &c->d_statmax, 1)) || (r = dalloc(c, &c->d_statsum, 1)) || (r = dalloc(c, &c->d_statcnt, 1))) {
        mqi_destroy(c);
        return r;
    }
    const float one = 1.0f, zero3[3] = { 0, 0, 0 };
    if ((r = dalloc(c, &c->d_shift, 3)) || (r = dalloc(c, &c->d_scale, 1))) {
        mqi_destroy(c);
        return r;
    }
    hipMemcpy(c->d_shift, zero3, sizeof zero3, hipMemcpyHostToDevice);
    hipMemcpy(c->d_scale, &one, sizeof one, hipMemcpyHostToDevice);
    // this device's no-allocation scratch (mqi_run_vertices_noalloc); a
    // failed lookup only disables that entry point
    void* sp = nullptr;
    if ((e = scratch_address(device, &sp)) == hipSuccess)
        c->scratch = (unsigned char*) sp;
    else
        c->scratch_err = hipGetErrorString(e);
    *out = c;
    return 0;
}

int
mqi_destroy(mqi_ctx* c) {
    if (!c) return MQI_E_ARG;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    dfree(c->d_xe);
    dfree(c->d_ye);
    dfree(c->d_ze);
    dfree(c->d_density);
    dfree(c->d_spots);
    dfree(c->d_shift);
    dfree(c->d_scale);
    for (auto& p : c->d_out) dfree(p);
    dfree(c->d_counters);
    dfree(c->d_queue);
    dfree(c->d_snap);
    dfree(c->d_status);
    dfree(c->d_hoff);
    dfree(c->d_statmax);
    dfree(c->d_statsum);
    dfree(c->d_statcnt);
    dfree(c->d_stack);
    dfree(c->d_pair);
    dfree(c->d_item_end);
    dfree(c->d_item_spot);
    dfree(c->d_item_scen);
    dfree(c->d_dij_keys);
    dfree(c->d_dij_vals);
    dfree(c->d_dij_off);
    dfree(c->d_roi_flags);
    dfree(c->d_bl_meta);
    dfree(c->d_bl_xf);
    dfree(c->d_bl_edges);
    dfree(c->d_bl_dens);
    free_dij(c);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->graph_exec) hipGraphExecDestroy(c->graph_exec);
    if (c->ev_mv0) hipEventDestroy(c->ev_mv0);
    if (c->ev_mv1) hipEventDestroy(c->ev_mv1);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

const char*
mqi_last_error(const mqi_ctx* c) {
    return c ? c->err.c_str() : g_create_error.c_str();
}

int
mqi_device_info(mqi_ctx* c, int* cus, size_t* mem) {
    if (!c) return MQI_E_ARG;
    if (cus) *cus = c->n_cu;
    if (mem) *mem = c->memory;
    return 0;
}

int
mqi_set_grid(mqi_ctx* c, int nx, int ny, int nz, const float* xe, const float* ye, const float* ze,
             const float* density, const float* rot, const float* tr) {
    if (!c || !density) return fail(c, MQI_E_ARG, "density is NULL");
    hipSetDevice(c->device);
    int r = upload_grid_common(c, nx, ny, nz, xe, ye, ze, rot, tr);
    if (r) return r;
    HIPCHK(c, hipMemcpyAsync(c->d_density, density, sizeof(float) * nvox(c), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->mask) return mqi_set_scorers(c, c->mask, c->per_scen);
    return 0;
}

int
mqi_set_grid_hu(mqi_ctx* c, int nx, int ny, int nz, const float* xe, const float* ye, const float* ze,
                const int16_t* hu, float density_scale, const float* rot, const float* tr) {
    if (!c || !hu) return fail(c, MQI_E_ARG, "hu is NULL");
    hipSetDevice(c->device);
    int r = upload_grid_common(c, nx, ny, nz, xe, ye, ze, rot, tr);
    if (r) return r;
    int16_t* d_hu = nullptr;
    if ((r = dalloc(c, &d_hu, nvox(c)))) return r;
    HIPCHK(c, hipMemcpyAsync(d_hu, hu, sizeof(int16_t) * nvox(c), hipMemcpyHostToDevice, c->stream));
    KCHK(c, mqi_aux::hu_to_density(d_hu, c->d_density, nvox(c), density_scale, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    hipFree(d_hu);
    if (c->mask) return mqi_set_scorers(c, c->mask, c->per_scen);
    return 0;
}

int
mqi_get_density(mqi_ctx* c, float* out) {
    int r = ready(c);
    if (r) return r;
    HIPCHK(c, hipMemcpy(out, c->d_density, sizeof(float) * nvox(c), hipMemcpyDeviceToHost));
    return 0;
}

int
mqi_set_spots(mqi_ctx* c, const mqi_spot* spots, int n) {
    if (!c || n < 0 || (n > 0 && !spots)) return fail(c, MQI_E_ARG, "bad spot table");
    hipSetDevice(c->device);
    for (int i = 0; i < n; ++i) {
        if (!(spots[i].energy > 0.5f) || !(spots[i].energy < 299.0f))
            return fail(c, MQI_E_ARG, "spot %d: energy %g MeV outside the tables (0.5, 299)", i, spots[i].energy);
        if (spots[i].histories >= (1ull << 32))
            return fail(c, MQI_E_ARG, "spot %d: more than 2^32 histories", i);
    }
    c->spots.assign(spots, spots + n);
    int r = dalloc(c, &c->d_spots, std::max(1, n));
    if (r) return r;
    if (n) HIPCHK(c, hipMemcpy(c->d_spots, spots, sizeof(mqi_spot) * n, hipMemcpyHostToDevice));
    return 0;
}

int
mqi_set_scenarios(mqi_ctx* c, int n, const float* shift, const float* scale, const float* budget) {
    if (!c || n < 0) return MQI_E_ARG;
    hipSetDevice(c->device);
    c->scen_budget.clear();
    if (n == 0) {
        c->n_scen   = 1;
        c->scen_set = false;
    } else {
        if (!shift || !scale) return fail(c, MQI_E_ARG, "scenario arrays are NULL");
        if (budget) {
            for (int k = 0; k < n; ++k)
                if (!(budget[k] >= 0.0f) || budget[k] > 1e6f) return fail(c, MQI_E_ARG, "bad history budget");
            c->scen_budget.assign(budget, budget + n);
        }
        int r;
        if ((r = dalloc(c, &c->d_shift, 3 * (size_t) n)) || (r = dalloc(c, &c->d_scale, (size_t) n))) return r;
        HIPCHK(c, hipMemcpy(c->d_shift, shift, sizeof(float) * 3 * n, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->d_scale, scale, sizeof(float) * n, hipMemcpyHostToDevice));
        c->n_scen   = n;
        c->scen_set = true;
    }
    if (c->mask && c->per_scen) return mqi_set_scorers(c, c->mask, c->per_scen);
    return 0;
}

int
mqi_set_scorers(mqi_ctx* c, int mask, int per_scenario) {
    int r = ready(c);
    if (r) return r;
    if (mask & ~(MQI_SCORE_DOSE | MQI_SCORE_LETD | MQI_SCORE_STAT | MQI_SCORE_LETT | MQI_SCORE_EDEP | MQI_SCORE_DM))
        return fail(c, MQI_E_ARG, "unknown scorer bits");
    {
        const int n = ((mask & MQI_SCORE_DOSE) ? 1 : 0) + ((mask & MQI_SCORE_LETD) ? 2 : 0) +
                      ((mask & MQI_SCORE_LETT) ? 2 : 0) + ((mask & MQI_SCORE_EDEP) ? 1 : 0) +
                      ((mask & MQI_SCORE_DM) ? 1 : 0) + ((mask & MQI_SCORE_STAT) ? 2 : 0);
        if (n > MQI_MAX_SCORERS) return fail(c, MQI_E_ARG, "%d scorer arrays requested, at most %d", n, MQI_MAX_SCORERS);
    }
    c->mask     = mask;
    c->per_scen = per_scenario ? 1 : 0;
    c->kinds.clear();
    if (mask & MQI_SCORE_DOSE) c->kinds.push_back(MQI_KIND_DOSE);
    if (mask & MQI_SCORE_LETD) {
        c->kinds.push_back(MQI_KIND_LETD_NUM);
        c->kinds.push_back(MQI_KIND_LETD_DEN);
    }
    if (mask & MQI_SCORE_LETT) {
        c->kinds.push_back(MQI_KIND_LETT_NUM);
        c->kinds.push_back(MQI_KIND_LETT_DEN);
    }
    if (mask & MQI_SCORE_EDEP) c->kinds.push_back(MQI_KIND_EDEP);
    if (mask & MQI_SCORE_DM) c->kinds.push_back(MQI_KIND_DM);
    if (mask & MQI_SCORE_STAT) {
        c->kinds.push_back(MQI_KIND_STAT_DOSE);
        c->kinds.push_back(MQI_KIND_STAT_DOSE2);
    }
    const uint64_t n = nvox(c) * (c->per_scen ? (uint64_t) c->n_scen : 1ull);
    for (int k = 0; k < MQI_MAX_SCORERS; ++k) {
        if (k < (int) c->kinds.size()) {
            if ((r = dalloc(c, &c->d_out[k], n))) return r;
        } else {
            dfree(c->d_out[k]);
        }
    }
    return mqi_zero_scorers(c);
}

int
mqi_zero_scorers(mqi_ctx* c) {
    int r = ready(c);
    if (r) return r;
    const uint64_t n = nvox(c) * (c->per_scen ? (uint64_t) c->n_scen : 1ull);
    for (size_t k = 0; k < c->kinds.size(); ++k) KCHK(c, mqi_aux::fill_u64(c->d_out[k], 0ull, n, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int
mqi_scorer_count(mqi_ctx* c) {
    return c ? (int) c->kinds.size() : 0;
}

int
mqi_run(mqi_ctx* c, const mqi_run_config* cfg_in, mqi_run_result* res) {
    int r = ready(c);
    if (r) return r;
    if (!cfg_in) return fail(c, MQI_E_ARG, "cfg is NULL");
    if (c->spots.empty()) return fail(c, MQI_E_STATE, "no spots: call mqi_set_spots first");
    hipSetDevice(c->device);
    mqi_run_config cfg = *cfg_in;
    if ((r = check_tuning(c, cfg))) return r;
    if (cfg.scoring_loops < 0 || cfg.scoring_loops > 2) return fail(c, MQI_E_ARG, "bad scoring_loops");
    const int n_sc = (int) c->kinds.size() + (cfg.dij ? 1 : 0);
    if (n_sc == 0) return fail(c, MQI_E_STATE, "no scorers");
    if (n_sc > MQI_MAX_SCORERS) return fail(c, MQI_E_ARG, "the Dij scorer needs one of the %d scorer slots", MQI_MAX_SCORERS);
    if (cfg.dij && (r = alloc_dij(c, cfg))) return r;
    KCHK(c, mqi_aux::fill_u64(c->d_counters, 0ull, 4, c->stream));
    KCHK(c, mqi_aux::fill_u32(c->d_status, 0u, 1, c->stream));
    mqi_run_result out;
    std::memset(&out, 0, sizeof out);
    out.stat          = -1.0;
    const int max_rounds = cfg.max_rounds > 0 ? cfg.max_rounds : 64;
    run_plan  p          = make_plan(c, -1);
    // Dose and Dose_stat receive the same per-step value except in voxel 0
    // (excluded from Dose, mqi_roi.hpp): with each scorer scored once, score
    // Dose_stat only and derive Dose at the end of the run
    int       s1 = -1, s2 = -1;
    const bool alias = cfg.scoring_loops == MQI_LOOPS_ONCE && !c->kinds.empty() && c->kinds[0] == MQI_KIND_DOSE &&
                       stat_index(c, &s1, &s2) == 0 && !c->roi_set[0] && !c->roi_set[1];
    const uint64_t n_all = nvox(c) * (c->per_scen ? (uint64_t) c->n_scen : 1ull);
    if (alias) {
        if (c->snap_n != n_all) {
            if ((r = dalloc(c, &c->d_snap, n_all))) return r;
            c->snap_n = n_all;
        }
        HIPCHK(c, hipMemcpyAsync(c->d_snap, c->d_out[s1], sizeof(double) * n_all, hipMemcpyDeviceToDevice, c->stream));
    }
    struct alias_guard {
        mqi_ctx* c;
        ~alias_guard() { c->alias_dose = false; }
    } ag{c};
    c->alias_dose = alias;
    uint64_t  round_len  = p.max_item;
    double    ms         = 0;
    int       rounds     = 0;
    for (;;) {
        if ((r = transport_pass(c, cfg, cfg.hist_offset + (uint64_t) rounds * round_len, &ms))) return r;
        ++rounds;
        if (!(cfg.stat_criterion > 0)) break;
        unsigned long long tracked = 0;
        HIPCHK(c, hipMemcpy(&tracked, c->d_counters + 1, sizeof tracked, hipMemcpyDeviceToHost));
        if ((r = stat_value(c, 0, tracked, cfg.stat_threshold, &out.stat))) return r;
        if (out.stat <= cfg.stat_criterion || rounds >= max_rounds) break;
        if ((uint64_t) (rounds + 1) * round_len + cfg.hist_offset >= (1ull << 32)) break;
    }
    unsigned long long cnt[4] = { 0, 0, 0, 0 };
    unsigned int       st     = 0;
    HIPCHK(c, hipMemcpy(cnt, c->d_counters, sizeof cnt, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(&st, c->d_status, sizeof st, hipMemcpyDeviceToHost));
    // calculate_average_results: scorers 0..n-3 scaled by N_target / N_tracked
    // (mqi_tps_env.hpp:1430-1476, mqi_variables.hpp:51-66)
    const bool   rescale = cfg.stat_criterion > 0 && rounds > 1;
    const double w       = rescale ? (double) (float) (double) (p.total * 1.0 / (double) cnt[1]) : 1.0;
    if (alias) {
        KCHK(c, mqi_aux::dose_from_stat(c->d_out[0], c->d_out[s1], c->d_snap, n_all, nvox(c), w, c->stream));
    }
    if (rescale) {
        for (size_t k = alias ? 1 : 0; k + 2 < c->kinds.size(); ++k)
            KCHK(c, mqi_aux::scale(c->d_out[k], n_all, w, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    out.histories = cnt[1];
    out.dropped   = cnt[2];
    out.steps     = cnt[3];
    out.status    = st;
    out.rounds    = rounds;
    out.kernel_ms = ms;
    if (res) *res = out;
    if (st & MQI_STATUS_DIJ_FULL)
        return fail(c, MQI_E_CAPACITY, "a Dij table overflowed: raise dij_capacity%s",
                    cfg.dij_slots_sqrt > 0 ? " or dij_slots_sqrt" : "");
    return 0;
}

// One timed batched pass of the current plan (counters cleared first);
// returns the device time and the number of dropped secondaries.
int
trial_pass(mqi_ctx* c, const mqi_run_config& cfg, double* ms, unsigned long long* dropped) {
    KCHK(c, mqi_aux::fill_u64(c->d_counters, 0ull, 4, c->stream));
    double t = 0;
    int    r = transport_pass(c, cfg, 0, &t);
    if (r) return r;
    unsigned long long cnt[4];
    HIPCHK(c, hipMemcpy(cnt, c->d_counters, sizeof cnt, hipMemcpyDeviceToHost));
    *ms      = t;
    *dropped = cnt[2];
    return 0;
}

int
mqi_autotune(mqi_ctx* c, mqi_run_config* cfg, double sample, double* best_ms) {
    int r = ready(c);
    if (r) return r;
    if (!cfg || !(sample > 0.0) || sample > 1.0) return fail(c, MQI_E_ARG, "bad autotune arguments");
    if (c->spots.empty()) return fail(c, MQI_E_STATE, "no spots: call mqi_set_spots first");
    if (c->kinds.empty()) return fail(c, MQI_E_STATE, "no scorers");
    if (cfg->exact) return fail(c, MQI_E_ARG, "the verification kernels are not tuned");
    hipSetDevice(c->device);
    mqi_run_config base = *cfg;
    const bool     tune_stack = base.stack_depth < 0;
    base.stack_depth          = tune_stack ? 0 : base.stack_depth;
    base.batched              = 1;
    base.dij                  = 0;
    base.stat_criterion       = 0;
    base.blocks               = 0;
    if ((r = check_tuning(c, base))) return r;
    // trial plan: every spot's histories scaled by `sample` (at least one)
    struct restore_spots {
        mqi_ctx*              c;
        std::vector<mqi_spot> saved;
        ~restore_spots() { c->spots = saved; }
    } guard{c, c->spots};
    for (auto& sp : c->spots)
        if (sp.histories) sp.histories = std::max<uint64_t>(1, (uint64_t) std::llround((double) sp.histories * sample));
    // the trials score what mqi_run scores: with Dose derived from Dose_stat
    // (each scorer once, the statistics pair requested, no ROI) the kernel
    // scores two arrays, not three
    int        s1 = -1, s2 = -1;
    const bool alias = base.scoring_loops == MQI_LOOPS_ONCE && c->kinds[0] == MQI_KIND_DOSE &&
                       stat_index(c, &s1, &s2) == 0 && !c->roi_set[0] && !c->roi_set[1];
    struct alias_guard {
        mqi_ctx* c;
        ~alias_guard() { c->alias_dose = false; }
    } ag{c};
    c->alias_dose = alias;
    double             ms = 0;
    unsigned long long dropped = 0;
    if ((r = trial_pass(c, base, &ms, &dropped))) return r;   // warm-up
    if (tune_stack) {
        // the smallest secondary stack that drops nothing on the sample
        const int depths[] = {16, 32, 64, 128, 256, 512};
        for (int d : depths) {
            base.stack_depth = d;
            if ((r = trial_pass(c, base, &ms, &dropped))) return r;
            if (dropped == 0) break;
        }
    }
    // coordinate descent: register budget x workgroup size, then the refill
    // gathering, then the histories per fetch
    double         best = 1e300;
    mqi_run_config pick = base;
    pick.refill_lanes   = 4;
    pick.fetch_chunk    = 1;
    auto try_cfg = [&](const mqi_run_config& t) -> int {
        int e = trial_pass(c, t, &ms, &dropped);
        if (e) return e;
        if (ms < best) {
            best = ms;
            pick = t;
        }
        return 0;
    };
    const mqi_run_config start = pick;
    for (int w : {3, 4, 5})
        for (int b : {128, 256}) {
            mqi_run_config t = start;
            t.waves_per_simd = w;
            t.block_size     = b;
            if ((r = try_cfg(t))) return r;
        }
    // (a wave refilled as a whole, or nearly, keeps its lanes at similar
    // depths, so their scoring atomics cluster: 5.7-6.9% on the unpaired
    // 2000-spot CT plan; with paired scoring its idle lanes cost more)
    const mqi_run_config stage2 = pick;
    for (int k : {1, 2, 8, 16, 32, 60, 62, 64}) {
        mqi_run_config t = stage2;
        t.refill_lanes   = k;
        if ((r = try_cfg(t))) return r;
    }
    const mqi_run_config stage3 = pick;
    for (int ch : {2, 4}) {
        mqi_run_config t = stage3;
        t.fetch_chunk    = ch;
        if ((r = try_cfg(t))) return r;
    }
    cfg->waves_per_simd = pick.waves_per_simd;
    cfg->block_size     = pick.block_size;
    cfg->refill_lanes   = pick.refill_lanes;
    cfg->fetch_chunk    = pick.fetch_chunk;
    if (tune_stack) cfg->stack_depth = pick.stack_depth;
    if (best_ms) *best_ms = best;
    // trial passes scored into the scorers
    return mqi_zero_scorers(c);
}

int
mqi_run_vertices(mqi_ctx* c, const float* vtx, uint64_t n, const uint32_t* spot_ids, const mqi_run_config* cfg_in,
                 mqi_run_result* res) {
    int r = ready(c);
    if (r) return r;
    if (!vtx || !cfg_in || n == 0) return fail(c, MQI_E_ARG, "bad vertex arguments");
    if (n >= (1ull << 32)) return fail(c, MQI_E_ARG, "more than 2^32 vertices");
    if (c->kinds.empty()) return fail(c, MQI_E_STATE, "no scorers");
    hipSetDevice(c->device);
    mqi_run_config cfg = *cfg_in;
    if ((r = check_tuning(c, cfg))) return r;
    if (cfg.dij) return fail(c, MQI_E_ARG, "Dij scoring needs a spot table (mqi_run)");
    float*    d_v = nullptr;
    uint32_t* d_s = nullptr;
    if ((r = dalloc(c, &d_v, 7 * n))) return r;
    HIPCHK(c, hipMemcpyAsync(d_v, vtx, sizeof(float) * 7 * n, hipMemcpyHostToDevice, c->stream));
    if (spot_ids) {
        if ((r = dalloc(c, &d_s, n))) return r;
        HIPCHK(c, hipMemcpyAsync(d_s, spot_ids, sizeof(uint32_t) * n, hipMemcpyHostToDevice, c->stream));
    }
    const uint64_t one_end = n;
    const uint32_t zero    = 0;
    if ((r = ensure_items(c, 1))) return r;
    HIPCHK(c, hipMemcpyAsync(c->d_item_end, &one_end, sizeof one_end, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_item_spot, &zero, sizeof zero, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_item_scen, &zero, sizeof zero, hipMemcpyHostToDevice, c->stream));
    transport_params P;
    fill_params(c, P, cfg);
    P.vertices    = d_v;
    P.vertex_spot = d_s;
    P.n_items     = 1;
    P.total       = n;
    KCHK(c, mqi_aux::fill_u64(c->d_counters, 0ull, 4, c->stream));
    KCHK(c, mqi_aux::fill_u32(c->d_status, 0u, 1, c->stream));
    KCHK(c, mqi_aux::fill_u64(c->d_hoff, cfg.hist_offset, 1, c->stream));
    const int threads = cfg.block_size > 0 ? cfg.block_size : 256;
    int       blocks;
    if (cfg.batched) {
        blocks = persistent_blocks(c, cfg, threads, 0);
    } else {
        uint64_t lanes = std::min<uint64_t>(n, 1ull << 18);
        blocks         = (int) ((lanes + threads - 1) / threads);
    }
    if ((r = ensure_stack(c, (size_t) P.stack_cap * MQI_STACK_FLOATS * sizeof(float) * (size_t) blocks * threads))) return r;
    P.stack = c->d_stack;
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    KCHK(c, launch(c, P, cfg, cfg.batched ? 1 : 0, 0, blocks, threads));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float t = 0;
    HIPCHK(c, hipEventElapsedTime(&t, c->ev0, c->ev1));
    unsigned long long cnt[4] = { 0, 0, 0, 0 };
    HIPCHK(c, hipMemcpy(cnt, c->d_counters, sizeof cnt, hipMemcpyDeviceToHost));
    hipFree(d_v);
    if (d_s) hipFree(d_s);
    if (res) {
        std::memset(res, 0, sizeof *res);
        res->histories = cnt[1];
        res->dropped   = cnt[2];
        res->steps     = cnt[3];
        res->rounds    = 1;
        res->kernel_ms = t;
        res->stat      = -1.0;
    }
    return 0;
}

int
mqi_run_vertices_noalloc(mqi_ctx* c, const float* d_vtx, uint64_t n, const uint32_t* d_spot_ids,
                         const mqi_run_config* cfg_in, mqi_run_result* res) {
    int r = ready(c);
    if (r) return r;
    if (!d_vtx || !cfg_in || n == 0) return fail(c, MQI_E_ARG, "bad vertex arguments");
    if (n >= (1ull << 32)) return fail(c, MQI_E_ARG, "more than 2^32 vertices");
    if (c->kinds.empty()) return fail(c, MQI_E_STATE, "no scorers");
    if (!c->scratch)
        return fail(c, MQI_E_HIP, "no scratch address on device %d: %s", c->device, c->scratch_err.c_str());
    hipSetDevice(c->device);
    mqi_run_config cfg = *cfg_in;
    if ((r = check_tuning(c, cfg))) return r;
    if (cfg.dij) return fail(c, MQI_E_ARG, "Dij scoring needs a spot table (mqi_run)");
    unsigned char*      base     = c->scratch;
    unsigned long long* counters = (unsigned long long*) (base + SCR_COUNTERS);
    transport_params    P;
    fill_params(c, P, cfg);
    P.vertices      = d_vtx;
    P.vertex_spot   = d_spot_ids;
    P.n_items       = 1;
    P.total         = n;
    P.item_end      = (uint64_t*) (base + SCR_ITEM_END);
    P.item_spot     = (uint32_t*) (base + SCR_ITEM_SPT);
    P.item_scenario = (uint32_t*) (base + SCR_ITEM_SCN);
    P.queue         = (unsigned long long*) (base + SCR_QUEUE);
    P.tracked       = counters + 1;
    P.drops         = counters + 2;
    P.steps         = counters + 3;
    P.status        = (unsigned int*) (base + SCR_STATUS);
    P.hist_offset   = (uint64_t*) (base + SCR_HOFF);
    P.stack         = (float*) (base + MQI_SCRATCH_HEADER);
    // as many lanes as the scratch holds stacks for, at most the launch
    // mqi_run_vertices makes
    const int      threads    = cfg.block_size > 0 ? cfg.block_size : 256;
    const uint64_t block_size = (uint64_t) threads * P.stack_cap * MQI_STACK_FLOATS * sizeof(float);
    const uint64_t max_blocks = (uint64_t) MQI_NOALLOC_STACK_BYTES / block_size;
    if (max_blocks == 0)
        return fail(c, MQI_E_CAPACITY,
                    "stack_depth %d x block_size %d needs %llu bytes of secondary stacks per block; the "
                    "no-allocation scratch holds %llu", P.stack_cap, threads, (unsigned long long) block_size,
                    (unsigned long long) MQI_NOALLOC_STACK_BYTES);
    uint64_t blocks;
    if (cfg.batched)
        blocks = (uint64_t) persistent_blocks(c, cfg, threads, 0);
    else
        blocks = (std::min<uint64_t>(n, 1ull << 18) + threads - 1) / threads;
    blocks = std::min(blocks, max_blocks);
    // the work list (one item of n vertices) and the counters, by fill kernels
    KCHK(c, mqi_aux::fill_u64(base + SCR_ITEM_END, (unsigned long long) n, 1, c->stream));
    KCHK(c, mqi_aux::fill_u32(base + SCR_ITEM_SPT, 0u, 1, c->stream));
    KCHK(c, mqi_aux::fill_u32(base + SCR_ITEM_SCN, 0u, 1, c->stream));
    KCHK(c, mqi_aux::fill_u64(counters, 0ull, 4, c->stream));
    KCHK(c, mqi_aux::fill_u32(base + SCR_STATUS, 0u, 1, c->stream));
    KCHK(c, mqi_aux::fill_u64(base + SCR_HOFF, cfg.hist_offset, 1, c->stream));
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    KCHK(c, launch_bound(c, P, cfg, cfg.batched ? 1 : 0, 0, (int) blocks, threads, MQI_NOALLOC_STACK_BYTES));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float t = 0;
    HIPCHK(c, hipEventElapsedTime(&t, c->ev0, c->ev1));
    unsigned long long cnt[4] = { 0, 0, 0, 0 };
    HIPCHK(c, hipMemcpy(cnt, counters, sizeof cnt, hipMemcpyDeviceToHost));
    if (res) {
        std::memset(res, 0, sizeof *res);
        res->histories = cnt[1];
        res->dropped   = cnt[2];
        res->steps     = cnt[3];
        res->rounds    = 1;
        res->kernel_ms = t;
        res->stat      = -1.0;
    }
    return 0;
}

int
mqi_trace(mqi_ctx* c, const mqi_run_config* cfg_in, const float* vertices, uint64_t n_hist, int max_steps,
          int64_t* cnb, float* rec, int32_t* n_steps) {
    int r = ready(c);
    if (r) return r;
    if (!cfg_in || !cnb || !rec || !n_steps || max_steps <= 0 || n_hist == 0)
        return fail(c, MQI_E_ARG, "bad trace arguments");
    hipSetDevice(c->device);
    mqi_run_config cfg = *cfg_in;
    if ((r = check_tuning(c, cfg))) return r;
    run_plan       p;
    float*         d_v = nullptr;
    if (vertices) {
        p.total = n_hist;
        p.item_end.push_back(n_hist);
        p.item_spot.push_back(0);
        p.item_scen.push_back(0);
        if ((r = dalloc(c, &d_v, 7 * n_hist))) return r;
        HIPCHK(c, hipMemcpyAsync(d_v, vertices, sizeof(float) * 7 * n_hist, hipMemcpyHostToDevice, c->stream));
    } else {
        if (c->spots.empty()) return fail(c, MQI_E_STATE, "no spots");
        p = make_plan(c, -1);
        if (p.total != n_hist) return fail(c, MQI_E_ARG, "trace: plan has %llu histories, caller expects %llu",
                                           (unsigned long long) p.total, (unsigned long long) n_hist);
    }
    if ((r = upload_plan(c, p))) return r;
    int64_t* d_cnb = nullptr;
    float*   d_rec = nullptr;
    int32_t* d_n   = nullptr;
    const uint64_t slots = n_hist * (uint64_t) max_steps;
    if ((r = dalloc(c, &d_cnb, slots)) || (r = dalloc(c, &d_rec, slots * 12)) || (r = dalloc(c, &d_n, n_hist))) return r;
    KCHK(c, mqi_aux::fill_u32(d_n, 0u, n_hist, c->stream));
    transport_params P;
    fill_params(c, P, cfg);
    P.n_items   = (int) p.item_end.size();
    P.total     = p.total;
    P.trace_cnb = d_cnb;
    P.trace_rec = d_rec;
    P.trace_cap = max_steps;
    P.trace_n   = d_n;
    P.vertices  = d_v;
    KCHK(c, mqi_aux::fill_u64(c->d_hoff, cfg.hist_offset, 1, c->stream));
    KCHK(c, mqi_aux::fill_u64(c->d_counters, 0ull, 4, c->stream));
    const int threads = 64;
    const int blocks  = (int) ((n_hist + threads - 1) / threads);
    if ((r = ensure_stack(c, (size_t) P.stack_cap * MQI_STACK_FLOATS * sizeof(float) * (size_t) blocks * threads))) return r;
    P.stack = c->d_stack;
    KCHK(c, launch(c, P, cfg, 0, 1, blocks, threads));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(cnb, d_cnb, sizeof(int64_t) * slots, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(rec, d_rec, sizeof(float) * 12 * slots, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(n_steps, d_n, sizeof(int32_t) * n_hist, hipMemcpyDeviceToHost));
    hipFree(d_cnb);
    hipFree(d_rec);
    hipFree(d_n);
    if (d_v) hipFree(d_v);
    return 0;
}

int
mqi_get_scorer(mqi_ctx* c, int k, int scenario, double* out) {
    int r = ready(c);
    if (r) return r;
    if (k < 0 || k >= (int) c->kinds.size() || !out) return fail(c, MQI_E_ARG, "no scorer %d", k);
    if (scenario < 0 || scenario >= c->n_scen) return fail(c, MQI_E_ARG, "no scenario %d", scenario);
    const uint64_t off = c->per_scen ? (uint64_t) scenario * nvox(c) : 0;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(out, c->d_out[k] + off, sizeof(double) * nvox(c), hipMemcpyDeviceToHost));
    return 0;
}

int
mqi_scorer_device_ptr(mqi_ctx* c, int k, int scenario, double** p) {
    int r = ready(c);
    if (r) return r;
    if (k < 0 || k >= (int) c->kinds.size() || !p) return fail(c, MQI_E_ARG, "no scorer %d", k);
    if (scenario < 0 || scenario >= c->n_scen) return fail(c, MQI_E_ARG, "no scenario %d", scenario);
    *p = c->d_out[k] + (c->per_scen ? (uint64_t) scenario * nvox(c) : 0);
    return 0;
}

int
mqi_stat(mqi_ctx* c, int scenario, uint64_t n, float thr, double* out) {
    int r = ready(c);
    if (r) return r;
    if (!out) return MQI_E_ARG;
    return stat_value(c, scenario, n, thr, out);
}

// The product structures of a CSR on the device (m.indptr, m.col, m.val set;
// indptr its row pointers on the host): rows sorted by column (unless they
// are already), the voxel-major copy and plan of D w, the compacted columns
// of D^T g and, for long rows, its pieces.
int
prepare_products(mqi_ctx* c, csr_dev& m, const std::vector<int64_t>& indptr, bool sort_rows) {
    int r = 0;
    if ((r = build_plan(c, indptr, m.plan))) return r;
    if (m.nnz > 0) {
        // sort every row by voxel (rows handed in by the caller are checked
        // sorted instead)
        if (sort_rows) {
            // the segmented radix sort counts items in 32 bits
            if (m.nnz > (int64_t) INT32_MAX)
                return fail(c, MQI_E_CAPACITY, "%lld nonzeros in one Dij exceed the row sort's 2^31 - 1",
                            (long long) m.nnz);
            int32_t* tk = nullptr;
            float*   tv = nullptr;
            if ((r = dalloc(c, &tk, m.nnz)) || (r = dalloc(c, &tv, m.nnz))) return r;
            size_t bytes = 0;
            KCHK(c, mqi_aux::segmented_sort(m.col, m.val, tk, tv, m.nnz, (int) m.rows, m.indptr, m.indptr + 1,
                                            nullptr, &bytes, c->stream));
            char* scratch = nullptr;
            if ((r = dalloc(c, &scratch, std::max<size_t>(bytes, 1)))) return r;
            KCHK(c, mqi_aux::segmented_sort(m.col, m.val, tk, tv, m.nnz, (int) m.rows, m.indptr, m.indptr + 1,
                                            scratch, &bytes, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            hipFree(scratch);
            hipFree(tk);
            hipFree(tv);
        }
        // voxel-major copy for D w: row lengths by counting, entries by a
        // stable sort on the voxel (each voxel's spots ascending); its rows
        // are the voxels with a nonzero only (D w zeroes the others)
        uint32_t* d_cc = nullptr;
        if ((r = dalloc(c, &d_cc, m.cols)) || (r = dalloc(c, &m.tcol, m.nnz)) || (r = dalloc(c, &m.tval, m.nnz)))
            return r;
        KCHK(c, mqi_aux::fill_u32(d_cc, 0u, m.cols, c->stream));
        KCHK(c, mqi_aux::csr_transpose_count(m.col, m.nnz, d_cc, c->stream));
        std::vector<uint32_t> cc(m.cols);
        HIPCHK(c, hipMemcpy(cc.data(), d_cc, sizeof(uint32_t) * m.cols, hipMemcpyDeviceToHost));
        // voxels with a nonzero (the rows of the voxel-major copy and the
        // compacted gather of D^T g), each one's rank, and their bitmask
        std::vector<int32_t>  nz, rank(m.cols, 0);
        std::vector<int64_t>  tptr(1, 0);
        std::vector<uint32_t> bits((m.cols + 31) / 32, 0u);
        for (int64_t v = 0; v < m.cols; ++v)
            if (cc[v]) {
                rank[v] = (int32_t) nz.size();
                nz.push_back((int32_t) v);
                tptr.push_back(tptr.back() + cc[v]);
                bits[v >> 5] |= 1u << (v & 31);
            }
        m.n_nz = (int64_t) nz.size();
        if ((r = dalloc(c, &m.tindptr, m.n_nz + 1)) || (r = dalloc(c, &m.tnonempty, bits.size()))) return r;
        HIPCHK(c, hipMemcpy(m.tindptr, tptr.data(), sizeof(int64_t) * (m.n_nz + 1), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(m.tnonempty, bits.data(), sizeof(uint32_t) * bits.size(), hipMemcpyHostToDevice));
        if ((r = build_plan(c, tptr, m.tplan))) return r;
        int32_t* d_rank = nullptr;
        if ((r = dalloc(c, &m.nzvox, m.n_nz)) || (r = dalloc(c, &m.ccol, m.nnz)) || (r = dalloc(c, &m.xc, m.n_nz)) ||
            (r = dalloc(c, &d_rank, m.cols)))
            return r;
        HIPCHK(c, hipMemcpy(m.nzvox, nz.data(), sizeof(int32_t) * m.n_nz, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(d_rank, rank.data(), sizeof(int32_t) * m.cols, hipMemcpyHostToDevice));
        KCHK(c, mqi_aux::renumber(m.col, m.nnz, d_rank, m.ccol, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        hipFree(d_rank);
        // long rows (the Dij of pencil beams): D^T g over byte column deltas
        if (m.nnz >= (int64_t) MQI_SEG_SHORT * m.rows) {
            if ((r = build_pieces(c, m, indptr))) return r;
            dfree(m.ccol);
        }
        int key_bits = 1;
        while (key_bits < 32 && (1ll << key_bits) < m.cols) ++key_bits;
        int32_t*  d_keys = nullptr;
        uint64_t *d_pk = nullptr, *d_pk2 = nullptr;
        char*     d_tmp = nullptr;
        size_t    tbytes = 0;
        KCHK(c, mqi_aux::csr_transpose_sorted(m.indptr, m.col, m.val, m.rows, m.nnz, key_bits, nullptr, nullptr,
                                              nullptr, nullptr, nullptr, nullptr, &tbytes, c->stream));
        if ((r = dalloc(c, &d_keys, m.nnz)) || (r = dalloc(c, &d_pk, m.nnz)) || (r = dalloc(c, &d_pk2, m.nnz)) ||
            (r = dalloc(c, &d_tmp, std::max<size_t>(tbytes, 1))))
            return r;
        KCHK(c, mqi_aux::csr_transpose_sorted(m.indptr, m.col, m.val, m.rows, m.nnz, key_bits, d_keys, d_pk, d_pk2,
                                              m.tcol, m.tval, d_tmp, &tbytes, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        hipFree(d_keys);
        hipFree(d_pk);
        hipFree(d_pk2);
        hipFree(d_tmp);
        hipFree(d_cc);
        // D w over 16-bit spot columns
        if (m.rows <= 65536) {
            if ((r = dalloc(c, &m.tcol16, m.nnz))) return r;
            KCHK(c, mqi_aux::low16(m.tcol, m.nnz, m.tcol16, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            dfree(m.tcol);
        }
    }
    return 0;
}

int
mqi_dij_build(mqi_ctx* c, double rel_thr) {
    int r = ready(c);
    if (r) return r;
    if (!c->d_dij_keys) return fail(c, MQI_E_STATE, "no Dij tables: run with cfg.dij = 1 first");
    if (c->dij_items != (size_t) c->n_scen * c->spots.size())
        return fail(c, MQI_E_STATE, "the spots or scenarios changed since the run that scored the Dij tables");
    hipSetDevice(c->device);
    free_dij(c);
    const uint32_t n_items = (uint32_t) c->dij_items;
    const int      n_spots = (int) c->spots.size();
    double*        d_max   = nullptr;
    uint32_t*      d_cnt   = nullptr;
    if ((r = dalloc(c, &d_max, n_items)) || (r = dalloc(c, &d_cnt, n_items))) return r;
    KCHK(c, mqi_aux::dij_compact(c->d_dij_keys, c->d_dij_vals, c->d_dij_off, n_items, rel_thr, d_max, d_cnt,
                                 c->stream));
    std::vector<uint32_t> cnt(n_items);
    HIPCHK(c, hipMemcpy(cnt.data(), d_cnt, sizeof(uint32_t) * n_items, hipMemcpyDeviceToHost));
    c->dij.resize(c->n_scen);
    for (int s = 0; s < c->n_scen; ++s) {
        csr_dev& m = c->dij[s];
        m.rows     = n_spots;
        m.cols     = (int64_t) nvox(c);
        std::vector<int64_t> indptr(n_spots + 1, 0);
        for (int j = 0; j < n_spots; ++j) indptr[j + 1] = indptr[j] + cnt[(size_t) s * n_spots + j];
        m.nnz = indptr[n_spots];
        if ((r = dij_size_check(m.nnz)))
            return fail(c, r, "scenario %d's Dij has %lld nonzeros: its rows are sorted with 32-bit counts "
                              "(at most %lld); raise the threshold or split the plan", s, (long long) m.nnz,
                        (long long) MQI_MAX_SORT_NNZ);
        if ((r = dalloc(c, &m.indptr, n_spots + 1)) || (r = dalloc(c, &m.col, std::max<int64_t>(1, m.nnz))) ||
            (r = dalloc(c, &m.val, std::max<int64_t>(1, m.nnz))))
            return r;
        HIPCHK(c, hipMemcpy(m.indptr, indptr.data(), sizeof(int64_t) * (n_spots + 1), hipMemcpyHostToDevice));
        KCHK(c, mqi_aux::dij_write(c->d_dij_keys, c->d_dij_vals, c->d_dij_off + (size_t) s * n_spots, n_spots, rel_thr,
                                   d_max + (size_t) s * n_spots, m.indptr, m.col, m.val, c->stream));
        if ((r = prepare_products(c, m, indptr, true))) return r;
    }
    hipFree(d_max);
    hipFree(d_cnt);
    return 0;
}

int
mqi_dij_set_csr(mqi_ctx* c, int s, int64_t rows, int64_t cols, const int64_t* indptr, const int32_t* indices,
                const float* values) {
    if (!c) return MQI_E_ARG;
    int r = 0;
    if (s < 0 || s >= (1 << 16)) return fail(c, MQI_E_ARG, "scenario %d out of range", s);
    if (rows < 0 || rows >= (1ll << 31) || cols < 1 || cols >= (1ll << 31))
        return fail(c, MQI_E_ARG, "rows %lld / cols %lld out of range", (long long) rows, (long long) cols);
    if (!indptr) return fail(c, MQI_E_ARG, "indptr is NULL");
    const int64_t nnz = indptr[rows];
    if (nnz > 0 && (!indices || !values)) return fail(c, MQI_E_ARG, "indices or values are NULL");
    int64_t bad = 0;
    const int e = check_csr(rows, cols, indptr, indices, &bad);
    if (e == 1) return fail(c, MQI_E_ARG, "indptr must start at 0 and not decrease (row %lld)", (long long) bad);
    if (e == 2) return fail(c, MQI_E_ARG, "column out of [0, cols) at entry %lld", (long long) bad);
    if (e == 3) return fail(c, MQI_E_ARG, "columns must ascend strictly within a row (entry %lld)", (long long) bad);
    if ((r = dij_size_check(nnz)))
        return fail(c, r, "%lld nonzeros: at most %lld", (long long) nnz, (long long) MQI_MAX_SORT_NNZ);
    hipSetDevice(c->device);
    if ((int) c->dij.size() <= s) c->dij.resize(s + 1);
    csr_dev& m = c->dij[s];
    free_csr(m);
    m.rows = rows;
    m.cols = cols;
    m.nnz  = nnz;
    if ((r = dalloc(c, &m.indptr, rows + 1)) || (r = dalloc(c, &m.col, std::max<int64_t>(1, nnz))) ||
        (r = dalloc(c, &m.val, std::max<int64_t>(1, nnz))))
        return r;
    std::vector<int64_t> ip(indptr, indptr + rows + 1);
    HIPCHK(c, hipMemcpy(m.indptr, ip.data(), sizeof(int64_t) * (rows + 1), hipMemcpyHostToDevice));
    if (nnz) {
        HIPCHK(c, hipMemcpy(m.col, indices, sizeof(int32_t) * nnz, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(m.val, values, sizeof(float) * nnz, hipMemcpyHostToDevice));
    }
    if ((r = prepare_products(c, m, ip, false))) {
        free_csr(m);
        return r;
    }
    return 0;
}

int
mqi_dij_info(mqi_ctx* c, int s, int64_t* nnz, int64_t* rows, int64_t* cols) {
    if (!c) return MQI_E_ARG;
    if (s < 0 || s >= (int) c->dij.size()) return fail(c, MQI_E_STATE, "no Dij for scenario %d", s);
    if (nnz) *nnz = c->dij[s].nnz;
    if (rows) *rows = c->dij[s].rows;
    if (cols) *cols = c->dij[s].cols;
    return 0;
}

int
mqi_dij_copy(mqi_ctx* c, int s, int64_t* indptr, int32_t* indices, float* values) {
    if (!c) return MQI_E_ARG;
    if (s < 0 || s >= (int) c->dij.size()) return fail(c, MQI_E_STATE, "no Dij for scenario %d", s);
    const csr_dev& m = c->dij[s];
    if (indptr) HIPCHK(c, hipMemcpy(indptr, m.indptr, sizeof(int64_t) * (m.rows + 1), hipMemcpyDeviceToHost));
    if (indices && m.nnz) HIPCHK(c, hipMemcpy(indices, m.col, sizeof(int32_t) * m.nnz, hipMemcpyDeviceToHost));
    if (values && m.nnz) HIPCHK(c, hipMemcpy(values, m.val, sizeof(float) * m.nnz, hipMemcpyDeviceToHost));
    return 0;
}

int
mqi_set_roi(mqi_ctx* c, int target, const uint8_t* mask) {
    if (!c) return MQI_E_ARG;
    if (target != 0 && target != 1) return fail(c, MQI_E_ARG, "roi target %d (0 scorers, 1 statistics)", target);
    c->roi_set[target] = false;
    if (!mask) return c->roi_set[1 - target] ? roi_upload(c) : 0;
    int r = ready(c);
    if (r) return r;
    // mask_reader::mask_to_roi (mqi_file_handler.hpp:177-217): in voxel order
    // a 1 opens a run and a 0 closes it; other values (overlapping structures
    // summed) neither open nor close one.  A run still open at the last voxel
    // ends there (the reference leaves that stride unset).
    std::vector<uint32_t> start, stride;
    const uint64_t        n    = nvox(c);
    bool                  open = false;
    for (uint64_t v = 0; v < n; ++v) {
        if (mask[v] == 1 && !open) {
            open = true;
            start.push_back((uint32_t) v);
        } else if (mask[v] == 0 && open) {
            open = false;
            stride.push_back((uint32_t) (v - start.back()));
        }
    }
    if (open) stride.push_back((uint32_t) (n - start.back()));
    c->roi_start[target]  = std::move(start);
    c->roi_stride[target] = std::move(stride);
    c->roi_set[target]    = true;
    return roi_upload(c);
}

int
mqi_set_beamline(mqi_ctx* c, int n, const mqi_grid_desc* d) {
    if (!c) return MQI_E_ARG;
    if (n < 0 || n > 8) return fail(c, MQI_E_ARG, "beamline children: %d (0 to 8)", n);
    if (n > 0 && !d) return fail(c, MQI_E_ARG, "children is NULL");
    hipSetDevice(c->device);
    std::vector<int>   meta;
    std::vector<float> xf, edges, dens;
    for (int k = 0; k < n; ++k) {
        const mqi_grid_desc& g = d[k];
        if (!g.xe || !g.ye || !g.ze || !g.density) return fail(c, MQI_E_ARG, "child %d: NULL array", k);
        int r;
        if ((r = check_edges(c, g.xe, g.nx, "beamline x")) || (r = check_edges(c, g.ye, g.ny, "beamline y")) ||
            (r = check_edges(c, g.ze, g.nz, "beamline z")))
            return r;
        const uint64_t nv = (uint64_t) g.nx * g.ny * g.nz;
        if (nv >= (1ull << 31)) return fail(c, MQI_E_ARG, "child %d: more than 2^31 voxels", k);
        const uint64_t doff = dens.size();
        meta.insert(meta.end(), {g.nx, g.ny, g.nz, (int) edges.size(), (int) (uint32_t) doff, (int) (uint32_t) (doff >> 32), 0, 0});
        edges.insert(edges.end(), g.xe, g.xe + g.nx + 1);
        edges.insert(edges.end(), g.ye, g.ye + g.ny + 1);
        edges.insert(edges.end(), g.ze, g.ze + g.nz + 1);
        dens.insert(dens.end(), g.density, g.density + nv);
        xf.insert(xf.end(), g.rotation, g.rotation + 9);
        // grid3d::rotation_matrix_inv = fwd.inverse() (the transpose, mqi_matrix.hpp:290-293)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) xf.push_back(g.rotation[3 * j + i]);
        xf.insert(xf.end(), g.translation, g.translation + 3);
    }
    if (edges.size() > 16384) return fail(c, MQI_E_ARG, "beamline edges: %zu (at most 16384 staged in LDS)", edges.size());
    int r;
    c->bl_n       = 0;
    c->bl_n_edges = 0;
    if (n == 0) return 0;
    if ((r = dalloc(c, &c->d_bl_meta, meta.size())) || (r = dalloc(c, &c->d_bl_xf, xf.size())) ||
        (r = dalloc(c, &c->d_bl_edges, edges.size())) || (r = dalloc(c, &c->d_bl_dens, dens.size())))
        return r;
    HIPCHK(c, hipMemcpy(c->d_bl_meta, meta.data(), sizeof(int) * meta.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_bl_xf, xf.data(), sizeof(float) * xf.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_bl_edges, edges.data(), sizeof(float) * edges.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_bl_dens, dens.data(), sizeof(float) * dens.size(), hipMemcpyHostToDevice));
    c->bl_n       = n;
    c->bl_n_edges = (int) edges.size();
    return 0;
}

int
mqi_set_grid_ct_mha(mqi_ctx* c, const char* path, float density_scale, const float* shift) {
    if (!c || !path) return MQI_E_ARG;
    mqi_hip::io::ct_image ct;
    try {
        ct = mqi_hip::io::read_mha_ct(path, shift);
    } catch (const std::exception& e) {
        return fail(c, MQI_E_IO, "%s", e.what());
    }
    return mqi_set_grid_hu(c, ct.nx, ct.ny, ct.nz, ct.xe.data(), ct.ye.data(), ct.ze.data(), ct.hu.data(),
                           density_scale, nullptr, nullptr);
}

int
mqi_set_roi_mha(mqi_ctx* c, int target, int n, const char* const* paths) {
    if (!c) return MQI_E_ARG;
    if (target != 0 && target != 1) return fail(c, MQI_E_ARG, "roi target %d (0 scorers, 1 statistics)", target);
    if (n == 0) return mqi_set_roi(c, target, nullptr);
    if (n < 0 || !paths) return fail(c, MQI_E_ARG, "mask files: %d", n);
    int r = ready(c);
    if (r) return r;
    std::vector<std::string> p;
    for (int k = 0; k < n; ++k) {
        if (!paths[k]) return fail(c, MQI_E_ARG, "mask file %d is NULL", k);
        p.emplace_back(paths[k]);
    }
    std::vector<uint8_t> total;
    try {
        total = mqi_hip::io::read_mask_total(p, c->nx, c->ny, c->nz);
    } catch (const std::exception& e) {
        return fail(c, MQI_E_IO, "%s", e.what());
    }
    return mqi_set_roi(c, target, total.data());
}

int
mqi_dij_save_npz(mqi_ctx* c, int s, const char* path, double scale) {
    if (!c || !path) return MQI_E_ARG;
    if (s < 0 || s >= (int) c->dij.size()) return fail(c, MQI_E_STATE, "no Dij for scenario %d", s);
    const csr_dev& m = c->dij[s];
    if (m.cols > (int64_t) UINT32_MAX) return fail(c, MQI_E_ARG, "%lld voxels exceed the uint32 format", (long long) m.cols);
    std::vector<int64_t> ip(m.rows + 1);
    std::vector<int32_t> ix((size_t) m.nnz);
    std::vector<float>   v((size_t) m.nnz);
    int                  r = mqi_dij_copy(c, s, ip.data(), ix.data(), v.data());
    if (r) return r;
    try {
        mqi_hip::io::save_npz_csr(path, (uint32_t) m.rows, (uint32_t) m.cols, ip.data(), ix.data(), v.data(), scale);
    } catch (const std::exception& e) {
        return fail(c, MQI_E_IO, "%s", e.what());
    }
    return 0;
}

int
mqi_save_scorer(mqi_ctx* c, int k, int scenario, const char* dir, const char* name, int format, double scale) {
    int r = ready(c);
    if (r) return r;
    if (!dir || !name) return MQI_E_ARG;
    if (format < MQI_FORMAT_RAW || format > MQI_FORMAT_MHA) return fail(c, MQI_E_ARG, "unknown format %d", format);
    std::vector<double> d(nvox(c));
    if ((r = mqi_get_scorer(c, k, scenario, d.data()))) return r;
    try {
        if (format == MQI_FORMAT_RAW)
            mqi_hip::io::save_to_bin(d.data(), d.size(), scale, dir, name);
        else
            mqi_hip::io::save_metaimage(d.data(), c->nx, c->ny, c->nz, c->h_xe.data(), c->h_ye.data(), c->h_ze.data(),
                                        scale, dir, name, format == MQI_FORMAT_MHA);
    } catch (const std::exception& e) {
        return fail(c, MQI_E_IO, "%s", e.what());
    }
    return 0;
}

int
mqi_dij_device_ptrs(mqi_ctx* c, int s, const int64_t** indptr, const int32_t** indices, const float** values) {
    if (!c) return MQI_E_ARG;
    if (s < 0 || s >= (int) c->dij.size()) return fail(c, MQI_E_STATE, "no Dij for scenario %d", s);
    if (indptr) *indptr = c->dij[s].indptr;
    if (indices) *indices = c->dij[s].col;
    if (values) *values = c->dij[s].val;
    return 0;
}

int
mqi_dij_matvec(mqi_ctx* c, int s, int transpose, const float* x, void* y, int y_double, double alpha, int accumulate,
               void* stream) {
    if (!c || !x || !y) return MQI_E_ARG;
    if (s < 0 || s >= (int) c->dij.size()) return fail(c, MQI_E_STATE, "no Dij for scenario %d", s);
    hipSetDevice(c->device);
    const csr_dev& m  = c->dij[s];
    hipStream_t    st = stream ? (hipStream_t) stream : c->stream;
    HIPCHK(c, hipEventRecord(c->ev_mv0, st));
    if (transpose) {
        // D^T g: rows of the spot-major matrix
        const csr_plan& p = m.plan;
        if (m.sval) {   // pieces over byte column deltas of the compacted x
            KCHK(c, mqi_aux::gather_x(x, m.nzvox, m.n_nz, m.xc, st));
            KCHK(c, mqi_aux::pieces_staged(m.sval, m.sd8, m.pesc, m.xc, m.phdr, m.n_piece, m.row_piece, m.rows,
                                           m.ppart, y, y_double, alpha, accumulate, st));
        } else {
            const int32_t* colp = m.col;
            const float*   xp   = x;
            if (m.ccol) {   // gather from the compacted x (12 MB instead of 33.5 MB on the 5000-spot plan)
                KCHK(c, mqi_aux::gather_x(x, m.nzvox, m.n_nz, m.xc, st));
                colp = m.ccol;
                xp   = m.xc;
            }
            KCHK(c, mqi_aux::csr_adaptive(m.indptr, colp, m.val, xp, p.seg_k, p.seg_info, p.n_seg, p.split_seg,
                                          p.n_split, p.part, y, y_double, alpha, accumulate, nullptr, nullptr, 0,
                                          st));
        }
    } else if (m.nnz == 0) {
        // D w of an empty matrix: zeros (or y unchanged)
        if (!accumulate) {
            if (y_double)
                KCHK(c, mqi_aux::fill_u64(y, 0ull, m.cols, st));
            else
                KCHK(c, mqi_aux::fill_u32(y, 0u, m.cols, st));
        }
    } else {
        // D w: rows of the voxel-major copy (the voxels with a nonzero),
        // zeros at the others
        const csr_plan& p = m.tplan;
        if (m.tcol16)
            KCHK(c, mqi_aux::csr_adaptive16(m.tindptr, m.tcol16, m.tval, x, p.seg_k, p.seg_info, p.n_seg,
                                            p.split_seg, p.n_split, p.part, y, y_double, alpha, accumulate, m.nzvox,
                                            m.tnonempty, m.cols, st));
        else
            KCHK(c, mqi_aux::csr_adaptive(m.tindptr, m.tcol, m.tval, x, p.seg_k, p.seg_info, p.n_seg, p.split_seg,
                                          p.n_split, p.part, y, y_double, alpha, accumulate, m.nzvox, m.tnonempty,
                                          m.cols, st));
    }
    HIPCHK(c, hipEventRecord(c->ev_mv1, st));
    c->matvec_pending = true;
    return 0;
}

int
mqi_dij_quadratic(mqi_ctx* c, int s, const float* w, const double* p, const float* cw, double alpha, float* g,
                  double* f, void* stream) {
    if (!c || !w || !p || !f) return MQI_E_ARG;
    if (s < 0 || s >= (int) c->dij.size()) return fail(c, MQI_E_STATE, "no Dij for scenario %d", s);
    hipSetDevice(c->device);
    csr_dev& m = c->dij[s];
    if (m.nnz == 0) return fail(c, MQI_E_STATE, "scenario %d's Dij has no nonzeros", s);
    csr_plan& P = m.tplan;
    int       r = 0;
    if (!P.qpart &&
        (r = dalloc(c, &P.qpart, P.n_seg + P.n_split + MQI_QUAD_EMPTY_BLOCKS + MQI_QUAD_SUM_BLOCKS)))
        return r;
    hipStream_t       st = stream ? (hipStream_t) stream : c->stream;
    mqi_aux::quad_obj o;
    o.p         = p;
    o.c         = cw;
    o.g         = g;
    o.part      = P.qpart;
    o.two_inv_n = 2.0 / (double) m.cols;
    o.n_seg     = P.n_seg;
    const double inv_n = 1.0 / (double) m.cols;
    HIPCHK(c, hipEventRecord(c->ev_mv0, st));
    if (m.tcol16)
        KCHK(c, mqi_aux::csr_quadratic16(m.tindptr, m.tcol16, m.tval, w, P.seg_k, P.seg_info, P.n_seg, P.split_seg,
                                         P.n_split, P.part, alpha, o, inv_n, f, m.nzvox, m.tnonempty, m.cols, st));
    else
        KCHK(c, mqi_aux::csr_quadratic(m.tindptr, m.tcol, m.tval, w, P.seg_k, P.seg_info, P.n_seg, P.split_seg,
                                       P.n_split, P.part, alpha, o, inv_n, f, m.nzvox, m.tnonempty, m.cols, st));
    HIPCHK(c, hipEventRecord(c->ev_mv1, st));
    c->matvec_pending = true;
    return 0;
}

int
mqi_last_matvec_ms(mqi_ctx* c, double* ms) {
    if (!c || !ms) return MQI_E_ARG;
    if (!c->matvec_pending) return fail(c, MQI_E_STATE, "no sparse product was run");
    hipSetDevice(c->device);
    HIPCHK(c, hipEventSynchronize(c->ev_mv1));
    float t = 0;
    HIPCHK(c, hipEventElapsedTime(&t, c->ev_mv0, c->ev_mv1));
    *ms = t;
    return 0;
}

int
mqi_math_probe(mqi_ctx* c, int fn, const double* x, const double* y, int64_t n, double* out) {
    if (!c || !x || !out || n <= 0) return MQI_E_ARG;
    hipSetDevice(c->device);
    double *dx = nullptr, *dy = nullptr, *dout = nullptr;
    int     r;
    if ((r = dalloc(c, &dx, n)) || (r = dalloc(c, &dy, n)) || (r = dalloc(c, &dout, n))) return r;
    HIPCHK(c, hipMemcpy(dx, x, sizeof(double) * n, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(dy, y ? y : x, sizeof(double) * n, hipMemcpyHostToDevice));
    KCHK(c, mqi_aux::math_probe(fn, dx, dy, dout, n, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(out, dout, sizeof(double) * n, hipMemcpyDeviceToHost));
    hipFree(dx);
    hipFree(dy);
    hipFree(dout);
    return 0;
}

int
mqi_atomic_probe(mqi_ctx* c, int64_t n_slots, int64_t n_updates, int64_t window, double* ms) {
    if (!c || !ms || n_slots <= 0 || n_updates <= 0 || n_slots > (int64_t(1) << 32)) return MQI_E_ARG;
    if (window < 0 || (window > 0 && (window >= n_slots || (window & (window - 1)) != 0))) return MQI_E_ARG;
    hipSetDevice(c->device);
    double* d = nullptr;
    int     r = dalloc(c, &d, (size_t) n_slots);
    if (r) return r;
    int e = mqi_aux::fill_u64(d, 0ull, (uint64_t) n_slots, c->stream);
    // a warm-up pass, then the timed one
    if (!e) e = mqi_aux::atomic_probe(d, (uint64_t) n_slots, (uint64_t) n_updates, (uint64_t) window, c->stream);
    hipEvent_t e0 = c->ev0, e1 = c->ev1;
    if (!e) e = (int) hipEventRecord(e0, c->stream);
    if (!e) e = mqi_aux::atomic_probe(d, (uint64_t) n_slots, (uint64_t) n_updates, (uint64_t) window, c->stream);
    if (!e) e = (int) hipEventRecord(e1, c->stream);
    if (!e) e = (int) hipEventSynchronize(e1);
    float t = 0;
    if (!e) e = (int) hipEventElapsedTime(&t, e0, e1);
    hipFree(d);
    if (e) return fail(c, MQI_E_HIP, "atomic probe: %s", hipGetErrorString((hipError_t) e));
    *ms = t;
    return 0;
}

int
mqi_physics_probe(mqi_ctx* c, int exact, int physics_debug, int fn, const float* rho, const float* ek, int64_t n,
                  float* out) {
    if (!c || !rho || !out || n <= 0 || fn < 0 || fn > 2) return MQI_E_ARG;
    if (fn != 1 && !ek) return MQI_E_ARG;
    hipSetDevice(c->device);
    const int64_t n_out = fn == 2 ? 10 * n : n;
    float *       drho = nullptr, *dek = nullptr, *dout = nullptr;
    int           r;
    if ((r = dalloc(c, &drho, n)) || (r = dalloc(c, &dek, n)) || (r = dalloc(c, &dout, n_out))) {
        hipFree(drho);
        hipFree(dek);
        return r;
    }
    HIPCHK(c, hipMemcpy(drho, rho, sizeof(float) * n, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(dek, ek ? ek : rho, sizeof(float) * n, hipMemcpyHostToDevice));
    if (exact)
        KCHK(c, mqi_exact::physics_probe(physics_debug, fn, drho, dek, n, dout, c->stream));
    else
        KCHK(c, mqi_fast::physics_probe(physics_debug, fn, drho, dek, n, dout, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(out, dout, sizeof(float) * n_out, hipMemcpyDeviceToHost));
    hipFree(drho);
    hipFree(dek);
    hipFree(dout);
    return 0;
}

}   // extern "C"
