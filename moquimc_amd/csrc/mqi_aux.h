// Host-side launchers of the auxiliary kernels (mqi_aux.hip).
#ifndef MQI_AUX_H
#define MQI_AUX_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mqi_aux
{
int fill_u64(void* p, unsigned long long v, uint64_t n, hipStream_t s);
// counters of the per-XCD work ranges of a dynamic launch over `total` histories
int queue_init(unsigned long long* q, uint64_t total, hipStream_t s);
int fill_u32(void* p, uint32_t v, uint64_t n, hipStream_t s);
int hu_to_density(const int16_t* hu, float* rho, uint64_t n, float scale, hipStream_t s);
int math_probe(int fn, const double* x, const double* y, double* out, int64_t n, hipStream_t s);
// n_updates FP64 atomic adds to hashed slots of a (n_slots doubles); with
// window > 0 (a power of two below n_slots) the 64 adds of each wave fall in
// one window of that many slots placed at random
int atomic_probe(double* a, uint64_t n_slots, uint64_t n_updates, uint64_t window, hipStream_t s);
int stat_reduce(const double* s1, const double* s2, uint64_t n, double N, float thr, float* d_max, double* d_sum,
                unsigned long long* d_cnt, hipStream_t s);
int scale(double* p, uint64_t n, double w, hipStream_t s);
// dose += w * (stat - snap) outside voxel 0 of each block of nvox
int dose_from_stat(double* dose, const double* stat, const double* snap, uint64_t n, uint64_t nvox, double w,
                   hipStream_t s);
int dij_compact(const uint32_t* keys, const double* vals, const uint64_t* off, uint32_t n_items, double rel_thr,
                double* item_max, uint32_t* counts, hipStream_t s);
int dij_write(const uint32_t* keys, const double* vals, const uint64_t* off, uint32_t n_items, double rel_thr,
              const double* item_max, const int64_t* row_start, int32_t* cols, float* out_vals, hipStream_t s);
int segmented_sort(int32_t* keys, float* vals, int32_t* keys_tmp, float* vals_tmp, int64_t nnz, int n_seg,
                   const int64_t* seg_begin, const int64_t* seg_end, void* scratch, size_t* scratch_bytes,
                   hipStream_t s);
// CSR-Adaptive segment sizes (entries of one long-row piece; entries and rows
// of one packed segment)
#define MQI_SEG_LONG 16384
#ifndef MQI_SEG_SHORT
#define MQI_SEG_SHORT 1024
#endif
#define MQI_SEG_ROWS 1024
// rowmap: the element of y each row writes (NULL: row r writes y[r]).  A
// matrix stored without its empty rows (rowmap set) writes y = 0 at the n_y
// elements whose bit in nonempty is clear, from MQI_ZERO_BLOCKS workgroups of
// the same launch (with accumulate it leaves them alone).
#define MQI_ZERO_BLOCKS 1024
int csr_adaptive(const int64_t* indptr, const int32_t* col, const float* val, const float* x, const int64_t* seg_k,
                 const uint32_t* seg_info, int64_t n_seg, const int32_t* split_seg, int64_t n_split, double* part,
                 void* y, int y_is_double, double alpha, int accumulate, const int32_t* rowmap,
                 const uint32_t* nonempty, int64_t n_y, hipStream_t s);
// the same over 16-bit columns (D w of a Dij with at most 65536 spots)
int csr_adaptive16(const int64_t* indptr, const uint16_t* col, const float* val, const float* x, const int64_t* seg_k,
                   const uint32_t* seg_info, int64_t n_seg, const int32_t* split_seg, int64_t n_split, double* part,
                   void* y, int y_is_double, double alpha, int accumulate, const int32_t* rowmap,
                   const uint32_t* nonempty, int64_t n_y, hipStream_t s);
// The quadratic objective of y = alpha * A x on the D w matrix, without y:
// f = inv_n * sum_r c_r (y_r - p_r)^2 (c NULL: 1) into *f (device), and
// g_r = two_inv_n * c_r (y_r - p_r) (float32, when g is set).  part: n_seg +
// n_split doubles of per-segment and per-split-row terms, summed in a fixed
// order.  With rowmap the matrix has no empty rows: the n_y elements whose
// bit in nonempty is clear (d = 0) get their terms from a fixed grid of
// MQI_QUAD_EMPTY_BLOCKS workgroups (part then holds n_seg + n_split +
// MQI_QUAD_EMPTY_BLOCKS doubles).
#define MQI_QUAD_EMPTY_BLOCKS 2048
// the terms are summed in MQI_QUAD_SUM_BLOCKS fixed chunks, then the chunk
// sums (part holds MQI_QUAD_SUM_BLOCKS more doubles for them)
#define MQI_QUAD_SUM_BLOCKS 256
struct quad_obj {
    const double* p;
    const float*  c;
    float*        g;
    double*       part;
    double        two_inv_n;
    int64_t       n_seg;
};
int csr_quadratic(const int64_t* indptr, const int32_t* col, const float* val, const float* x, const int64_t* seg_k,
                  const uint32_t* seg_info, int64_t n_seg, const int32_t* split_seg, int64_t n_split, double* part,
                  double alpha, quad_obj o, double inv_n, double* f, const int32_t* rowmap, const uint32_t* nonempty,
                  int64_t n_y, hipStream_t s);
int csr_quadratic16(const int64_t* indptr, const uint16_t* col, const float* val, const float* x,
                    const int64_t* seg_k, const uint32_t* seg_info, int64_t n_seg, const int32_t* split_seg,
                    int64_t n_split, double* part, double alpha, quad_obj o, double inv_n, double* f,
                    const int32_t* rowmap, const uint32_t* nonempty, int64_t n_y, hipStream_t s);
// D^T g over column deltas (mqi_aux.hip): the 16-bit deltas of each row's
// columns (0 at row starts and where a row skips more than 65535 columns,
// those positions listed in gaps), the column before each piece's first
// delta, the pieces restaged lane-major in whole 256-nonzero batches (pb:
// staged starts, n_p + 1 entries) with one byte per delta, deltas of 255
// and more kept in per-piece escape lists (ecount: escapes per piece; ebase:
// their exclusive prefix sums), and the product: one wave per piece into
// part[piece] in the order of the headers, then each row's pieces summed in
// a fixed order.
#define MQI_PIECE 4096          // nonzeros per piece
#define MQI_PIECE_WAVES 8       // waves per workgroup of the product
#define MQI_PIECE_BATCHES 8     // 256-nonzero batches a wave has in flight
int deltas16(const int64_t* indptr, const int32_t* col, int64_t rows, int64_t nnz, uint16_t* dcol,
             unsigned long long* n_gaps, int64_t* gaps, int64_t cap, hipStream_t s);
int piece_base(const int32_t* col, const uint16_t* dcol, const int64_t* pk, int64_t n_p, int32_t* pbase,
               hipStream_t s);
int piece_escapes(const uint16_t* dcol, const int64_t* pk, int64_t n_p, uint32_t* ecount, hipStream_t s);
// esc: the escapes, plus 64 entries of padding the product may read
int stage_pieces(const float* val, const uint16_t* dcol, const int64_t* pk, const int64_t* pb, const uint32_t* ebase,
                 int64_t n_p, float* sval, uint8_t* sd8, uint16_t* esc, hipStream_t s);
// a piece as the product reads it: batches [batch0, batch0 + batches) of
// the staged arrays, the column its first delta adds to, its index, and the
// position of its first escape
struct alignas(16) piece_hdr {
    uint32_t batch0, batches;
    int32_t  base, piece;
    uint32_t esc0, pad[3];
};
int pieces_staged(const float* sval, const uint8_t* sd8, const uint16_t* esc, const float* x, const piece_hdr* hdr,
                  int64_t n_p, const int32_t* row_piece, int64_t rows, double* part, void* y, int y_is_double,
                  double alpha, int accumulate, hipStream_t s);
// low 16 bits of int32 columns (D w over at most 65536 spots)
int low16(const int32_t* col, int64_t nnz, uint16_t* out, hipStream_t s);
// D^T g on compacted x: xc = x[nzvox]; ccol = rank[col]
int gather_x(const float* x, const int32_t* nzvox, int64_t n, float* xc, hipStream_t s);
int renumber(const int32_t* col, int64_t nnz, const int32_t* rank, int32_t* ccol, hipStream_t s);
int csr_transpose_count(const int32_t* col, int64_t nnz, uint32_t* cnt, hipStream_t s);
// voxel-major copy (tcol = spot, tval) with each voxel's spots ascending:
// stable radix sort of the nonzeros by voxel (key_bits bits); with scratch
// NULL only *scratch_bytes is set.  keys_tmp: nnz ints, packed / packed_tmp:
// nnz 64-bit words.
int csr_transpose_sorted(const int64_t* indptr, const int32_t* col, const float* val, int64_t n_rows, int64_t nnz,
                         int key_bits, int32_t* keys_tmp, uint64_t* packed, uint64_t* packed_tmp, int32_t* tcol,
                         float* tval, void* scratch, size_t* scratch_bytes, hipStream_t s);
// grouped scoring: out[s][i] += pair[stride i + s] for s < ns, i < n (stride
// 2 or 4), pair cleared
int pair_fold(double* const* out, int ns, int stride, double* pair, uint64_t n, hipStream_t s);
}   // namespace mqi_aux

#endif
