// Host-side planning of the sparse products, as plain functions (no HIP), so
// they are testable on any host.
#ifndef MQI_PLAN_H
#define MQI_PLAN_H

#include <stddef.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

// Pieces of the D^T g product over column deltas (each at most 65535 within
// a piece): each row of a CSR
// (row pointers ip[0..rows]) is cut at its first nonzero, at every position
// in `gaps` (ascending; nonzeros whose column is more than 65535 past the
// previous one of the row; positions at a row start are ignored) and every
// `plen` nonzeros after a cut; an empty row gets one empty piece (the product
// writes every row through its pieces).  pk receives the first nonzero of
// each piece in order plus the sentinel ip[rows]; row_piece[r] (rows + 1
// entries) the first piece of row r.  Returns false when the pieces would not
// fit int32.
static inline bool
cut_pieces(const std::vector<int64_t>& ip, const std::vector<int64_t>& gaps, int64_t plen,
           std::vector<int64_t>& pk, std::vector<int32_t>& row_piece) {
    const int64_t rows = (int64_t) ip.size() - 1;
    pk.clear();
    row_piece.assign((size_t) rows + 1, 0);
    if (plen < 1) plen = 1;
    size_t gi = 0;
    for (int64_t i = 0; i < rows; ++i) {
        row_piece[i]    = (int32_t) pk.size();
        const int64_t b = ip[i + 1];
        if (ip[i] == b) pk.push_back(b);
        for (int64_t k = ip[i]; k < b;) {
            while (gi < gaps.size() && gaps[gi] <= k) ++gi;
            const int64_t cut = (gi < gaps.size() && gaps[gi] < b) ? gaps[gi] : b;
            pk.push_back(k);
            k = std::min(k + plen, cut);
        }
        if (pk.size() >= (size_t) INT32_MAX) return false;
    }
    row_piece[rows] = (int32_t) pk.size();
    pk.push_back(rows >= 0 ? ip[rows] : 0);
    return true;
}

// A CSR handed in by a caller (rows x cols, row pointers ip[0..rows], column
// indices idx[0 .. ip[rows])): 0 when ip starts at 0 and does not decrease
// and every row's columns lie in [0, cols) and ascend strictly; else 1 (ip,
// *bad = the row), 2 (a column out of range) or 3 (columns not ascending),
// *bad = the entry.
static inline int
check_csr(int64_t rows, int64_t cols, const int64_t* ip, const int32_t* idx, int64_t* bad) {
    // the row pointers first: idx is read only up to ip[rows]
    if (ip[0] != 0) {
        *bad = 0;
        return 1;
    }
    for (int64_t i = 0; i < rows; ++i)
        if (ip[i + 1] < ip[i]) {
            *bad = i;
            return 1;
        }
    for (int64_t i = 0; i < rows; ++i) {
        for (int64_t k = ip[i]; k < ip[i + 1]; ++k) {
            if (idx[k] < 0 || (int64_t) idx[k] >= cols) {
                *bad = k;
                return 2;
            }
            if (k > ip[i] && idx[k] <= idx[k - 1]) {
                *bad = k;
                return 3;
            }
        }
    }
    return 0;
}

#endif
