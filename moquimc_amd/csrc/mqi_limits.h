// Size limits of the host runtime that callers can run into, as plain
// functions (no HIP), so they are testable on any host.
#ifndef MQI_LIMITS_H
#define MQI_LIMITS_H

#include <stdint.h>

#include "../../include/mqi_hip.h"

// Nonzeros of one scenario's Dij: its rows are sorted by voxel with a
// segmented radix sort that counts items in 32-bit ints.
#define MQI_MAX_SORT_NNZ 2147483647ll

// MQI_OK, or MQI_E_CAPACITY for a Dij too large to sort
static inline int
dij_size_check(int64_t nnz) {
    return (nnz < 0 || nnz > MQI_MAX_SORT_NNZ) ? MQI_E_CAPACITY : MQI_OK;
}

#endif
