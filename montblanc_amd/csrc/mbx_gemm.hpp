// Point-source coherency sums as a Hermitian matrix product on the matrix
// cores (mbx_gemm.hip).
#pragma once

#include "mbx_tile.hpp"

namespace mbx {

template <typename T> struct GemmArgs {
    const int32_t *bmap;   // (na x na) pair -> baseline map, device
    cx<T> *dst;            // running source sum, (ntime, nbl, nchan, 4)
    const cx<T> *vin;      // sum to add to (may alias dst) or NULL
    int rows, st, ntp, tpw;
};

// True when the antenna count fits the kernel's tile assignment.
template <typename T> bool point_gemm_supported(const mbx_dims &d);

// dst[t, b, c] = (vin ? vin[t, b, c] : 0) + sum over the point sources
// [0, npsrc extent) of J_p J_q^H, for every baseline of the map.
template <typename T>
int launch_point_gemm(const RimeArgs<T> &A, const int32_t *bmap, cx<T> *dst, const cx<T> *vin,
                      hipStream_t stream);

}  // namespace mbx
