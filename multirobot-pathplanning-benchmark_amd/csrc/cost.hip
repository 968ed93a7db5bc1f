// cost.hip — costs and distances of configuration pairs on the device.
//
// The planners' metric functions (core/configuration.py): batch_config_cost / config_cost
// (:437-510, :352-363), batch_config_dist / config_dist (:303-349) and the numba reductions they
// call (compute_sliced_euclidean_dists :101-126, compute_sum_reduction :129-147,
// compute_max_sum_reduction :150-171, compute_abs_max_reduction :174-197, compute_max_reduction
// :200-219).  Per row r the difference d = P[r] - Q[r] is reduced per slice [s, e) to one value,
// and the slice values are reduced to the row's result.
//
// HBM-bound: each row is read once.  A workgroup is one wave; the wave copies a tile of rows
// (flat, coalesced) as differences into LDS, then lane r walks row r of the tile in coordinate
// order, so the sums keep the reference's sequential order.
#include "../../include/mrcc.h"
#include "engine.hpp"

// one LDS tile of a workgroup: rows x stride doubles, stride = D | 1 (odd: lanes walking rows
// r * stride + k spread over the banks)
__host__ __device__ inline int cost_stride(int D) { return D | 1; }
__host__ __device__ inline int cost_tile_rows(int D) {
    const int r = MRCC_COST_TILE_BYTES / (8 * cost_stride(D));
    return r > 64 ? 64 : r;
}

__global__ void __launch_bounds__(64) k_config_cost(CostArgs A) {
#pragma clang fp contract(off)
    extern __shared__ double tile[];
    const int lane = threadIdx.x;
    const int D = A.D, stride = cost_stride(D), R = cost_tile_rows(D);
    // this lane's first element of a tile (flat index lane) and the step of 64 elements
    const int r_first = lane / D, c_first = lane % D;
    const int r_step = 64 / D, c_step = 64 % D;
    const long long ntiles = (A.n + R - 1) / R;
    for (long long tix = blockIdx.x; tix < ntiles; tix += gridDim.x) {
        const long long row0 = tix * R;
        const long long left = A.n - row0;
        const int rows = left < R ? (int)left : R;
        const int total = rows * D;
        int r = r_first, c = c_first;
#pragma unroll 4
        for (int t = lane; t < total; t += 64) {
            const long long g = row0 + r;
            tile[r * stride + c] = A.P[g * A.ps + c] - A.Q[g * A.qs + c];
            r += r_step;
            c += c_step;
            if (c >= D) {
                c -= D;
                ++r;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        if (lane < rows) {
            const double* d = tile + lane * stride;
            double mx = 0.0, sm = 0.0;
            for (int i = 0; i < A.ns; ++i) {
                const int s = A.sl[2 * i], e = A.sl[2 * i + 1];
                double v;
                if (A.metric == MRCC_COST_EUCLIDEAN) {
                    double acc = 0.0;  // sum_squared += d * d, contracted, in coordinate order
                    for (int k = s; k < e; ++k) acc = __builtin_fma(d[k], d[k], acc);
                    v = __builtin_sqrt(acc);
                } else {
                    // MAX: np.max(np.abs(.)) (a NaN propagates); ABSMAX: m = |d_s|, then
                    // m = |d_k| > m ? |d_k| : m
                    const bool nan_sticks = A.metric == MRCC_COST_MAX;
                    v = __builtin_fabs(d[s]);
                    for (int k = s + 1; k < e; ++k) {
                        const double u = __builtin_fabs(d[k]);
                        if (u > v || (nan_sticks && u != u)) v = u;
                    }
                }
                if (i == 0) {
                    mx = v;
                    sm = A.red == MRCC_REDUCE_SUM ? 0.0 + v : v;
                } else {
                    mx = v > mx ? v : mx;
                    sm = sm + v;
                }
            }
            double res;
            if (A.red == MRCC_REDUCE_SUM) res = sm;
            else if (A.red == MRCC_REDUCE_MAX) res = mx;
            else res = __builtin_fma(A.w, sm, mx);  // max + w * sum, contracted
            A.out[row0 + lane] = res;
        }
        // the next tile overwrites the LDS this one was read from
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
}

void launch_config_cost(const CostArgs& A, int max_grid, hipStream_t st) {
    const int R = cost_tile_rows(A.D);
    const long long ntiles = (A.n + R - 1) / R;
    const int grid = (int)(ntiles < max_grid ? ntiles : max_grid);
    const size_t lds = (size_t)R * cost_stride(A.D) * 8;
    k_config_cost<<<grid, 64, lds, st>>>(A);
}
