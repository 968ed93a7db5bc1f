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

// One LDS tile of a workgroup: rows x stride doubles, stride = D | 1 (odd: lanes walking rows
// r * stride + k spread over the banks).  Tiles of about 8 KiB (at most 64 rows) let many waves
// per CU keep loads in flight; wide rows get at least 16 rows per tile (at most 56 KiB), so that
// enough lanes walk rows.  A broadcast configuration P is kept in LDS after the tile.
__host__ __device__ inline int cost_stride(int D) { return D | 1; }
__host__ __device__ inline int cost_tile_rows(int D) {
    const int row_bytes = 8 * cost_stride(D);
    int r = MRCC_COST_TILE_BYTES / row_bytes;
    if (r < 16) r = 57344 / row_bytes < 16 ? 57344 / row_bytes : 16;
    return r > 64 ? 64 : (r < 1 ? 1 : r);
}

#define COST_CH 4  // 16-byte loads per lane in flight per round of the vector staging

__device__ inline void cost_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// MODE 0: any strides; the tile holds d = P - Q, element by element.
// MODE 1: Q rows contiguous (qs = D), D even, 16-byte aligned, P broadcast (ps = 0): the tile holds
//         Q, copied with 16-byte loads; d = P[k] - tile, P from LDS.
// MODE 2: as 1 with P rows contiguous too (ps = D): the tile holds d = P - Q.
template <int MODE>
__global__ void __launch_bounds__(64) k_config_cost(CostArgs A) {
#pragma clang fp contract(off)
    extern __shared__ double tile[];
    const int lane = threadIdx.x;
    const int D = A.D, stride = cost_stride(D), R = cost_tile_rows(D);
    double* pb = tile + R * stride;  // broadcast configuration (MODE 1)
    if (MODE == 1) {
        for (int k = lane; k < D; k += 64) pb[k] = A.P[k];
    }
    const long long ntiles = (A.n + R - 1) / R;
    for (long long tix = blockIdx.x; tix < ntiles; tix += gridDim.x) {
        const long long row0 = tix * R;
        const long long left = A.n - row0;
        const int rows = left < R ? (int)left : R;
        if (MODE == 0) {
            // flat element t = lane + 64 i of the tile: row t / D, column t % D
            int r = lane / D, c = lane % D;
            const int r_step = 64 / D, c_step = 64 % D, total = rows * D;
#pragma unroll 8
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
        } else {
            // the tile's rows are one contiguous run of rows * D doubles: pairs u = 0 .. rows*D/2
            const int n2 = rows * D / 2;
            const double2* Q2 = (const double2*)(A.Q + row0 * D);
            const double2* P2 = (const double2*)(A.P + row0 * D);
            const int r_step = 128 / D, c_step = 128 % D;
            for (int u0 = 0; u0 < n2; u0 += 64 * COST_CH) {
                double2 v[COST_CH], w[COST_CH];
#pragma unroll
                for (int j = 0; j < COST_CH; ++j) {
                    const int u = u0 + lane + 64 * j;
                    if (u < n2) {
                        v[j] = Q2[u];
                        if (MODE == 2) w[j] = P2[u];
                    }
                }
                int r = (2 * (u0 + lane)) / D, c = (2 * (u0 + lane)) % D;
#pragma unroll
                for (int j = 0; j < COST_CH; ++j) {
                    const int u = u0 + lane + 64 * j;
                    if (u < n2) {
                        double* t = tile + r * stride + c;  // c even, c + 1 < D: same row
                        if (MODE == 2) {
                            t[0] = w[j].x - v[j].x;
                            t[1] = w[j].y - v[j].y;
                        } else {
                            t[0] = v[j].x;
                            t[1] = v[j].y;
                        }
                    }
                    r += r_step;
                    c += c_step;
                    if (c >= D) {
                        c -= D;
                        ++r;
                    }
                }
            }
        }
        cost_lds_sync();
        if (lane < rows) {
            const double* row = tile + lane * stride;
            double mx = 0.0, sm = 0.0;
            for (int i = 0; i < A.ns; ++i) {
                const int s = A.sl[2 * i], e = A.sl[2 * i + 1];
                double v;
                if (A.metric == MRCC_COST_EUCLIDEAN) {
                    double acc = 0.0;  // sum_squared += d * d, contracted, in coordinate order
                    for (int k = s; k < e; ++k) {
                        const double d = MODE == 1 ? pb[k] - row[k] : row[k];
                        acc = __builtin_fma(d, d, acc);
                    }
                    v = __builtin_sqrt(acc);
                } else {
                    // MAX: np.max(np.abs(.)) (a NaN propagates); ABSMAX: m = |d_s|, then
                    // m = |d_k| > m ? |d_k| : m
                    const bool nan_sticks = A.metric == MRCC_COST_MAX;
                    v = __builtin_fabs(MODE == 1 ? pb[s] - row[s] : row[s]);
                    for (int k = s + 1; k < e; ++k) {
                        const double u = __builtin_fabs(MODE == 1 ? pb[k] - row[k] : row[k]);
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
        cost_lds_sync();  // the next tile overwrites the LDS this one was read from
    }
}

void launch_config_cost(const CostArgs& A, int max_grid, hipStream_t st) {
    const int R = cost_tile_rows(A.D);
    const long long ntiles = (A.n + R - 1) / R;
    const int grid = (int)(ntiles < max_grid ? ntiles : max_grid);
    const bool vec = A.qs == A.D && (A.D & 1) == 0 && ((uintptr_t)A.Q & 15) == 0;
    if (vec && A.ps == 0) {
        const size_t lds = ((size_t)R * cost_stride(A.D) + A.D) * 8;
        k_config_cost<1><<<grid, 64, lds, st>>>(A);
    } else if (vec && A.ps == A.D && ((uintptr_t)A.P & 15) == 0) {
        k_config_cost<2><<<grid, 64, (size_t)R * cost_stride(A.D) * 8, st>>>(A);
    } else {
        k_config_cost<0><<<grid, 64, (size_t)R * cost_stride(A.D) * 8, st>>>(A);
    }
}
