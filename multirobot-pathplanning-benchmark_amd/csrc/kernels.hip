// kernels.hip — kernels of the collision engine (gfx950, wave64).
//
// Edge batches run as one pass-ordered work queue (see PlanWork):
//   prep      : one thread per edge: N (abstract_env.py:321-323), window (:325-331), slot count M,
//               passes S = ceil(M / G), histogram of S
//   scan      : one block: E_p = #{edges with S > p}, pass offsets PS[p] = G * sum_{p'<p} E_p'
//   scatter   : counting sort, edges longest-first
//   persistent: waves claim chunks of queue slots from 8 XCD-interleaved counters; slot -> (pass p,
//               rank r, k);
//               edge = order[r]; state = k * S_e + p of the edge (coarse-to-fine: pass 0 visits G
//               evenly spread states of every edge, later passes refine; an edge found in
//               collision is skipped by all later slots)
// The same state evaluators serve the direct kernels (one workgroup per edge: the plain per-call
// form and the unsorted single-launch ablation) and the configuration kernels.
#include "engine.hpp"

using namespace mrcc;

__device__ inline void stat_add(unsigned long long* stats, int idx, unsigned long long v) {
    if (stats && v) atomicAdd(stats + idx, v);
}

// row pointers of edge e
__device__ inline const double* row1(const EdgeBatchView& B, long long e) {
    long long r = B.idx1 ? (long long)B.idx1[e] : e;
    return B.q1 + r * B.rs;
}
__device__ inline const double* row1(const ArmQueue& B, long long e) {
    long long r = B.idx1 ? (long long)B.idx1[e] : e;
    return B.q1 + r * B.rs;
}
__device__ inline const double* row2(const EdgeBatchView& B, long long e) {
    long long r = B.idx2 ? (long long)B.idx2[e] : e;
    return B.q2 + r * B.rs;
}

__device__ inline int edge_N_dev(const EdgeBatchView& B, long long e, int D) {
    if (B.N) return B.N[e];
    if (B.N0 > 0) return B.N0;
    const double* a = row1(B, e);
    const double* b = row2(B, e);
    double m = __builtin_fabs(a[0] - b[0]);
#pragma unroll 8
    for (int k = 1; k < D; ++k) {  // unrolled: the row loads of several coordinates in flight
        double v = __builtin_fabs(a[k * B.ds] - b[k * B.ds]);
        m = v > m ? v : m;
    }
    return edge_N_from_max(m, B.res);
}

// interpolated state (abstract_env.py:341, :348): q1 + ((q2 - q1) / (N - 1)) * i, unfused
__device__ inline double interp_f64(double a, double b, int N, int i) {
#pragma clang fp contract(off)
    double dir = (b - a) / (double)(N - 1);
    return a + dir * (double)i;
}
__device__ inline double interp_dir_f64(double a, double dir, int i) {
#pragma clang fp contract(off)
    return a + dir * (double)i;
}

// ============================================================================================
// queue construction
__global__ void k_plan_init(PlanWork W) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = gridDim.x * blockDim.x;
    for (int b = t; b <= MRCC_MAX_PASSES; b += stride) W.hist[b] = 0;
    if (t < 8) W.meta[t] = 0;
    if (t < 8) W.ticket[32 * t] = 0;
    if (t < 8 && W.stats) W.stats[t] = 0;
}

// Pass counts of small edges are histogrammed in LDS first (one global atomic per block and
// count): edges of a batch share a few pass counts, and one global atomic per edge on the same few
// addresses serialises in L2.
#define PLAN_LDS_BINS 1024

// A group of `lpe` lanes (a power of two) shares an edge: lane `sub` holds coordinates sub,
// sub + lpe, ... of both rows in registers, the maximum is reduced across the group, the direction
// is written from the same registers, and lane sub == 0 writes the edge's queue entries.  The rows
// of the group's next edge are loaded before the current edge is processed.  D <= PREP_MAXC << lg_lpe.
#define PREP_MAXC 8

__device__ inline void prep_load_rows(const EdgeBatchView& B, long long e, int D, int sub, int lg_lpe,
                                      double* av, double* bv) {
    const double* a = row1(B, e);
    const double* b = row2(B, e);
#pragma unroll
    for (int j = 0; j < PREP_MAXC; ++j) {
        const int k = sub + (j << lg_lpe);
        av[j] = k < D ? a[k * B.ds] : 0.0;
        bv[j] = k < D ? b[k * B.ds] : 0.0;
    }
}

// nm > 0: mode ids outside [0, nm) make the edge an error (verdict 0, counted in meta[2])
__global__ void __launch_bounds__(256) k_plan_prep(EdgeBatchView B, PlanWork W, int D, int lgG,
                                                   int want_dir, int lg_lpe, int nm_modes) {
    __shared__ int h[PLAN_LDS_BINS];
    __shared__ int smax;
    for (int i = threadIdx.x; i < PLAN_LDS_BINS; i += blockDim.x) h[i] = 0;
    if (threadIdx.x == 0) smax = 0;
    __syncthreads();
    const int lpe = 1 << lg_lpe;
    const int sub = threadIdx.x & (lpe - 1);
    int mymax = 0;  // largest pass count seen by this thread
    const long long per_grid = (long long)gridDim.x * (blockDim.x >> lg_lpe);
    // the lanes of a group share e, so they run the same iterations and their shuffles stay within
    // active lanes
    const bool need_rows = (!B.N && B.N0 <= 0) || (want_dir && W.dir);
    const long long e_first = (long long)blockIdx.x * (blockDim.x >> lg_lpe) + (threadIdx.x >> lg_lpe);
    double an[PREP_MAXC], bn[PREP_MAXC];  // rows of the next edge
    if (need_rows && e_first < B.n) prep_load_rows(B, e_first, D, sub, lg_lpe, an, bn);
    for (long long e = e_first; e < B.n; e += per_grid) {
        double av[PREP_MAXC], bv[PREP_MAXC];
#pragma unroll
        for (int j = 0; j < PREP_MAXC; ++j) {
            av[j] = an[j];
            bv[j] = bn[j];
        }
        if (need_rows && e + per_grid < B.n) prep_load_rows(B, e + per_grid, D, sub, lg_lpe, an, bn);
        int N;
        if (B.N || B.N0 > 0) {
            N = B.N ? B.N[e] : B.N0;
        } else {
            // the sequential rule (edge_N_dev): m = |a0 - b0|, then m = v > m ? v : m, so a NaN
            // in coordinate 0 sticks and a NaN elsewhere is skipped.  Lane 0 starts from
            // coordinate 0, the other lanes from 0.0; lane 0's reduction keeps the rule.
            double m = sub == 0 ? __builtin_fabs(av[0] - bv[0]) : 0.0;
#pragma unroll
            for (int j = 0; j < PREP_MAXC; ++j) {
                const int k = sub + (j << lg_lpe);
                if (k > 0 && k < D) {
                    const double v = __builtin_fabs(av[j] - bv[j]);
                    m = v > m ? v : m;
                }
            }
            for (int o = 1; o < lpe; o <<= 1) {
                const double v = __shfl_xor(m, o, lpe);
                m = v > m ? v : m;
            }
            m = __shfl(m, 0, lpe);
            N = edge_N_from_max(m, B.res);
        }
        if (want_dir && W.dir) {
            const double den = (double)(N - 1);
#pragma unroll
            for (int j = 0; j < PREP_MAXC; ++j) {
                const int k = sub + (j << lg_lpe);
                if (k < D) {
#pragma clang fp contract(off)
                    W.dir[e * D + k] = (bv[j] - av[j]) / den;
                }
            }
        }
        if (sub != 0) continue;
        const int ns = batch_N_start(B, e);
        const int nm = batch_N_max(B, e);
        EdgeWin w = edge_window(N, ns, nm, B.incl);
        if (nm_modes > 0) {
            const int md = batch_mode(B, e);
            if (md < 0 || md >= nm_modes) {
                w.err = 1;
                w.M = 0;
            }
        }
        int M = w.M;
        long long S = ((long long)M + (1 << lgG) - 1) >> lgG;
        if (w.err) atomicAdd((unsigned long long*)&W.meta[2], 1ull);
        if (S > MRCC_MAX_PASSES) {
            unsigned long long slot = atomicAdd((unsigned long long*)&W.meta[3], 1ull);
            W.long_list[slot] = (int)e;
            S = 0;
        }
        W.erec[e] = make_int4(N, M, (int)S, w.w0 | (w.full << 31));
        B.out[e] = w.err ? 0 : 1;
        if (S < PLAN_LDS_BINS) atomicAdd(&h[S], 1);
        else atomicAdd(&W.hist[S], 1);
        mymax = S > mymax ? (int)S : mymax;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {  // wave maximum, then one LDS atomic per wave
        const int u = __shfl_xor(mymax, o, 64);
        mymax = u > mymax ? u : mymax;
    }
    if ((threadIdx.x & 63) == 0 && mymax > 0) atomicMax(&smax, mymax);
    __syncthreads();
    for (int i = threadIdx.x; i < PLAN_LDS_BINS; i += blockDim.x)
        if (h[i]) atomicAdd(&W.hist[i], h[i]);
    if (threadIdx.x == 0 && smax > 0) atomicMax((unsigned long long*)&W.meta[5], (unsigned long long)smax);
}

// Inclusive prefix sum over the threads of a 1024-thread block: shuffles within each wave, then
// the 16 wave totals through LDS (two barriers).
__device__ inline long long block_prefix_1024(long long v, long long* wsum) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const long long u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    if (lane == 63) wsum[wid] = v;
    __syncthreads();
    long long off = 0;
    for (int w = 0; w < wid; ++w) off += wsum[w];
    __syncthreads();  // wsum is reused by the next call
    return v + off;
}

// One block of 1024 threads: E_p = number of edges with more than p passes (suffix sums of the
// pass histogram) and the slot offset of each pass (prefix sums of E_p).  Only passes below the
// largest pass count (written by the prep kernel) are scanned; slot lookups never read further.
__global__ void __launch_bounds__(1024) k_plan_scan(PlanWork W, int lgG) {
    constexpr int PER_MAX = MRCC_MAX_PASSES / 1024;
    __shared__ long long wsum[16];
    const int t = threadIdx.x;
    const int maxS = (int)W.meta[5];             // largest pass count, <= MRCC_MAX_PASSES
    const int nb = maxS + 1;                     // passes 0..maxS (E_maxS = 0)
    const int per = (nb + 1023) / 1024;          // <= PER_MAX + 1
    // thread t owns passes [t*per, (t+1)*per) of [0, nb), in reverse order for the suffix sums
    long long loc[PER_MAX + 1];
    long long s = 0;
    for (int i = per - 1; i >= 0; --i) {
        const int p = t * per + i;
        s += (p < maxS) ? (long long)W.hist[p + 1] : 0;
        loc[i] = s;  // sum_{b in (p, (t+1)*per]}
    }
    // suffix sums over threads: the total minus the prefix sum
    const long long rs = block_prefix_1024(s, wsum);  // sum of s over threads <= t
    __shared__ long long total_s;
    if (t == 1023) total_s = rs;
    __syncthreads();
    const long long after = total_s - rs;             // sum of s over threads > t
    long long esum = 0;
    for (int i = 0; i < per; ++i) {
        const int p = t * per + i;
        if (p >= nb) break;
        const long long E = loc[i] + after;
        loc[i] = E;
        esum += E;
        W.cursor[p] = (int)E;  // descending-order start of bucket p
    }
    // exclusive prefix of E over p
    long long run = block_prefix_1024(esum, wsum) - esum;
    for (int i = 0; i < per; ++i) {
        const int p = t * per + i;
        if (p >= nb) break;
        W.pass_start[p] = run << lgG;
        run += loc[i];
    }
    // pass_start[maxS] (written above, E_maxS = 0) is the slot count; lookups never read beyond
    if (t == 1023) {
        W.meta[1] = maxS;
        W.meta[0] = run << lgG;
    }
}

// Tiles of 256 x SCATTER_EPT edges: edges of a small pass count are ranked within the tile in
// LDS; the tile then reserves one range of the count's bucket with one global atomic.
#define SCATTER_EPT 4

__global__ void __launch_bounds__(256) k_plan_scatter(PlanWork W, long long n) {
    __shared__ int cnt[PLAN_LDS_BINS];
    __shared__ int base[PLAN_LDS_BINS];
    const long long tile_edges = (long long)blockDim.x * SCATTER_EPT;
    for (long long t0 = (long long)blockIdx.x * tile_edges; t0 < n; t0 += (long long)gridDim.x * tile_edges) {
        for (int i = threadIdx.x; i < PLAN_LDS_BINS; i += blockDim.x) cnt[i] = 0;
        __syncthreads();
        int S[SCATTER_EPT], r[SCATTER_EPT];
#pragma unroll
        for (int k = 0; k < SCATTER_EPT; ++k) {
            const long long e = t0 + (long long)k * blockDim.x + threadIdx.x;
            S[k] = e < n ? W.erec[e].z : 0;
            r[k] = (S[k] > 0 && S[k] < PLAN_LDS_BINS) ? atomicAdd(&cnt[S[k]], 1) : 0;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < PLAN_LDS_BINS; i += blockDim.x)
            if (cnt[i]) base[i] = atomicAdd(&W.cursor[i], cnt[i]);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < SCATTER_EPT; ++k) {
            const long long e = t0 + (long long)k * blockDim.x + threadIdx.x;
            if (S[k] <= 0) continue;
            const int pos = S[k] < PLAN_LDS_BINS ? base[S[k]] + r[k] : atomicAdd(&W.cursor[S[k]], 1);
            W.order[pos] = (int)e;
        }
        __syncthreads();
    }
}

// Work queue of the persistent kernels: the slots are cut into chunks; chunk c belongs to counter
// c % 8.  A workgroup starts on counter blockIdx % 8 (the XCD it is dispatched to) and claims its
// chunks in increasing order, so all counters advance through the passes together; when its
// counter is exhausted it takes chunks from the others.  Returns the chunk, or -1 when every
// counter is exhausted.  Must be called by all lanes of the wave (wave-uniform result).
template <class WV>
__device__ inline long long claim_chunk(const WV& W, int& q, long long nchunks) {
    const int home = blockIdx.x & 7;
    while (q < 8) {
        const int x = (home + q) & 7;
        unsigned long long k = 0;
        if ((threadIdx.x & 63) == 0) k = atomicAdd(W.ticket + 32 * x, 1ull);
        k = ((unsigned long long)__builtin_amdgcn_readfirstlane((unsigned)(k >> 32)) << 32) |
            (unsigned)__builtin_amdgcn_readfirstlane((unsigned)k);
        const long long c = (long long)x + 8 * (long long)k;
        if (c < nchunks) return c;
        ++q;
    }
    return -1;
}

// The verdict byte of an edge in the persistent kernels, as every XCD sees it: another wave
// writes an edge's 0 when one of its states collides, and the later slots of the edge are skipped
// when they read it.  Each XCD has its own L2 and plain accesses may stay in it (a plain read can
// return an XCD's stale copy until the kernel ends); agent-scope accesses reach the level that all
// XCDs share.  (The direct kernels keep plain accesses: there the agent-scope ones measured slower,
// 47 -> 248 us for a list of 30 arm edges with 64 waves per edge.)
__device__ inline bool edge_dead(const unsigned char* out, int e) {
    return __hip_atomic_load(out + e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
}
__device__ inline void edge_kill(unsigned char* out, int e) {
    __hip_atomic_store(out + e, (unsigned char)0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct SlotRef {
    int e;      // edge, -1 when the slot is padding
    int local;  // slot within the edge
};

// `p` is a pass index with pass_start[p] <= s (a hint carried forward: slots claimed by one wave
// only increase, so the walk over all slots of a wave is at most n_pass steps)
__device__ inline SlotRef slot_ref(const PlanWork& W, int& p, long long s, int lgG, long long n) {
    SlotRef ref;
    ref.e = -1;
    ref.local = 0;
    if (W.pass_start[p] > s) p = 0;  // the wave moved to another queue
    while (p < MRCC_MAX_PASSES && W.pass_start[p + 1] <= s) ++p;
    const long long off = s - W.pass_start[p];
    const long long r = off >> lgG;
    const int k = (int)(off & ((1 << lgG) - 1));
    if (off < 0 || r >= n) return ref;  // unreachable for a consistent queue
    ref.e = W.order[r];
    if (ref.e < 0 || ref.e >= n) {
        ref.e = -1;
        return ref;
    }
    const int4 rec = W.erec[ref.e];
    ref.local = k * rec.z + p;
    if (ref.local >= rec.y) ref.e = -1;
    return ref;
}

__device__ inline EdgeWin win_of(const PlanWork& W, int e, int incl) {
    const int4 rec = W.erec[e];
    EdgeWin w;
    w.N = rec.x;
    w.w0 = erec_w0(rec);
    w.M = rec.y;
    w.full = erec_full(rec);
    w.first = incl ? 0 : 1;
    w.err = 0;
    return w;
}

// ============================================================================================
// lane-per-state kinds
template <class View, int T1, int T2>
__device__ inline bool lane_collides(const View& S, const double* q, unsigned mask, unsigned& rc);

// register length of the per-lane state: exact for the specialised shapes
template <class View, int T1, int T2>
struct LaneDim {
    static constexpr int value = MRCC_MAX_D;
};
template <>
struct LaneDim<DiskView, 2, 2> {
    static constexpr int value = 4;
};
template <>
struct LaneDim<PlanarView, 4, 0> {
    static constexpr int value = 12;
};
template <>
struct LaneDim<PlanarView, 0, 0> {
    static constexpr int value = 3 * MRCC_PLANAR_MAX_ROBOTS;
};
template <>
struct LaneDim<DiskView, 0, 0> {
    static constexpr int value = MRCC_DISK_MAX_RD * 4;
};

template <>
__device__ inline bool lane_collides<DiskView, 2, 2>(const DiskView& S, const double* q, unsigned mask, unsigned& rc) {
    return disk_collides<2, 2>(S, q, mask, rc);
}
template <>
__device__ inline bool lane_collides<DiskView, 0, 0>(const DiskView& S, const double* q, unsigned mask, unsigned& rc) {
    return disk_collides<0, 0>(S, q, mask, rc);
}
template <>
__device__ inline bool lane_collides<PlanarView, 4, 0>(const PlanarView& S, const double* q, unsigned mask, unsigned& rc) {
    return planar_collides<4>(S, q, mask, rc);
}
template <>
__device__ inline bool lane_collides<PlanarView, 0, 0>(const PlanarView& S, const double* q, unsigned mask, unsigned& rc) {
    return planar_collides<0>(S, q, mask, rc);
}

template <class View, int T1, int T2>
__global__ void __launch_bounds__(256) k_lane_persistent(View S, EdgeBatchView B, PlanWork W,
                                                         int lgG, int spt) {
    constexpr int MAXD = LaneDim<View, T1, T2>::value;
    const int D = S.D;
    const long long T = W.meta[0];
    unsigned long long n_eval = 0, n_dead = 0, n_slots = 0;
    unsigned rechecks = 0;
    const int lane = threadIdx.x & 63;
    int pw = 0;  // pass hint of this wave
    int q = 0;   // queues exhausted so far
    const long long chunk = (long long)spt * 64;
    const long long nchunks = (T + chunk - 1) / chunk;
    for (;;) {
        const long long c = claim_chunk(W, q, nchunks);
        if (c < 0) break;
        const long long s0 = c * chunk;
        if (W.pass_start[pw] > s0) pw = 0;
        while (pw < MRCC_MAX_PASSES && W.pass_start[pw + 1] <= s0) ++pw;
        for (int j = 0; j < spt; ++j) {
            const long long s = s0 + (long long)j * 64 + lane;
            if (s >= T) break;
            int p = pw;
            SlotRef ref = slot_ref(W, p, s, lgG, B.n);
            ++n_slots;
            if (ref.e < 0) continue;
            if (edge_dead(B.out, ref.e)) { ++n_dead; continue; }
            EdgeWin w = win_of(W, ref.e, B.incl);
            const int i = edge_state_index(w, ref.local, B.incl);
            if (i < 0) continue;
            const double* a = row1(B, ref.e);
            double q[MAXD];
#pragma unroll
            for (int k = 0; k < MAXD; ++k)
                if (k < D) q[k] = interp_dir_f64(a[k * B.ds], W.dir[(long long)ref.e * D + k], i);
            ++n_eval;
            const int mode = B.modes ? B.modes[ref.e] : 0;
            (void)mode;
            if (lane_collides<View, T1, T2>(S, q, B.mask, rechecks)) edge_kill(B.out, ref.e);
        }
    }
    if (W.stats) {
        stat_add(W.stats, 0, n_eval);
        stat_add(W.stats, 1, rechecks);
        stat_add(W.stats, 2, n_dead);
        stat_add(W.stats, 3, n_slots);
    }
}

// direct form: block = 64 lanes, `wpe` blocks per edge (block b -> edge b / wpe); lane-per-state
template <class View, int T1, int T2>
__global__ void __launch_bounds__(64) k_lane_direct(View S, EdgeBatchView B, const int* list,
                                                    long long n_list, const long long* n_dev,
                                                    int wpe, int init_out,
                                                    unsigned long long* stats) {
    constexpr int MAXD = LaneDim<View, T1, T2>::value;
    const int D = S.D;
    const int lane = threadIdx.x;
    unsigned long long n_eval = 0;
    unsigned rechecks = 0;
    if (n_dev) n_list = *n_dev;
    const long long nblocks = n_list * wpe;
    for (long long blk = blockIdx.x; blk < nblocks; blk += gridDim.x) {
        const long long li = blk / wpe;
        const int u = (int)(blk % wpe);
        const long long e = list ? (long long)list[li] : li;
        if (init_out && lane == 0) B.out[e] = 1;  // only with wpe == 1
        const int N = edge_N_dev(B, e, D);
        EdgeWin w = edge_window(N, batch_N_start(B, e), batch_N_max(B, e), B.incl);
        if (w.err) { B.out[e] = 0; continue; }
        const double* a = row1(B, e);
        const double* b = row2(B, e);
        for (int base = u * 64; base < w.M; base += wpe * 64) {
            if (B.out[e] == 0) break;  // wave-uniform: same address
            const int local = base + lane;
            bool coll = false;
            if (local < w.M) {
                const int i = edge_state_index(w, local, B.incl);
                if (i >= 0) {
                    double q[MAXD];
#pragma unroll
                    for (int k = 0; k < MAXD; ++k)
                        if (k < D) q[k] = interp_f64(a[k * B.ds], b[k * B.ds], N, i);
                    ++n_eval;
                    coll = lane_collides<View, T1, T2>(S, q, B.mask, rechecks);
                }
            }
            if (__any(coll)) {
                if (lane == 0) B.out[e] = 0;
                break;
            }
        }
    }
    if (stats) {
        stat_add(stats, 0, n_eval);
        stat_add(stats, 1, rechecks);
    }
}

template <class View, int T1, int T2>
__global__ void __launch_bounds__(256) k_lane_configs(View S, const double* qv, long long n,
                                                      long long rs, long long ds, unsigned mask,
                                                      unsigned char* out) {
    constexpr int MAXD = LaneDim<View, T1, T2>::value;
    const int D = S.D;
    unsigned rechecks = 0;
    for (long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x; r < n;
         r += (long long)gridDim.x * blockDim.x) {
        double q[MAXD];
#pragma unroll
        for (int k = 0; k < MAXD; ++k)
            if (k < D) q[k] = qv[r * rs + k * ds];
        out[r] = lane_collides<View, T1, T2>(S, q, mask, rechecks) ? 0 : 1;
    }
}

// ============================================================================================
// arms: one state per wavefront.  A workgroup of nw = blockDim.x / 64 waves stages the scene into
// LDS once; each wave then works as an independent worker (worker id blockIdx.x * nw + wave)
// with wave-level synchronisation only.
// coordinate `lane` of state i of edge e (lanes >= Dact: 0)
__device__ inline double arm_load_state(const EdgeBatchView& B, long long e, int N, int i, int Dact) {
    const int lane = wave_lane();
    double qv = 0.0;
    if (lane < Dact) {
        const double* a = row1(B, e);
        const double* b = row2(B, e);
        qv = interp_f64(a[lane * B.ds], b[lane * B.ds], N, i);
    }
    return qv;
}

template <bool STATS>
__device__ inline void arm_stats(unsigned long long* stats, const RecheckT<STATS>& rechecks,
                                 const ArmCountT<STATS>& cnt) {
    if constexpr (STATS) {
        if (!stats) return;
        stat_add(stats, 1, rechecks);
        stat_add(stats, 4, cnt.lb);
        stat_add(stats, 5, cnt.sb);
        stat_add(stats, 6, cnt.lp);
        stat_add(stats, 7, cnt.sp);
    }
}

// Persistent arm kernel.  A wave claims a chunk of `spt` (<= 64) consecutive slots (after chunks
// without live states, up to 64 / spt chunks of one counter at once); lane j resolves slot j
// (slot -> edge, state, mode; edge already dead?) so the dependent loads of all slots overlap;
// the wave then evaluates the live states one after another, k-major.  The ticket of the next
// chunk is requested before those evaluations and read after them, so the atomic's round trip
// overlaps the work.
template <bool GRASP, bool CHAIN, bool STATS>
__device__ inline void arm_persistent_body(const ArmK& V, const ArmQueue& B, const ArmQueue& W, int lgG,
                                           int spt) {
    extern __shared__ double lds_arm[];
    arm_stage_scene(lds_arm, V);
    const ArmLds Wl = arm_lds(lds_arm, V);
    const ArmRegs R = arm_regs(V, Wl);
    const int lane = wave_lane();
    const long long T = W.meta[0];
    // statistics (compiled away without STATS)
    typename std::conditional<STATS, unsigned long long, NoCount>::type n_eval{}, n_dead{}, n_slots{};
    RecheckT<STATS> rechecks{};
    ArmCountT<STATS> cnt{};
    int p = 0;  // pass hint (wave-uniform)
    int q = 0;  // queues exhausted so far
    // outcomes of the wave's last 16 states (1: collision); mostly colliding -> lazy placement
    unsigned recent = 0xffffu;
    const long long nchunks = (T + spt - 1) / spt;
    const int home = blockIdx.x & 7;
    long long ps_lo = W.pass_start[0], ps_hi = W.pass_start[1];  // pass p's slots (scalar)
    // A wave takes `nb` tickets of one counter at a time (chunks c, c + 8, ...; lane l resolves
    // slot l % spt of the batch's chunk l / spt).  One ticket while the chunks hold states to
    // evaluate; after a batch without one (the edges of its slots already collided), twice as many
    // next time, up to 64 slots: such batches cost their round trips whatever their size.
    const int nb_max = 64 / spt;
    const int lj = lane / spt, lo_slot = lane - lj * spt;
    int nb = 1;
    long long c = claim_chunk(W, q, nchunks);
    while (c >= 0) {
        const long long cj = c + 8 * (long long)lj;
        const long long sl = cj * spt + lo_slot;
        const bool valid = lj < nb && cj < nchunks && sl < T;
        // lane-parallel resolution of the batch's slots: slot -> (pass, rank, k) by arithmetic
        // (the current pass's bounds are in scalar registers), rank -> edge (one load), then every
        // per-edge word in one round of loads
        int e = -1, i = -1, N = 0, mode = 0, Dact = 0;
        int kl = 0;  // the slot's place among its edge's slots of the pass
        int pl = p;
        bool dead = false;
        if (valid) {
            long long lo = ps_lo;
            if (sl < ps_lo || sl >= ps_hi) {  // another pass: walk from the hint
                if (W.pass_start[pl] > sl) pl = 0;
                while (pl < MRCC_MAX_PASSES && W.pass_start[pl + 1] <= sl) ++pl;
                lo = W.pass_start[pl];
            }
            const long long off = sl - lo;
            const long long r = off >> lgG;
            const int k = (int)(off & ((1 << lgG) - 1));
            kl = k;
            const int ee = (off >= 0 && r < B.n) ? W.order[r] : -1;
            if (ee >= 0 && ee < B.n) {
                const int4 rec = W.erec[ee];
                const int S = rec.z, Mv = rec.y, Nv = rec.x, w0 = erec_w0(rec);
                const int full = erec_full(rec);
                const unsigned char o = edge_dead(B.out, ee) ? 0 : 1;
                const int md0 = batch_mode(B, ee);
                const int local = k * S + pl;
                if (local < Mv) {
                    if (o == 0) {
                        dead = true;
                    } else {
                        EdgeWin w;
                        w.N = Nv;
                        w.w0 = w0;
                        w.M = Mv;
                        w.full = full;
                        w.first = B.incl ? 0 : 1;
                        w.err = 0;
                        i = edge_state_index(w, local, B.incl);
                        if (i >= 0) {
                            e = ee;
                            N = Nv;
                            mode = md0;
                            const int md = (mode >= 0 && mode < V.nm) ? mode : 0;
                            Dact = arm_mact(V, Wl, md) * V.dof;
                        }
                    }
                }
            }
        }
        const unsigned long long vm = __ballot(valid);  // lane 0 is valid: c < nchunks
        {
            const int pn = __builtin_amdgcn_readlane(pl, 63 - __builtin_clzll(vm));  // the batch's largest pass
            if (pn != p) {
                p = pn;
                ps_lo = W.pass_start[p];
                ps_hi = W.pass_start[p + 1];
            }
        }
        n_slots += __popcll(vm);
        n_dead += __popcll(__ballot(dead));
        // The live slots are evaluated in turn, k-major: the first slot of every edge of the chunk,
        // then the second, ...  Most states of random edges collide, so an edge usually dies at
        // its first state; in this order the state after it belongs to another edge, and the inputs
        // loaded for it ahead of time (its edge's row and step, binary64) are used.  The slots of
        // edges that died in this chunk are skipped.  The next chunk's ticket is requested after
        // the first slot's loads, so none of these round trips sits on the evaluation path.
        const int G = 1 << lgG;
        int kc = 0;
        unsigned long long cur = __ballot(e >= 0 && kl == 0);
        unsigned long long deadm = 0;  // lanes whose edge was found in collision in this chunk
        auto next_slot = [&]() -> int {
            for (;;) {
                n_dead += __popcll(cur & deadm);
                cur &= ~deadm;
                if (cur) {
                    const int j = __builtin_ctzll(cur);
                    cur &= cur - 1ull;
                    return j;
                }
                if (++kc >= G) return -1;
                cur = __ballot(e >= 0 && kl == kc);
            }
        };
        int jn = next_slot();
        int en = -1, in_ = 0, mn = 0;
        double an = 0.0, dn = 0.0;
        if (jn >= 0) {
            en = __builtin_amdgcn_readlane(e, jn);
            in_ = __builtin_amdgcn_readlane(i, jn);
            mn = __builtin_amdgcn_readlane(mode, jn);
            const int Dn = __builtin_amdgcn_readlane(Dact, jn);
            if (lane < Dn) {
                an = row1(B, en)[lane * B.ds];
                dn = W.dir[(long long)en * V.D + lane];
            }
        }
        const int nb_next = jn >= 0 ? 1 : (2 * nb < nb_max ? 2 * nb : nb_max);
        const int x = (home + q) & 7;
        unsigned long long tk = 0;
        if (lane == 0) tk = atomicAdd(W.ticket + 32 * x, (unsigned long long)nb_next);
        while (jn >= 0) {
            const int jc = jn, ej = en, ij = in_, mj = mn;
            const double aj = an, dj = dn;
            jn = next_slot();
            if (jn >= 0) {
                en = __builtin_amdgcn_readlane(e, jn);
                in_ = __builtin_amdgcn_readlane(i, jn);
                mn = __builtin_amdgcn_readlane(mode, jn);
                const int Dn = __builtin_amdgcn_readlane(Dact, jn);
                an = 0.0;
                dn = 0.0;
                if (lane < Dn) {
                    an = row1(B, en)[lane * B.ds];
                    dn = W.dir[(long long)en * V.D + lane];
                }
            }
            if ((deadm >> jc) & 1ull) {  // picked before its edge was found in collision
                ++n_dead;
                continue;
            }
            const double qv = interp_dir_f64(aj, dj, ij);  // lanes >= D: unused
            ++n_eval;
            const bool col = arm_state_collides<GRASP, CHAIN, STATS>(V, Wl, R, qv, mj, B.mask, rechecks, cnt,
                                                                     __builtin_popcount(recent & 0xffffu) >= 8);
            recent = (recent << 1) | (col ? 1u : 0u);
            if (col) {
                if (lane == 0) edge_kill(B.out, ej);
                deadm |= __ballot(e == ej);
            }
        }
        const unsigned long long k = ((unsigned long long)__builtin_amdgcn_readfirstlane((unsigned)(tk >> 32)) << 32) |
                                     (unsigned)__builtin_amdgcn_readfirstlane((unsigned)tk);
        c = (long long)x + 8 * (long long)k;
        nb = nb_next;
        if (c >= nchunks) {  // counter x is exhausted: continue on the others
            ++q;
            c = claim_chunk(W, q, nchunks);
            nb = 1;
        }
    }
    if constexpr (STATS) {
        if (W.stats && lane == 0) {
            stat_add(W.stats, 0, n_eval);
            stat_add(W.stats, 2, n_dead);
            stat_add(W.stats, 3, n_slots);
        }
    }
    arm_stats<STATS>(W.stats, rechecks, cnt);
}

// Scenes without held boxes fit 4 waves per SIMD (<= 128 VGPRs), in workgroups of up to 4 waves;
// the grasp variant 3, in workgroups of up to MRCC_PERSISTENT_GRASP_NW waves (10: 3, 3, 2 and 2 on
// the CU's SIMDs).  DH and chain joints, held boxes and statistics are template parameters.
template <bool GRASP, bool CHAIN, bool STATS>
__global__ void __launch_bounds__(GRASP ? 64 * MRCC_PERSISTENT_GRASP_NW : 256, GRASP ? 3 : 4)
    k_arm_persistent(ArmK V, ArmQueue Q, int lgG, int spt) {
    arm_persistent_body<GRASP, CHAIN, STATS>(V, Q, Q, lgG, spt);
}

// Work item k of n_list * wpe: edge list[k / wpe] (or k / wpe), states k % wpe, k % wpe + wpe, ...
// With init_out (wpe == 1 only) the wave owns its edge: it writes the verdict without reading it.
// A mode id outside [0, nm) makes the edge's verdict 0 (the error contract of mujoco_env.py:797-807,
// "exception => collision"); entry points with host-side mode ids reject them before launching.
template <bool GRASP, bool CHAIN>
__global__ void __launch_bounds__(256) k_arm_direct(ArmK V, EdgeBatchView B, const int* list,
                                                    long long n_list, const long long* n_dev,
                                                    int wpe, int init_out,
                                                    unsigned long long* stats) {
    extern __shared__ double lds_arm[];
    arm_stage_scene(lds_arm, V);
    const ArmLds Wl = arm_lds(lds_arm, V);
    const ArmRegs R = arm_regs(V, Wl);
    const int nw = blockDim.x >> 6;
    unsigned long long n_eval = 0;
    unsigned rechecks = 0;
    ArmCountT<true> cnt = {0, 0, 0, 0};
    if (n_dev) n_list = *n_dev;
    const long long nitems = n_list * wpe;
    const long long w0 = (long long)blockIdx.x * nw + (threadIdx.x >> 6);
    for (long long it = w0; it < nitems; it += (long long)gridDim.x * nw) {
        const long long li = it / wpe;
        const int u = (int)(it % wpe);
        const long long e = list ? (long long)list[li] : li;
        const int mode = batch_mode(B, e);
        const int md = (mode >= 0 && mode < V.nm) ? mode : 0;
        const int N = edge_N_dev(B, e, V.D);
        EdgeWin w = edge_window(N, batch_N_start(B, e), batch_N_max(B, e), B.incl);
        if (w.err || md != mode) {
            if (wave_lane() == 0) B.out[e] = 0;
            continue;
        }
        bool hit = false;
        for (int local = u; local < w.M; local += wpe) {
            if (!init_out && B.out[e] == 0) break;
            const int i = edge_state_index(w, local, B.incl);
            if (i < 0) continue;
            const double qv = arm_load_state(B, e, N, i, arm_mact(V, Wl, md) * V.dof);
            ++n_eval;
            if (arm_state_collides<GRASP, CHAIN, true>(V, Wl, R, qv, mode, B.mask, rechecks, cnt)) {
                hit = true;
                break;
            }
        }
        if (wave_lane() == 0) {
            if (hit) B.out[e] = 0;
            else if (init_out) B.out[e] = 1;
        }
    }
    if (stats && wave_lane() == 0) stat_add(stats, 0, n_eval);
    arm_stats<true>(stats, rechecks, cnt);
}

// One wave per configuration.  The wave's first configuration (its row, its mode and the joint
// sines and cosines) is loaded and computed while the scene image is in flight, as in the per-call
// edge kernel; a single configuration (the per-call form) is then one staging round trip plus
// the evaluation.
template <bool GRASP, bool CHAIN>
__global__ void __launch_bounds__(256, 3) k_arm_configs(ArmK V, const double* qv, long long n,
                                                     long long rs, long long ds, const int* modes,
                                                     unsigned mask, unsigned char* out) {
    extern __shared__ double lds_arm[];
    const int nw = blockDim.x >> 6;
    const int lane = wave_lane();
    const long long r0 = (long long)blockIdx.x * nw + (threadIdx.x >> 6);
    // coordinates of inactive arms are loaded too and zeroed once the mode is known
    const double q0 = r0 < n && lane < V.D ? qv[r0 * rs + (long long)lane * ds] : 0.0;
    const int mode0 = r0 < n && modes ? modes[r0] : 0;
    const auto stage = arm_stage_load<4>(V);
    const SinCos t0 = sincos_f64(q0);
    arm_stage_store(lds_arm, V, stage);  // ends with a barrier
    const ArmLds Wl = arm_lds(lds_arm, V);
    const ArmRegs R = arm_regs(V, Wl);
    NoCount rechecks;
    ArmCountT<false> cnt;
    for (long long r = r0; r < n; r += (long long)gridDim.x * nw) {
        const bool first = r == r0;
        const int mode = first ? mode0 : (modes ? modes[r] : 0);
        const int md = (mode >= 0 && mode < V.nm) ? mode : 0;
        const int Dact = arm_mact(V, Wl, md) * V.dof;
        const double q = lane < Dact ? (first ? q0 : qv[r * rs + (long long)lane * ds]) : 0.0;
        // a mode outside [0, nm): verdict 0 (arm_state_collides reports it as a collision)
        const bool c = arm_state_collides<GRASP, CHAIN, false>(V, Wl, R, q, mode, mask, rechecks, cnt, false,
                                                               first, t0);
        if (lane == 0) out[r] = c ? 0 : 1;
    }
}

// Plain per-call form: one launch checks edge 0 of B with one workgroup of nw waves; wave w takes
// the edge's slots w, w + nw, ...  The waves share the verdict through LDS: a wave that finds a
// collision raises the flag, the others stop before their next state, and thread 0 writes the
// verdict once every wave is done, so the output needs no initialisation.  With several
// workgroups (long edges from the host entry point), wave w of workgroup g takes the slots
// g nw + w, then every G nw further; the caller presets the verdict to 1, a workgroup writes 0
// only, and completion is the caller's to observe (stream order).  Spheres beyond the
// first 64 are placed when a candidate pair needs them (most states of a per-call stream collide
// within the first arm).  The rows are loaded first, then the scene image; while the image is in
// flight each wave computes the edge's state count and the joint sines and cosines of its first
// state, so the staging round trip overlaps the rows' round trip and that arithmetic.
template <bool GRASP, bool CHAIN>
__global__ void __launch_bounds__(512) k_arm_edge(ArmK V, EdgeBatchView B) {
    extern __shared__ double lds_arm[];
    __shared__ int s_hit;
    const int lane = wave_lane();
    const int D = V.D;
    // coordinate `lane` of both rows (lanes >= D: 0)
    const double* a = row1(B, 0);
    const double* b = row2(B, 0);
    const double av = lane < D ? a[lane * B.ds] : 0.0;
    const double bv = lane < D ? b[lane * B.ds] : 0.0;
    const auto stage = arm_stage_load<4>(V);  // 16 KiB of the image at 256 threads
    if (threadIdx.x == 0) s_hit = 0;
    int N;
    if (B.N || B.N0 > 0) {
        N = B.N ? B.N[0] : B.N0;
    } else {
        // edge_N_dev's rule (m = |a0 - b0|, then m = v > m ? v : m over the coordinates): a NaN in
        // coordinate 0 sticks and a NaN elsewhere is skipped, as in the sequential loop; the other
        // values are >= 0, so the maximum of the non-NaN values (0 for lanes >= D) is the loop's
        const double m = lane < D ? __builtin_fabs(av - bv) : 0.0;
        const double m0 = readlane_f64(m, 0);
        const double mx = wave_max_f64(__builtin_isnan(m) ? 0.0 : m);
        N = edge_N_from_max(__builtin_isnan(m0) ? m0 : mx, B.res);
    }
    const int nw = blockDim.x >> 6, wv = threadIdx.x >> 6;
    const int g0 = (int)blockIdx.x * nw + wv, gstep = (int)gridDim.x * nw;  // the wave's first slot, stride
    const int mode = batch_mode(B, 0);
    const bool bad_mode = mode < 0 || mode >= V.nm;
    const EdgeWin w = edge_window(N, batch_N_start(B, 0), batch_N_max(B, 0), B.incl);
    const bool run = !w.err && !bad_mode;
    // the state of slot `local` (abstract_env.py:341, :348); i < 0: the slot has no state
    auto state_q = [&](int i) {
#pragma clang fp contract(off)
        return av + ((bv - av) / (double)(N - 1)) * (double)i;
    };
    // the wave's first state and its joint sines / cosines, before the image is stored
    const int i0 = run && g0 < w.M ? edge_state_index(w, g0, B.incl) : -1;
    double q0 = 0.0;
    SinCos t0 = SinCos{0.0, 1.0};
    if (i0 >= 0) {  // wave-uniform
        q0 = state_q(i0);
        t0 = sincos_f64(q0);
    }
    arm_stage_store(lds_arm, V, stage);  // ends with a barrier
    const ArmLds Wl = arm_lds(lds_arm, V);
    const ArmRegs R = arm_regs(V, Wl);
    NoCount rechecks;
    ArmCountT<false> cnt;
    if (run) {
        volatile int* hit = &s_hit;
        for (int local = g0; local < w.M; local += gstep) {
            if (*hit) break;  // wave-uniform: one LDS word
            const bool first = local == g0;
            const int i = first ? i0 : edge_state_index(w, local, B.incl);
            if (i < 0) continue;
            const double qv = first ? q0 : state_q(i);
            if (arm_state_collides<GRASP, CHAIN, false, true>(V, Wl, R, qv, mode, B.mask, rechecks, cnt, true,
                                                              first, t0)) {
                if (lane == 0) *hit = 1;
                break;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const bool coll = w.err || bad_mode || s_hit;
        if (gridDim.x == 1) B.out[0] = coll ? 0 : 1;
        else if (coll) B.out[0] = 0;
    }
}

// ---- per-call lane kernels: the scene's binary32 tables staged in LDS ----------------------------
// A per-call launch starts with cold caches, and the evaluators read their tables inside loops with
// early exits, one dependent round trip per cache line (100 planar obstacles: 25 of them).  The
// staged form issues every load of the tables at once, before the edge rows, and the evaluators
// then read LDS.  The binary64 tables of the rechecks stay in global memory.

#define LANE_STAGE_FLOATS 8192  // 32 KiB of LDS; larger scenes run unstaged

// One table: K loads per thread issued by load(), stored by store() (the rest of a longer table
// after them, K loads in flight per thread).
template <int K>
struct TabPre {
    float v[K];
    __device__ void load(const float* src, int n) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int e = (int)threadIdx.x + k * (int)blockDim.x;
            v[k] = e < n ? src[e] : 0.0f;
        }
    }
    __device__ void store(float* dst, const float* src, int n) const {
        const int nt = (int)blockDim.x;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int e = (int)threadIdx.x + k * nt;
            if (e < n) dst[e] = v[k];
        }
        for (int e0 = (int)threadIdx.x + K * nt; e0 < n; e0 += K * nt) {
            float r[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int e = e0 + k * nt;
                r[k] = e < n ? src[e] : 0.0f;
            }
#pragma unroll
            for (int k = 0; k < K; ++k)
                if (e0 + k * nt < n) dst[e0 + k * nt] = r[k];
        }
    }
};

__host__ __device__ inline int pad4(int n) { return (n + 3) & ~3; }

template <class View> struct LaneStage;

// planar scenes: obstacle boxes (8 floats each), then the robot half-extents
template <>
struct LaneStage<PlanarView> {
    TabPre<4> ob;
    TabPre<1> rh;
    __host__ __device__ static int floats(const PlanarView& S) { return pad4(8 * S.no) + 2 * S.nr; }
    __device__ void load(const PlanarView& S) {
        ob.load(S.ob32, 8 * S.no);
        rh.load(S.rh32, 2 * S.nr);
    }
    __device__ void store(float* t, const PlanarView& S) const {
        ob.store(t, S.ob32, 8 * S.no);
        rh.store(t + pad4(8 * S.no), S.rh32, 2 * S.nr);
    }
    __device__ static PlanarView redirect(PlanarView S, const float* t) {
        S.rh32 = t + pad4(8 * S.no);
        S.ob32 = t;
        return S;
    }
};

// disk scenes: robot-pair thresholds, sphere centres, robot-sphere thresholds, rectangle corners,
// robot radii squared
template <>
struct LaneStage<DiskView> {
    TabPre<1> rr, lo, hi, rp;
    TabPre<2> sc, sr;
    struct Off { int rr, sc, sr, lo, hi, rp, end; };
    __host__ __device__ static Off offsets(const DiskView& S) {
        Off o;
        o.rr = 0;
        o.sc = pad4(S.nr * S.nr);
        o.sr = o.sc + pad4(S.ns * S.rd);
        o.lo = o.sr + pad4(S.nr * S.ns);
        o.hi = o.lo + pad4(S.nrect * S.rd);
        o.rp = o.hi + pad4(S.nrect * S.rd);
        o.end = o.rp + S.nr;
        return o;
    }
    __host__ __device__ static int floats(const DiskView& S) { return offsets(S).end; }
    __device__ void load(const DiskView& S) {
        rr.load(S.rrT32, S.nr * S.nr);
        sc.load(S.sc32, S.ns * S.rd);
        sr.load(S.srT32, S.nr * S.ns);
        lo.load(S.lo32, S.nrect * S.rd);
        hi.load(S.hi32, S.nrect * S.rd);
        rp.load(S.rp2_32, S.nr);
    }
    __device__ void store(float* t, const DiskView& S) const {
        const Off o = offsets(S);
        rr.store(t + o.rr, S.rrT32, S.nr * S.nr);
        sc.store(t + o.sc, S.sc32, S.ns * S.rd);
        sr.store(t + o.sr, S.srT32, S.nr * S.ns);
        lo.store(t + o.lo, S.lo32, S.nrect * S.rd);
        hi.store(t + o.hi, S.hi32, S.nrect * S.rd);
        rp.store(t + o.rp, S.rp2_32, S.nr);
    }
    __device__ static DiskView redirect(DiskView S, const float* t) {
        const Off o = offsets(S);
        S.rrT32 = t + o.rr;
        S.sc32 = t + o.sc;
        S.srT32 = t + o.sr;
        S.lo32 = t + o.lo;
        S.hi32 = t + o.hi;
        S.rp2_32 = t + o.rp;
        return S;
    }
};

// the scene's tables fit the staged kernel's LDS table
template <class View>
inline bool lane_stage_fits(const View& S) { return LaneStage<View>::floats(S) <= LANE_STAGE_FLOATS; }

template <bool STAGED, class View>
__device__ inline View lane_view(const View& S, const float* t) {
    if constexpr (STAGED) return LaneStage<View>::redirect(S, t);
    else return S;
}

// binary64 interpolation step of one coordinate, (b - a) / (N - 1), as interp_f64 forms it
__device__ inline double interp_step_f64(double a, double b, int N) {
#pragma clang fp contract(off)
    return (b - a) / (double)(N - 1);
}

template <class View> __device__ inline int lane_pose_robots(const View&) { return 1; }
template <> __device__ inline int lane_pose_robots<PlanarView>(const PlanarView& S) { return S.nr; }

// Plain per-call form of the lane-per-state scenes: one workgroup of NT threads (1024 for the
// specialised scenes), the verdict through LDS as in k_arm_edge.
// - the interpolation steps (b - a) / (N - 1) once per coordinate, in LDS;
// - each round takes R states; a state's obstacle tests split over G lanes (a power of two, at
//   most 64, R G <= NT), so that a free state, which runs every test, does not set the launch's
//   latency alone;
// - planar scenes: before the tests, one lane per (state, robot) forms that robot's binary32 pose
//   (binary64 interpolation and sincos) in LDS, once instead of once per lane of the state.
// CONFIG: the row q1 itself is the one state to check (the one-configuration host call), taken as
// is, with no interpolation.
template <class View, int T1, int T2, bool STAGED, int NT, bool CONFIG = false>
__global__ void __launch_bounds__(NT) k_lane_edge(View S0, EdgeBatchView B) {
    constexpr int MAXD = LaneDim<View, T1, T2>::value;
    constexpr bool PLANAR = std::is_same<View, PlanarView>::value;
    __shared__ int s_hit, s_N;
    __shared__ float s_tab[STAGED ? LANE_STAGE_FLOATS : 1];
    __shared__ double s_a[MAXD], s_dir[MAXD];
    __shared__ float4 s_pose[PLANAR ? NT : 1];
    __shared__ unsigned char s_bad[PLANAR ? NT : 1];
    if (threadIdx.x == 0) s_hit = 0;
    LaneStage<View> pre;
    if constexpr (STAGED) pre.load(S0);  // every load of the tables in flight before the rows
    const int D = S0.D;  // <= MAXD <= 48: one wave holds every coordinate
    const int t = (int)threadIdx.x;
    if (t < 64) {
        // wave 0 reads the rows once (lane k: coordinate k; one round trip, also when the rows
        // live in host memory), forms the state count by edge_N_dev's rule (a NaN in coordinate 0
        // sticks, a NaN elsewhere is skipped) and the steps
        const double av = t < D ? row1(B, 0)[t * B.ds] : 0.0;
        const double bv = (!CONFIG && t < D) ? row2(B, 0)[t * B.ds] : 0.0;
        int N = 3;
        if constexpr (!CONFIG) {
            if (B.N || B.N0 > 0) {
                N = B.N ? B.N[0] : B.N0;
            } else {
                const double m = t < D ? __builtin_fabs(av - bv) : 0.0;
                const double m0 = readlane_f64(m, 0);
                const double mx = wave_max_f64(__builtin_isnan(m) ? 0.0 : m);
                N = edge_N_from_max(__builtin_isnan(m0) ? m0 : mx, B.res);
            }
        }
        if (t < D && t < MAXD) {
            s_a[t] = av;
            s_dir[t] = CONFIG ? 0.0 : interp_step_f64(av, bv, N);
        }
        if (t == 0) s_N = N;
    }
    if constexpr (STAGED) pre.store(s_tab, S0);
    __syncthreads();  // s_hit, the state count, the steps and the staged tables
    const int N = s_N;
    const EdgeWin w = CONFIG ? edge_window(3, 0, -1, 0)  // one slot, state 1
                             : edge_window(N, batch_N_start(B, 0), batch_N_max(B, 0), B.incl);
    const View S = lane_view<STAGED>(S0, s_tab);
    unsigned rechecks = 0;
    if (!w.err && w.M > 0) {
        const int nrp = PLANAR ? (T1 ? T1 : lane_pose_robots(S0)) : 1;
        const int R = max(1, min(w.M, NT / nrp));
        int lgG = 0;
        while (lgG < 6 && (R << (lgG + 1)) <= NT) ++lgG;
        const int G = 1 << lgG, part = t & (G - 1), sl = t >> lgG;
        volatile int* hit = &s_hit;
        for (int base = 0; base < w.M; base += R) {
            if (*hit) break;  // the same LDS word for every lane
            if constexpr (PLANAR) {
                if (t < R * nrp) {
                    const int ps = t / nrp, r = t - ps * nrp, local = base + ps;
                    if (local < w.M) {
                        const int i = edge_state_index(w, local, B.incl);
                        if (i >= 0) {
                            const double px = CONFIG ? s_a[3 * r] : interp_dir_f64(s_a[3 * r], s_dir[3 * r], i);
                            const double py = CONFIG ? s_a[3 * r + 1] : interp_dir_f64(s_a[3 * r + 1], s_dir[3 * r + 1], i);
                            const double th = CONFIG ? s_a[3 * r + 2] : interp_dir_f64(s_a[3 * r + 2], s_dir[3 * r + 2], i);
                            const SinCos sc = sincos_f64(th);
                            s_pose[t] = make_float4((float)px, (float)py, (float)sc.c, (float)sc.s);
                            s_bad[t] = !(__builtin_isfinite(px) && __builtin_isfinite(py) && __builtin_isfinite(th));
                        }
                    }
                }
                __syncthreads();
            }
            bool coll = false;
            const int local = base + sl;
            if (sl < R && local < w.M) {
                const int i = edge_state_index(w, local, B.incl);
                if (i >= 0) {
                    double q[MAXD];
#pragma unroll
                    for (int k = 0; k < MAXD; ++k)
                        if (k < D) q[k] = CONFIG ? s_a[k] : interp_dir_f64(s_a[k], s_dir[k], i);
                    if constexpr (PLANAR) {
                        bool bad = false;
                        for (int r = 0; r < nrp; ++r) bad = bad || s_bad[sl * nrp + r];
                        coll = planar_collides<T1, true, true>(S, q, B.mask, rechecks, part, G, s_pose + sl * nrp,
                                                               bad);
                    } else {
                        coll = disk_collides<T1, T2, true>(S, q, B.mask, rechecks, part, G);
                    }
                }
            }
            if (__any(coll)) *hit = 1;
            __syncthreads();
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) B.out[0] = (w.err || s_hit) ? 0 : 1;
}

// binary64 forward kinematics of every link, out[r][L][12]
__global__ void k_arm_fk(ArmView V, const double* qv, long long n, long long rs, long long ds,
                         double* out) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long r = t / V.na;
    const int a = (int)(t % V.na);
    if (r >= n) return;
    double X[12];
    for (int k = 0; k < 12; ++k) X[k] = V.base[12 * a + k];
    for (int k = 0; k < V.dof; ++k) {
        const int l = a * V.dof + k;
        const double q = qv[r * rs + (long long)l * ds];
        if (V.chain) {
            const bool pr = (V.prism >> l) & 1ull;
            SinCos j = sincos_f64(q);
            if (pr) j.s = 0.0, j.c = 1.0;
            chain_step_f64(X, V.dh + 12 * l, j, pr, q);
        } else {
            const SinCos j = sincos_f64(q);
            const double* p = V.dh + 4 * l;
            dh_step_f64(X, p[0], p[1], p[2], p[3], j);
        }
        for (int c = 0; c < 12; ++c) out[(r * V.L + l) * 12 + c] = X[c];
    }
}

__global__ void k_sincos(const double* x, long long n, double* s, double* c) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    SinCos v = sincos_f64(x[t]);
    s[t] = v.s;
    c[t] = v.c;
}

// ============================================================================================
// launchers
static int grid_for(long long n, int block) {
    long long g = (n + block - 1) / block;
    if (g > (1 << 30)) g = 1 << 30;
    return (int)(g < 1 ? 1 : g);
}

void launch_plan_build(const SceneRec& sc, const EdgeBatchView& B, const PlanWork& W, int lgG,
                       hipStream_t st) {
    const int want_dir = 1;
    k_plan_init<<<64, 256, 0, st>>>(W);
    // 4 lanes per edge (fewer for D < 4): measured faster than one group of >= D lanes, whose
    // per-edge tail runs on one lane of a wide group; a few edges per group, so each block's LDS
    // histogram covers many edges
    const int lg_lpe = sc.D > 32 ? 3 : (sc.D >= 4 ? 2 : (sc.D >= 2 ? 1 : 0));
    // D <= MRCC_MAX_D = 64 = PREP_MAXC << 3, so every coordinate has a register in k_plan_prep
    int gp = grid_for(B.n, 4 * (256 >> lg_lpe));
    if (gp > 1024) gp = 1024;
    int gs = grid_for(B.n, 256 * SCATTER_EPT);
    if (gs > 1024) gs = 1024;
    k_plan_prep<<<gp, 256, 0, st>>>(B, W, sc.D, lgG, want_dir, lg_lpe, sc.kind == KIND_ARMS ? sc.arm.nm : 0);
    k_plan_scan<<<1, 1024, 0, st>>>(W, lgG);
    k_plan_scatter<<<gs, 256, 0, st>>>(W, B.n);
}

void launch_plan_run(const SceneRec& sc, const EdgeBatchView& B, const PlanWork& W, int lgG,
                     int spt, int grid, hipStream_t st) {
    switch (sc.kind) {
        case KIND_DISKS:
            if (sc.disk.rd == 2 && sc.disk.nr == 2)
                k_lane_persistent<DiskView, 2, 2><<<grid, 256, 0, st>>>(sc.disk, B, W, lgG, spt);
            else
                k_lane_persistent<DiskView, 0, 0><<<grid, 256, 0, st>>>(sc.disk, B, W, lgG, spt);
            break;
        case KIND_PLANAR:
            if (sc.planar.nr == 4)
                k_lane_persistent<PlanarView, 4, 0><<<grid, 256, 0, st>>>(sc.planar, B, W, lgG, spt);
            else
                k_lane_persistent<PlanarView, 0, 0><<<grid, 256, 0, st>>>(sc.planar, B, W, lgG, spt);
            break;
        case KIND_ARMS: {
            // grid counts waves; workgroups of arm_nw_pers waves, no more of them than can be resident
            const int nw = sc.arm_nw_pers;
            int blocks = (grid + nw - 1) / nw;
            if (blocks > sc.arm_blocks_pers) blocks = sc.arm_blocks_pers;
            const bool stats = W.stats != nullptr;
            ArmQueue Q;
            Q.q1 = B.q1;
            Q.idx1 = B.idx1;
            Q.modes = B.modes;
            Q.out = B.out;
            Q.erec = W.erec;
            Q.order = W.order;
            Q.pass_start = W.pass_start;
            Q.ticket = W.ticket;
            Q.dir = W.dir;
            Q.stats = W.stats;
            Q.meta = W.meta;
            Q.rs = B.rs;
            Q.ds = (int)B.ds;
            Q.n = (int)B.n;
            Q.md0 = B.md0;
            Q.incl = B.incl;
            Q.mask = B.mask;
#define MRCC_ARM_PERSISTENT(G, C)                                                                     \
    do {                                                                                          \
        if (stats) k_arm_persistent<G, C, true><<<blocks, 64 * nw, sc.arm_lds_pers, st>>>(sc.armk, Q, lgG, spt);  \
        else k_arm_persistent<G, C, false><<<blocks, 64 * nw, sc.arm_lds_pers, st>>>(sc.armk, Q, lgG, spt);      \
    } while (0)
            if (sc.arm.chain) {
                if (sc.arm_has_grasp) MRCC_ARM_PERSISTENT(true, true);
                else MRCC_ARM_PERSISTENT(false, true);
            } else if (sc.arm_has_grasp) {
                MRCC_ARM_PERSISTENT(true, false);
            } else {
                MRCC_ARM_PERSISTENT(false, false);
            }
#undef MRCC_ARM_PERSISTENT
            break;
        }
    }
}

__global__ void k_fill_u8(unsigned char* p, long long n, unsigned char v) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

void launch_fill_u8(unsigned char* p, long long n, unsigned char v, hipStream_t st) {
    k_fill_u8<<<grid_for(n, 256), 256, 0, st>>>(p, n, v);
}

void launch_direct(const SceneRec& sc, const EdgeBatchView& B, const int* list, long long n_list,
                   const long long* n_dev, int wpe, int init_out, int max_grid,
                   unsigned long long* stats, hipStream_t st) {
    long long blocks = n_dev ? (long long)max_grid : n_list * wpe;
    int grid = (int)(blocks < max_grid ? blocks : max_grid);
    if (grid < 1) grid = 1;
    switch (sc.kind) {
        case KIND_DISKS:
            if (sc.disk.rd == 2 && sc.disk.nr == 2)
                k_lane_direct<DiskView, 2, 2><<<grid, 64, 0, st>>>(sc.disk, B, list, n_list, n_dev, wpe, init_out, stats);
            else
                k_lane_direct<DiskView, 0, 0><<<grid, 64, 0, st>>>(sc.disk, B, list, n_list, n_dev, wpe, init_out, stats);
            break;
        case KIND_PLANAR:
            if (sc.planar.nr == 4)
                k_lane_direct<PlanarView, 4, 0><<<grid, 64, 0, st>>>(sc.planar, B, list, n_list, n_dev, wpe, init_out, stats);
            else
                k_lane_direct<PlanarView, 0, 0><<<grid, 64, 0, st>>>(sc.planar, B, list, n_list, n_dev, wpe, init_out, stats);
            break;
        case KIND_ARMS: {
            // `grid` waves (at most one per work item), in workgroups of arm_nw waves, or of
            // one wave for a launch of fewer waves than that
            const int nw = grid < sc.arm_nw ? 1 : sc.arm_nw;
            const size_t lds = nw == 1 ? sc.arm_lds1 : sc.arm_lds;
            const int blocks = (grid + nw - 1) / nw;
#define MRCC_ARM_DIRECT(G, C) k_arm_direct<G, C><<<blocks, 64 * nw, lds, st>>>(sc.armk, B, list, n_list, n_dev, wpe, init_out, stats)
            if (sc.arm.chain) {
                if (sc.arm_has_grasp) MRCC_ARM_DIRECT(true, true);
                else MRCC_ARM_DIRECT(false, true);
            } else if (sc.arm_has_grasp) {
                MRCC_ARM_DIRECT(true, false);
            } else {
                MRCC_ARM_DIRECT(false, false);
            }
#undef MRCC_ARM_DIRECT
            break;
        }
    }
}

void launch_configs(const SceneRec& sc, const double* q, long long n, long long rs, long long ds,
                    const int* modes, unsigned mask, unsigned char* out, int max_grid,
                    hipStream_t st) {
    switch (sc.kind) {
        case KIND_DISKS: {
            int g = grid_for(n, 256);
            if (g > max_grid) g = max_grid;
            if (sc.disk.rd == 2 && sc.disk.nr == 2)
                k_lane_configs<DiskView, 2, 2><<<g, 256, 0, st>>>(sc.disk, q, n, rs, ds, mask, out);
            else
                k_lane_configs<DiskView, 0, 0><<<g, 256, 0, st>>>(sc.disk, q, n, rs, ds, mask, out);
            break;
        }
        case KIND_PLANAR: {
            int g = grid_for(n, 256);
            if (g > max_grid) g = max_grid;
            if (sc.planar.nr == 4)
                k_lane_configs<PlanarView, 4, 0><<<g, 256, 0, st>>>(sc.planar, q, n, rs, ds, mask, out);
            else
                k_lane_configs<PlanarView, 0, 0><<<g, 256, 0, st>>>(sc.planar, q, n, rs, ds, mask, out);
            break;
        }
        case KIND_ARMS: {
            // one wave per configuration up to 4 * max_grid waves
            long long g = n < (long long)max_grid * 4 ? n : (long long)max_grid * 4;
            if (g < 1) g = 1;
            // a full workgroup also for fewer configurations than waves: the extra waves only help
            // stage the scene (one round trip of four loads per thread instead of several)
            const int nw = sc.arm_nw;
            const size_t lds = sc.arm_lds;
            const int blocks = (int)((g + nw - 1) / nw);
#define MRCC_ARM_CONFIGS(G, C) k_arm_configs<G, C><<<blocks, 64 * nw, lds, st>>>(sc.armk, q, n, rs, ds, modes, mask, out)
            if (sc.arm.chain) {
                if (sc.arm_has_grasp) MRCC_ARM_CONFIGS(true, true);
                else MRCC_ARM_CONFIGS(false, true);
            } else if (sc.arm_has_grasp) {
                MRCC_ARM_CONFIGS(true, false);
            } else {
                MRCC_ARM_CONFIGS(false, false);
            }
#undef MRCC_ARM_CONFIGS
            break;
        }
    }
}

void launch_arm_edge_spread(const SceneRec& sc, const EdgeBatchView& B, int blocks, hipStream_t st) {
    const int nw = sc.arm_nw_edge;
    const size_t lds = sc.arm_lds_edge;
#define MRCC_ARM_EDGE(G, C) k_arm_edge<G, C><<<blocks, 64 * nw, lds, st>>>(sc.armk, B)
    if (sc.arm.chain) {
        if (sc.arm_has_grasp) MRCC_ARM_EDGE(true, true);
        else MRCC_ARM_EDGE(false, true);
    } else if (sc.arm_has_grasp) {
        MRCC_ARM_EDGE(true, false);
    } else {
        MRCC_ARM_EDGE(false, false);
    }
#undef MRCC_ARM_EDGE
}

// The per-call lane kernel for edge 0 of B (CONFIG: row q1 itself as the one state).  Scenes of
// at most LANE_SMALL_SCENE obstacles run one 256-thread workgroup: their states are quick, and
// 1024 threads (16 waves to start, larger barriers) cost more than the split saves there.
#define LANE_SMALL_SCENE 8

template <bool CONFIG>
static void launch_lane1(const SceneRec& sc, const EdgeBatchView& B, hipStream_t st) {
#define MRCC_LANE1(VIEW, T1, T2, STG, NT, S) k_lane_edge<VIEW, T1, T2, STG, NT, CONFIG><<<1, NT, 0, st>>>(S, B)
    if (sc.kind == KIND_DISKS) {
        const bool stg = lane_stage_fits(sc.disk);
        const bool small = sc.disk.ns + sc.disk.nrect <= LANE_SMALL_SCENE;
        if (sc.disk.rd == 2 && sc.disk.nr == 2) {
            if (small) {
                if (stg) MRCC_LANE1(DiskView, 2, 2, true, 256, sc.disk);
                else MRCC_LANE1(DiskView, 2, 2, false, 256, sc.disk);
            } else if (stg) {
                MRCC_LANE1(DiskView, 2, 2, true, 1024, sc.disk);
            } else {
                MRCC_LANE1(DiskView, 2, 2, false, 1024, sc.disk);
            }
        } else if (stg) {
            MRCC_LANE1(DiskView, 0, 0, true, 256, sc.disk);
        } else {
            MRCC_LANE1(DiskView, 0, 0, false, 256, sc.disk);
        }
    } else if (sc.kind == KIND_PLANAR) {
        const bool stg = lane_stage_fits(sc.planar);
        const bool small = sc.planar.no <= LANE_SMALL_SCENE;
        if (sc.planar.nr == 4) {
            if (small) {
                if (stg) MRCC_LANE1(PlanarView, 4, 0, true, 256, sc.planar);
                else MRCC_LANE1(PlanarView, 4, 0, false, 256, sc.planar);
            } else if (stg) {
                MRCC_LANE1(PlanarView, 4, 0, true, 1024, sc.planar);
            } else {
                MRCC_LANE1(PlanarView, 4, 0, false, 1024, sc.planar);
            }
        } else if (stg) {
            MRCC_LANE1(PlanarView, 0, 0, true, 256, sc.planar);
        } else {
            MRCC_LANE1(PlanarView, 0, 0, false, 256, sc.planar);
        }
    }
#undef MRCC_LANE1
}

void launch_config1(const SceneRec& sc, const EdgeBatchView& B, hipStream_t st) { launch_lane1<true>(sc, B, st); }

void launch_edge1(const SceneRec& sc, const EdgeBatchView& B, hipStream_t st) {
    switch (sc.kind) {
        case KIND_DISKS:
        case KIND_PLANAR:
            launch_lane1<false>(sc, B, st);
            break;
        case KIND_ARMS: {
            const int nw = sc.arm_nw_edge;
            const size_t lds = sc.arm_lds_edge;
#define MRCC_ARM_EDGE(G, C) k_arm_edge<G, C><<<1, 64 * nw, lds, st>>>(sc.armk, B)
            if (sc.arm.chain) {
                if (sc.arm_has_grasp) MRCC_ARM_EDGE(true, true);
                else MRCC_ARM_EDGE(false, true);
            } else if (sc.arm_has_grasp) {
                MRCC_ARM_EDGE(true, false);
            } else {
                MRCC_ARM_EDGE(false, false);
            }
#undef MRCC_ARM_EDGE
            break;
        }
    }
}

void launch_arm_fk(const SceneRec& sc, const double* q, long long n, long long rs, long long ds,
                   double* out, hipStream_t st) {
    long long t = n * sc.arm.na;
    k_arm_fk<<<grid_for(t, 128), 128, 0, st>>>(sc.arm, q, n, rs, ds, out);
}

__global__ void k_set_flag(unsigned char* flag) {
    if (threadIdx.x == 0) *flag = 1;
}

void launch_flag(unsigned char* flag, hipStream_t st) { k_set_flag<<<1, 64, 0, st>>>(flag); }

void launch_sincos(const double* x, long long n, double* s, double* c, hipStream_t st) {
    k_sincos<<<grid_for(n, 256), 256, 0, st>>>(x, n, s, c);
}
