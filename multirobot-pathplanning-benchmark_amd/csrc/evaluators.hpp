// evaluators.hpp — per-scene state evaluators of the collision engine (device code).
//
// Disk scenes (the reference's AbstractEnvironment, abstract_env.py:41-109, :255-276) and planar box
// scenes are evaluated one state per lane; arm scenes one state per wavefront (64 lanes share the
// FK and split the pair set).  Every evaluator returns true when the state is in collision.
#pragma once

#include <type_traits>

#include "common.hpp"

using namespace mrcc;

// --------------------------------------------------------------------------------------------
// binary64 predicates with the reference's numpy rounding sequence
__device__ inline double np_sqnorm_f64(const double* v, int n) {
    // np.dot(v, v) (OpenBLAS scalar tail loop): sequential fused chain
    double s = 0.0;
    for (int k = 0; k < n; ++k) s = __builtin_fma(v[k], v[k], s);
    return s;
}

__device__ inline double np_sum_f64(const double* a, int n) {
#pragma clang fp contract(off)
    // np.sum: identity 0 + numpy pairwise_sum (n <= 128)
    double res;
    if (n < 8) {
        res = -0.0;
        for (int i = 0; i < n; ++i) res += a[i];
    } else {
        double r[8];
        int i;
        for (int j = 0; j < 8; ++j) r[j] = a[j];
        for (i = 8; i < n - (n % 8); i += 8)
            for (int j = 0; j < 8; ++j) r[j] += a[i + j];
        res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
        for (; i < n; ++i) res += a[i];
    }
    return 0.0 + res;
}

__device__ inline double np_clip_f64(double x, double lo, double hi) {
    // NaN-propagating _NPY_MIN(_NPY_MAX(x, lo), hi)
    double m = __builtin_isnan(x) ? x : (x > lo ? x : lo);
    return __builtin_isnan(m) ? m : (m < hi ? m : hi);
}

// robot pair: np.linalg.norm(q[i] - q[j]) < r_i + r_j  <=>  sqnorm < T(r_i + r_j)
__device__ __attribute__((noinline)) bool disk_rr_f64(const DiskView& S, const double* q, int i, int j) {
#pragma clang fp contract(off)
    double d[MRCC_DISK_MAX_RD];
    for (int k = 0; k < S.rd; ++k) d[k] = q[i * S.rd + k] - q[j * S.rd + k];
    return np_sqnorm_f64(d, S.rd) < S.rrT64[i * S.nr + j];
}

// Sphere.collides_with_sphere: np.linalg.norm(pos - center) < r_o + r_i
__device__ __attribute__((noinline)) bool disk_rs_f64(const DiskView& S, const double* q, int i, int o) {
#pragma clang fp contract(off)
    double d[MRCC_DISK_MAX_RD];
    for (int k = 0; k < S.rd; ++k) d[k] = S.sc64[o * S.rd + k] - q[i * S.rd + k];
    return np_sqnorm_f64(d, S.rd) < S.srT64[i * S.ns + o];
}

// Rectangle.collides_with_sphere: sum((clip(p, min, max) - p)**2) <= r**2
__device__ __attribute__((noinline)) bool disk_rect_f64(const DiskView& S, const double* q, int i, int o) {
#pragma clang fp contract(off)
    double sq[MRCC_DISK_MAX_RD];
    for (int k = 0; k < S.rd; ++k) {
        const double p = q[i * S.rd + k];
        const double e = np_clip_f64(p, S.lo64[o * S.rd + k], S.hi64[o * S.rd + k]) - p;
        sq[k] = e * e;
    }
    return np_sum_f64(sq, S.rd) <= S.rp2_64[i];
}

// The same three predicates for 2-D disks with the coordinates by value, inlined: a pointer to the
// caller's state vector would keep that vector in scratch memory for the whole kernel, and a call
// needs a stack frame; the 2-D disk kernels have registers to spare.  The operations are those
// above for rd = 2: np.dot's fused chain from 0, and np.sum of two squares (-0.0 + a + b, then
// 0.0 + that).
__device__ __attribute__((always_inline)) inline bool disk_rr2_f64(const DiskView& S, double xi, double yi, double xj,
                                                       double yj, int i, int j) {
#pragma clang fp contract(off)
    const double d0 = xi - xj, d1 = yi - yj;
    return __builtin_fma(d1, d1, __builtin_fma(d0, d0, 0.0)) < S.rrT64[i * S.nr + j];
}

__device__ __attribute__((always_inline)) inline bool disk_rs2_f64(const DiskView& S, double x, double y, int i, int o) {
#pragma clang fp contract(off)
    const double d0 = S.sc64[2 * o] - x, d1 = S.sc64[2 * o + 1] - y;
    return __builtin_fma(d1, d1, __builtin_fma(d0, d0, 0.0)) < S.srT64[i * S.ns + o];
}

__device__ __attribute__((always_inline)) inline bool disk_rect2_f64(const DiskView& S, double x, double y, int i, int o) {
#pragma clang fp contract(off)
    const double e0 = np_clip_f64(x, S.lo64[2 * o], S.hi64[2 * o]) - x;
    const double e1 = np_clip_f64(y, S.lo64[2 * o + 1], S.hi64[2 * o + 1]) - y;
    const double s0 = e0 * e0, s1 = e1 * e1;
    double res = -0.0;
    res += s0;
    res += s1;
    return 0.0 + res <= S.rp2_64[i];
}

// literal binary64 evaluation of two 2-D disks (used for non-finite coordinates)
__device__ __attribute__((always_inline)) inline bool disk_collides22_f64(const DiskView& S, double q0, double q1,
                                                              double q2, double q3, unsigned mask) {
    if ((mask & 3u) == 3u && disk_rr2_f64(S, q0, q1, q2, q3, 0, 1)) return true;
    for (int i = 0; i < 2; ++i) {
        if (!((mask >> i) & 1u)) continue;
        const double x = i ? q2 : q0, y = i ? q3 : q1;
        for (int o = 0; o < S.ns; ++o)
            if (disk_rs2_f64(S, x, y, i, o)) return true;
        for (int o = 0; o < S.nrect; ++o)
            if (disk_rect2_f64(S, x, y, i, o)) return true;
    }
    return false;
}

// literal binary64 evaluation (used for non-finite coordinates)
__device__ __attribute__((noinline)) bool disk_collides_f64(const DiskView& S, const double* q, unsigned mask) {
    for (int i = 0; i < S.nr; ++i)
        for (int j = i + 1; j < S.nr; ++j)
            if (((mask >> i) & 1u) && ((mask >> j) & 1u) && disk_rr_f64(S, q, i, j)) return true;
    for (int i = 0; i < S.nr; ++i) {
        if (!((mask >> i) & 1u)) continue;
        for (int o = 0; o < S.ns; ++o)
            if (disk_rs_f64(S, q, i, o)) return true;
        for (int o = 0; o < S.nrect; ++o)
            if (disk_rect_f64(S, q, i, o)) return true;
    }
    return false;
}

// Margin of the binary32 squared-distance tests of disk scenes.  Coordinates are rounded from
// binary64 (error <= 2^-24 B each), so a squared distance near a threshold R^2 is off by at most
// ~2^-20 rd (B + 2) in binary32; M keeps a factor >= 4 over that bound for B >= 1.
__device__ inline float disk_margin(float B, int rd) {
    float t = 1.0f + B;
    return 0x1p-18f * t * t * (float)(rd > 2 ? rd : 2) * 0.5f;
}

// SPLIT: the state's tests split over G lanes, this lane taking the robot pairs when part == 0
// and the spheres and rectangles part, part + G, ...; without it every test.
template <int RDT, int NRT, bool SPLIT = false>
__device__ inline bool disk_collides(const DiskView& S, const double* q, unsigned mask,
                                     unsigned& rechecks, int part_ = 0, int G_ = 1) {
    const int part = SPLIT ? part_ : 0, G = SPLIT ? G_ : 1;
    constexpr int MAXD = (RDT ? RDT : MRCC_DISK_MAX_RD) * (NRT ? NRT : MRCC_DISK_MAX_ROBOTS);
    const int rd = RDT ? RDT : S.rd;
    const int nr = NRT ? NRT : S.nr;
    const int D = rd * nr;
    float p[MAXD];
    float B = S.Bscene;
    bool finite = true;
#pragma unroll
    for (int k = 0; k < MAXD; ++k) {
        if (k < D) {
            p[k] = (float)q[k];
            B = fmaxf(B, fabsf(p[k]));
            finite = finite && __builtin_isfinite(q[k]);
        }
    }
    constexpr bool by_value = RDT == 2 && NRT == 2;  // rechecks without a pointer to q
    if (!finite) {
        if constexpr (by_value) return disk_collides22_f64(S, q[0], q[1], q[2], q[3], mask);
        else return disk_collides_f64(S, q, mask);
    }
    const float M = disk_margin(B, rd);
#pragma unroll
    for (int i = 0; i < (NRT ? NRT : MRCC_DISK_MAX_ROBOTS); ++i) {
        if (i >= nr || part != 0) break;
#pragma unroll
        for (int j = i + 1; j < (NRT ? NRT : MRCC_DISK_MAX_ROBOTS); ++j) {
            if (j >= nr) break;
            if (!((mask >> i) & 1u) || !((mask >> j) & 1u)) continue;
            float d2 = 0.0f;
#pragma unroll
            for (int k = 0; k < (RDT ? RDT : MRCC_DISK_MAX_RD); ++k) {
                if (k >= rd) break;
                float d = p[i * rd + k] - p[j * rd + k];
                d2 = fmaf(d, d, d2);
            }
            const float T = S.rrT32[i * nr + j];
            if (d2 < T + M) {
                if (d2 < T - M) return true;
                ++rechecks;
                if constexpr (by_value) {
                    if (disk_rr2_f64(S, q[0], q[1], q[2], q[3], 0, 1)) return true;
                } else if (disk_rr_f64(S, q, i, j)) {
                    return true;
                }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < (NRT ? NRT : MRCC_DISK_MAX_ROBOTS); ++i) {
        if (i >= nr) break;
        if (!((mask >> i) & 1u)) continue;
        for (int o = part; o < S.ns; o += G) {
            float d2 = 0.0f;
#pragma unroll
            for (int k = 0; k < (RDT ? RDT : MRCC_DISK_MAX_RD); ++k) {
                if (k >= rd) break;
                float d = S.sc32[o * rd + k] - p[i * rd + k];
                d2 = fmaf(d, d, d2);
            }
            const float T = S.srT32[i * S.ns + o];
            if (d2 < T + M) {
                if (d2 < T - M) return true;
                ++rechecks;
                if constexpr (by_value) {
                    if (disk_rs2_f64(S, q[2 * i], q[2 * i + 1], i, o)) return true;
                } else if (disk_rs_f64(S, q, i, o)) {
                    return true;
                }
            }
        }
        const float T = S.rp2_32[i];
        for (int o = part; o < S.nrect; o += G) {
            float d2 = 0.0f;
#pragma unroll
            for (int k = 0; k < (RDT ? RDT : MRCC_DISK_MAX_RD); ++k) {
                if (k >= rd) break;
                const float pk = p[i * rd + k];
                const float e = fminf(fmaxf(pk, S.lo32[o * rd + k]), S.hi32[o * rd + k]) - pk;
                d2 = fmaf(e, e, d2);
            }
            if (d2 <= T + M) {
                if (d2 < T - M) return true;
                ++rechecks;
                if constexpr (by_value) {
                    if (disk_rect2_f64(S, q[2 * i], q[2 * i + 1], i, o)) return true;
                } else if (disk_rect_f64(S, q, i, o)) {
                    return true;
                }
            }
        }
    }
    return false;
}

// --------------------------------------------------------------------------------------------
// planar oriented boxes: closed 2D separating-axis test.  Box = [x y ux uy vx vy h0 h1].
// inlined into its callers, the rechecks below (noinline themselves), so that their box arrays
// stay in registers
__device__ __attribute__((always_inline)) inline bool obb2_overlap_f64(const double* A, const double* B) {
#pragma clang fp contract(off)
    double tx = B[0] - A[0], ty = B[1] - A[1];
    double C = A[2] * B[2] + A[3] * B[3];
    double Sv = A[4] * B[2] + A[5] * B[3];
    double aC = __builtin_fabs(C), aS = __builtin_fabs(Sv);
    double tA0 = tx * A[2] + ty * A[3];
    double tA1 = tx * A[4] + ty * A[5];
    double tB0 = tx * B[2] + ty * B[3];
    double tB1 = tx * B[4] + ty * B[5];
    double s0 = __builtin_fabs(tA0) - ((A[6] + B[6] * aC) + B[7] * aS);
    double s1 = __builtin_fabs(tA1) - ((A[7] + B[6] * aS) + B[7] * aC);
    double s2 = __builtin_fabs(tB0) - ((B[6] + A[6] * aC) + A[7] * aS);
    double s3 = __builtin_fabs(tB1) - ((B[7] + A[6] * aS) + A[7] * aC);
    return s0 <= 0.0 && s1 <= 0.0 && s2 <= 0.0 && s3 <= 0.0;
}

// binary64 robot box [x y ux uy vx vy h0 h1] of robot i at pose (x, y, theta)
__device__ inline void planar_robot_box_f64(const PlanarView& S, double x, double y, double th, int i,
                                            double* A) {
    SinCos t = sincos_f64(th);
    A[0] = x;
    A[1] = y;
    A[2] = t.c; A[3] = t.s;
    A[4] = -t.s; A[5] = t.c;
    A[6] = S.rh64[2 * i];
    A[7] = S.rh64[2 * i + 1];
}

// The rechecks take the poses by value: a pointer to the caller's state vector would keep that
// vector in scratch memory for the whole kernel.
__device__ __attribute__((noinline)) bool planar_rr_f64(const PlanarView& S, double xi, double yi, double ti,
                                                        int i, double xj, double yj, double tj, int j) {
    double A[8], B[8];
    planar_robot_box_f64(S, xi, yi, ti, i, A);
    planar_robot_box_f64(S, xj, yj, tj, j, B);
    return obb2_overlap_f64(A, B);
}

__device__ __attribute__((noinline)) bool planar_ro_f64(const PlanarView& S, double x, double y, double th,
                                                        int i, int o) {
    double A[8];
    planar_robot_box_f64(S, x, y, th, i, A);
    return obb2_overlap_f64(A, S.ob64 + 8 * o);
}

// binary32 separation max over the 4 axes (> 0: separated)
__device__ inline float obb2_sep_f32(float ax, float ay, float ac, float as, float ah0, float ah1,
                                     float bx, float by, float bc, float bs, float bh0, float bh1) {
    // A axes u=(ac,as), v=(-as,ac); B axes likewise
    const float tx = bx - ax, ty = by - ay;
    const float C = ac * bc + as * bs;
    const float Sv = ac * bs - as * bc;  // va.ub = -as*bc + ac*bs
    const float aC = fabsf(C), aS = fabsf(Sv);
    const float tA0 = tx * ac + ty * as;
    const float tA1 = ty * ac - tx * as;
    const float tB0 = tx * bc + ty * bs;
    const float tB1 = ty * bc - tx * bs;
    float s = fabsf(tA0) - (ah0 + bh0 * aC + bh1 * aS);
    s = fmaxf(s, fabsf(tA1) - (ah1 + bh0 * aS + bh1 * aC));
    s = fmaxf(s, fabsf(tB0) - (bh0 + ah0 * aC + ah1 * aS));
    s = fmaxf(s, fabsf(tB1) - (bh1 + ah0 * aS + ah1 * aC));
    return s;
}

// Margin of the binary32 axis values: positions and axes are rounded from binary64, so each axis
// value is off by ~2^-22 (B + H); M = 2^-18 (1 + B) keeps a factor >= 8.
__device__ inline float planar_margin(float B) { return 0x1p-18f * (1.0f + B); }

// One planar state: 3 binary64 coordinates per robot (x, y, yaw) in q; non-finite coordinates are
// reported as collisions.
// SPLIT: the state's tests split over G lanes, this lane taking the robot pairs when part == 0 and
// the obstacles part, part + G, ...; without it every test.
// POSED: the binary32 poses (x, y, cos, sin of the yaw, each rounded from binary64) come from
// pose[robot], computed once per robot by the caller, and bad says that a coordinate is not
// finite; q serves the binary64 rechecks only.
template <int NRT, bool SPLIT = false, bool POSED = false>
__device__ inline bool planar_collides(const PlanarView& S, const double* q, unsigned mask,
                                       unsigned& rechecks, int part_ = 0, int G_ = 1,
                                       const float4* pose = nullptr, bool bad = false) {
    constexpr int NR_MAX = NRT ? NRT : MRCC_PLANAR_MAX_ROBOTS;
    const int nr = NRT ? NRT : S.nr;
    const int part = SPLIT ? part_ : 0, G = SPLIT ? G_ : 1;
    float x[NR_MAX], y[NR_MAX], c[NR_MAX], s[NR_MAX], h0[NR_MAX], h1[NR_MAX];
    float B = S.Bscene;
    bool finite = true;
#pragma unroll
    for (int i = 0; i < NR_MAX; ++i) {
        if (i >= nr) break;
        if constexpr (POSED) {
            const float4 p = pose[i];
            x[i] = p.x;
            y[i] = p.y;
            c[i] = p.z;
            s[i] = p.w;
        } else {
            finite = finite && __builtin_isfinite(q[3 * i]) && __builtin_isfinite(q[3 * i + 1]) &&
                     __builtin_isfinite(q[3 * i + 2]);
            SinCos t = sincos_f64(q[3 * i + 2]);
            x[i] = (float)q[3 * i];
            y[i] = (float)q[3 * i + 1];
            c[i] = (float)t.c;
            s[i] = (float)t.s;
        }
        h0[i] = S.rh32[2 * i];
        h1[i] = S.rh32[2 * i + 1];
        B = fmaxf(B, fmaxf(fabsf(x[i]), fabsf(y[i])));
    }
    if (POSED ? bad : !finite) return true;
    const float M = planar_margin(B);
#pragma unroll
    for (int i = 0; i < NR_MAX; ++i) {
        if (i >= nr || part != 0) break;
#pragma unroll
        for (int j = i + 1; j < NR_MAX; ++j) {
            if (j >= nr) break;
            if (!((mask >> i) & 1u) || !((mask >> j) & 1u)) continue;
            float v = obb2_sep_f32(x[i], y[i], c[i], s[i], h0[i], h1[i], x[j], y[j], c[j], s[j], h0[j], h1[j]);
            if (v <= M) {
                if (v < -M) return true;
                ++rechecks;
                if (planar_rr_f64(S, q[3 * i], q[3 * i + 1], q[3 * i + 2], i, q[3 * j], q[3 * j + 1],
                                  q[3 * j + 2], j))
                    return true;
            }
        }
    }
    for (int o = part; o < S.no; o += G) {
        const float* ob = S.ob32 + 8 * o;
        const float ox = ob[0], oy = ob[1], oc = ob[2], os = ob[3], oh0 = ob[6], oh1 = ob[7];
        // every robot's separation, then one branch on the least (fminf skips NaN, as v <= M
        // does); the generic form (up to 16 robots) branches per robot, which needs fewer registers
        float v[NR_MAX];
        float vmin = NRT ? __builtin_inff() : -__builtin_inff();
        if constexpr (NRT != 0) {
#pragma unroll
            for (int i = 0; i < NR_MAX; ++i) {
                v[i] = obb2_sep_f32(x[i], y[i], c[i], s[i], h0[i], h1[i], ox, oy, oc, os, oh0, oh1);
                if ((mask >> i) & 1u) vmin = fminf(vmin, v[i]);
            }
        }
        if (vmin <= M) {
#pragma unroll
            for (int i = 0; i < NR_MAX; ++i) {
                if (i >= nr) break;
                if (!((mask >> i) & 1u)) continue;
                if constexpr (NRT == 0)
                    v[i] = obb2_sep_f32(x[i], y[i], c[i], s[i], h0[i], h1[i], ox, oy, oc, os, oh0, oh1);
                if (v[i] <= M) {
                    if (v[i] < -M) return true;
                    ++rechecks;
                    if (planar_ro_f64(S, q[3 * i], q[3 * i + 1], q[3 * i + 2], i, o)) return true;
                }
            }
        }
    }
    return false;
}

// --------------------------------------------------------------------------------------------
// arms: one state per wavefront; workgroups of 1, 2 or 4 waves.
//
// LDS of a workgroup = one static region (the scene, staged once per kernel by the whole
// workgroup) + one dynamic region per wave (the state being evaluated).  Waves advance
// independently: inside the evaluator they synchronise with wave-level barriers only.
// LDS holds the obstacle boxes when they form clusters (at most 16 x 8 boxes), the link-pair list
// of up to ARM_PAIRS_LDS_MAX pairs, and the per-mode arm counts of up to MRCC_MODES_LDS_MAX modes.
// Larger tables stay in global memory, read through separately typed pointers.  Pointers that could point to
// either LDS or global memory would compile to flat accesses, so each array has one home.

#define ARM_PAIRS_LDS_MAX 1024  // link-pair lists up to this length are staged in LDS

struct ArmStaticLayout {
    int o_dh, o_base, o_sphL, o_linkBC, o_box, o_clu, o_sphR, o_sphLink, o_linkR, o_linkFC, o_mact, o_sets, o_pairs;
    int bytes;
};


__host__ __device__ inline int align16(int x) { return (x + 15) & ~15; }

__host__ __device__ inline ArmStaticLayout arm_static_layout(int L, int na, int nb, int S, int NP, int nm,
                                                           int nclu, int chain) {
    ArmStaticLayout o;
    int b = 0;
    o.o_dh = b;      b += 8 * (chain ? 12 : 4) * L;
    o.o_base = b;    b += 8 * 12 * na;
    o.o_sphL = b;    b += 8 * 3 * S;
    o.o_linkBC = b;  b += 8 * 3 * L;
    b = align16(b);
    o.o_box = b;     b += (nclu > 0 ? 32 * nb : 0);  // lo xyz pad, hi xyz pad
    o.o_clu = b;     b += 32 * nclu;
    o.o_sphR = b;    b += 4 * S;
    o.o_sphLink = b; b += 4 * S;
    o.o_linkR = b;   b += 4 * L;
    o.o_linkFC = b;  b += 4 * L;
    o.o_mact = b;    b += (nm <= MRCC_MODES_LDS_MAX ? 4 * nm : 0);
    o.o_sets = b;    b += 4 * (na + 1 + na * (na - 1) / 2);
    b = align16(b);
    o.o_pairs = b;   b += (NP <= ARM_PAIRS_LDS_MAX ? 8 * NP : 0);
    o.bytes = align16(b);
    return o;
}

#define ARM_PEND_MAX 64  // band pairs awaiting their binary64 recheck, per wave
#define MRCC_MAX_SLOTS 32  // mode boxes per mode (held + resting): one bit each in a mask

// Wave region: X [12 L] doubles, mode boxes G64 [12 nslot] doubles (scenes with boxes in some
// mode only), the band-pair queue, one area that holds the joints' kinematics coefficients [8 D]
// doubles and the chain joints' A23 [D] doubles during the kinematics and the sphere table [S] float4 afterwards, link spheres [L]
// float4, three candidate lists [64] int, mode boxes G32 [16 nslot] float (same scenes).  Every
// part is a multiple of 16 bytes.  Mode-box slots: slot a < na is the box held by arm a, slot
// na + j the j-th resting box.
__host__ __device__ inline int arm_sph_bytes(int D, int S) {
    const int fk = (72 * D + 15) & ~15;  // 8 coefficients + A23 per joint
    return 16 * S > fk ? 16 * S : fk;
}

__host__ __device__ inline int arm_wave_bytes(int D, int L, int nslot, int S, int grasp) {
    return 96 * L + (grasp ? 96 * nslot : 0) + 8 * ARM_PEND_MAX + arm_sph_bytes(D, S) + 16 * L + 3 * 4 * 64 +
           (grasp ? 64 * nslot : 0);
}

// dynamic LDS of an arm kernel launched with `nw` waves per workgroup
__host__ __device__ inline size_t arm_lds_bytes(int D, int L, int na, int nb, int S, int NP, int nm,
                                                int nclu, int grasp, int chain, int nw, int nslot) {
    return (size_t)arm_static_layout(L, na, nb, S, NP, nm, nclu, chain).bytes +
           (size_t)nw * arm_wave_bytes(D, L, nslot, S, grasp);
}

struct ArmLds {
    // per wave
    double* fkc;    // [D*8] per joint: cos, sin, and the products of the DH step, then [D] A23 of
                    // the chain joints (kinematics phase; shares memory with sph)
    double* X;
    double* G64;
    float4* sph;
    float4* linkC;  // [L] bounding sphere (x, y, z, R) of each link
    int2* pend;     // [ARM_PEND_MAX] (kind, i << 16 | j): kind 0 sphere i - box j, 1 sphere pair
    int* cand;      // [64] compacted candidates of one broad-phase round
    int* lst;       // [64] links that pass the cull against the box around all obstacles
    int* lbc;       // [64] (link, box) candidates of the two-level link-box test
    float4* abox;   // [2 na] box (lo, hi) around each arm's link spheres (shares memory with lst)
    float* G32;
    // per workgroup (the scene)
    const double* dh;      // joint table: [L*4] d a cos(alpha) sin(alpha) (DH), or [L*12] the fixed
                           // transform C (3 x 4, rows) after Rz(q) of each joint (chains)
    const double* base;    // [na*12]
    const double* sphL;    // [S*3] sphere centres in the link frame
    const double* linkBC;  // [L*3] bounding-sphere centres in the link frame
    const float4* box;     // [nb*2] lo, hi (LDS; scenes with box clusters only)
    const float4* boxG;    // [nb*2] lo, hi (global)
    const float* sphR;     // [S]
    const int* sphLink;    // [S] link | arm << 8
    const float4* clu;     // [nclu*2] box clusters: lo xyz | first box, hi xyz | box count
    const float* linkR;    // [L]
    const int* linkFC;     // [L] first sphere | count << 16
    const int2* pairs;     // [NP] link pairs with their sphere ranges (NP <= ARM_PAIRS_LDS_MAX)
    const int* mact;       // [nm] active arms of each mode (nm <= MRCC_MODES_LDS_MAX)
    const int* sets;       // self-pair counts and arm-pair segments of the pair list
};

// A counter that counts nothing: kernels built without statistics keep no registers for them.
struct NoCount {
    __device__ NoCount& operator+=(unsigned) { return *this; }
    __device__ NoCount& operator++() { return *this; }
};

// tests executed by one lane (broad phase link-box / sphere-box / link-pair / sphere-pair)
template <bool STATS>
struct ArmCountT {
    unsigned lb, sb, lp, sp;
};
template <>
struct ArmCountT<false> {
    NoCount lb, sb, lp, sp;
};
template <bool STATS>
using RecheckT = typename std::conditional<STATS, unsigned, NoCount>::type;

__device__ inline ArmLds arm_lds(void* base, const ArmK& V) {
    ArmLds W;
    unsigned char* b = (unsigned char*)base;
    const ArmStaticLayout o = arm_static_layout(V.L, V.na, V.nb, V.S, V.NP, V.nm, V.nclu, V.chain);
    W.dh = (const double*)(b + o.o_dh);
    W.base = (const double*)(b + o.o_base);
    W.sphL = (const double*)(b + o.o_sphL);
    W.linkBC = (const double*)(b + o.o_linkBC);
    W.box = (const float4*)(b + o.o_box);
    W.boxG = (const float4*)V.box32();
    W.sphR = (const float*)(b + o.o_sphR);
    W.sphLink = (const int*)(b + o.o_sphLink);
    W.clu = (const float4*)(b + o.o_clu);
    W.linkR = (const float*)(b + o.o_linkR);
    W.linkFC = (const int*)(b + o.o_linkFC);
    W.pairs = (const int2*)(b + o.o_pairs);
    W.mact = (const int*)(b + o.o_mact);
    W.sets = (const int*)(b + o.o_sets);
    unsigned char* w = b + o.bytes + (threadIdx.x >> 6) * arm_wave_bytes(V.D, V.L, V.nslot, V.S, V.grasp);
    W.X = (double*)w;
    w += 96 * V.L;
    W.G64 = (double*)w;
    if (V.grasp) w += 96 * V.nslot;
    W.pend = (int2*)w;
    w += 8 * ARM_PEND_MAX;
    W.sph = (float4*)w;
    W.fkc = (double*)w;
    w += arm_sph_bytes(V.D, V.S);
    W.linkC = (float4*)w;
    W.cand = (int*)(W.linkC + V.L);
    W.lst = W.cand + 64;
    W.lbc = W.lst + 64;
    W.abox = (float4*)W.lst;  // 2 MRCC_MAX_ARMS float4 = 64 ints
    W.G32 = (float*)(W.lbc + 64);
    return W;
}

// Stages the scene into the static LDS region: a copy of the scene's LDS image (all threads of the
// workgroup, 16-byte loads, four per thread issued before their stores), then a barrier.
__device__ inline void arm_stage_scene(void* base, const ArmK& V) {
    float4* dst = (float4*)base;
    const float4* src = V.img();
    const int n16 = V.img_bytes >> 4;
    const int nt = blockDim.x;
    const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int i0 = threadIdx.x; i0 < n16; i0 += 4 * nt) {
        const int i1 = i0 + nt, i2 = i1 + nt, i3 = i2 + nt;
        const float4 a = src[i0];
        const float4 b = i1 < n16 ? src[i1] : z;
        const float4 c = i2 < n16 ? src[i2] : z;
        const float4 d = i3 < n16 ? src[i3] : z;
        dst[i0] = a;
        if (i1 < n16) dst[i1] = b;
        if (i2 < n16) dst[i2] = c;
        if (i3 < n16) dst[i3] = d;
    }
    __syncthreads();
}

// The same copy in two halves, for a kernel that computes while the image is in flight: the first
// K blocks of the image (K 16-byte loads per thread, indices clamped into the image, so the loads
// are issued unconditionally) into registers, then the stores, the rest of the image and the
// barrier.
template <int K>
struct ArmStage {
    float4 v[K];
};

template <int K>
__device__ inline ArmStage<K> arm_stage_load(const ArmK& V) {
    ArmStage<K> r;
    const float4* src = V.img();
    const int last = (V.img_bytes >> 4) - 1;  // >= 0: the joint table is never empty
    const int nt = blockDim.x;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i = threadIdx.x + k * nt;
        r.v[k] = src[i < last ? i : last];
    }
    return r;
}

template <int K>
__device__ inline void arm_stage_store(void* base, const ArmK& V, const ArmStage<K>& r) {
    float4* dst = (float4*)base;
    const float4* src = V.img();
    const int n16 = V.img_bytes >> 4;
    const int nt = blockDim.x;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i = threadIdx.x + k * nt;
        if (i < n16) dst[i] = r.v[k];
    }
    // the rest, four loads in flight per thread
    const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int i0 = threadIdx.x + K * nt; i0 < n16; i0 += 4 * nt) {
        const int i1 = i0 + nt, i2 = i1 + nt, i3 = i2 + nt;
        const float4 a = src[i0];
        const float4 b = i1 < n16 ? src[i1] : z;
        const float4 c = i2 < n16 ? src[i2] : z;
        const float4 d = i3 < n16 ? src[i3] : z;
        dst[i0] = a;
        if (i1 < n16) dst[i1] = b;
        if (i2 < n16) dst[i2] = c;
        if (i3 < n16) dst[i3] = d;
    }
    __syncthreads();
}

// LDS written by some lanes of a wave is read by other lanes of the same wave after this point
// (LDS-only fences: outstanding global loads, such as the next state's prefetched inputs, are
// not waited for)
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ inline int wave_lane() { return threadIdx.x & 63; }

// value of lane (lane & ~3) + j of the lane's quad (DPP quad_perm broadcast; all 4 lanes active)
__device__ inline double quad_bcast_f64(double v, int j) {
    const unsigned long long u = __double_as_longlong(v);
    const int lo = (int)(unsigned)u, hi = (int)(unsigned)(u >> 32);
    int rlo, rhi;
    switch (j) {
        case 0:
            rlo = __builtin_amdgcn_mov_dpp(lo, 0x00, 0xf, 0xf, false);
            rhi = __builtin_amdgcn_mov_dpp(hi, 0x00, 0xf, 0xf, false);
            break;
        case 1:
            rlo = __builtin_amdgcn_mov_dpp(lo, 0x55, 0xf, 0xf, false);
            rhi = __builtin_amdgcn_mov_dpp(hi, 0x55, 0xf, 0xf, false);
            break;
        case 2:
            rlo = __builtin_amdgcn_mov_dpp(lo, 0xaa, 0xf, 0xf, false);
            rhi = __builtin_amdgcn_mov_dpp(hi, 0xaa, 0xf, 0xf, false);
            break;
        default:
            rlo = __builtin_amdgcn_mov_dpp(lo, 0xff, 0xf, 0xf, false);
            rhi = __builtin_amdgcn_mov_dpp(hi, 0xff, 0xf, 0xf, false);
            break;
    }
    return __longlong_as_double(((long long)(unsigned)rhi << 32) | (unsigned)rlo);
}

__device__ inline unsigned lanes_below(unsigned long long m) {
    return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

__device__ inline float readlane_f(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

__device__ inline double readlane_f64(double v, int l) {
    const unsigned long long u = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | lo);
}

// v of another lane of the row of 16 by a DPP pattern (all lanes active)
template <int CTRL>
__device__ inline double dpp_f64(double v) {
    const unsigned long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)u, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)(unsigned)hi << 32) | (unsigned)lo);
}

// Largest of the 64 lanes' values, none of them NaN, in every lane: DPP steps within each row of
// 16 lanes (quad swaps, half-row and row mirrors), then the four row maxima.  The order of a max
// does not change its value, so any lane order gives the sequential loop's result.
__device__ inline double wave_max_f64(double m) {
    m = fmax(m, dpp_f64<0xb1>(m));   // quad_perm [1, 0, 3, 2]
    m = fmax(m, dpp_f64<0x4e>(m));   // quad_perm [2, 3, 0, 1]
    m = fmax(m, dpp_f64<0x141>(m));  // row_half_mirror
    m = fmax(m, dpp_f64<0x140>(m));  // row_mirror
    return fmax(fmax(readlane_f64(m, 0), readlane_f64(m, 16)), fmax(readlane_f64(m, 32), readlane_f64(m, 48)));
}

// Band pairs (binary32 test within the margin) are queued in LDS and rechecked in binary64 a
// batch at a time, lane-parallel: the recheck code then runs where little else is live, instead
// of a call inside the candidate loops (whose register cost would apply to the whole kernel).
// Appends the pairs of the lanes with `band` set; `np` is the wave-uniform queue length.
__device__ inline void arm_pend(const ArmLds& W, bool band, int kind, int i, int j, int& np) {
    const unsigned long long m = __ballot(band);
    if (band) W.pend[np + (int)lanes_below(m)] = make_int2(kind, (i << 16) | j);
    np += __popcll(m);
}

// Rechecks the queued pairs in binary64 (same predicates as the oracle); empties the queue.
// Returns the wave-uniform "some pair collides".
template <class RC>
__device__ inline bool arm_flush(const ArmK& V, const ArmLds& W, int& np, RC& rechecks) {
    wave_sync();
    bool hit = false;
    for (int k = wave_lane(); k < np; k += 64) {
        const int2 e = W.pend[k];
        const int i = e.y >> 16, j = e.y & 0xffff;
        double a[3];
        xform_point_f64(W.X + 12 * (W.sphLink[i] & 255), W.sphL[3 * i], W.sphL[3 * i + 1], W.sphL[3 * i + 2], a);
        if (e.x == 0) {
            hit |= sph_box_f64(a, V.sph64()[4 * i + 3], V.box64() + 6 * j, V.box64() + 6 * j + 3);
        } else {
            double b[3];
            xform_point_f64(W.X + 12 * (W.sphLink[j] & 255), W.sphL[3 * j], W.sphL[3 * j + 1], W.sphL[3 * j + 2], b);
            hit |= sph_sph_f64(a, V.sph64()[4 * i + 3], b, V.sph64()[4 * j + 3]);
        }
    }
    if (wave_lane() == 0) rechecks += np;
    np = 0;
    wave_sync();
    return __any(hit);
}

// Rechecks queued mode-box pairs in binary64: kind 2 held box i - obstacle box j, 3 sphere i -
// mode box j, 4 mode boxes i, j.  Empties the queue; returns the wave-uniform "some pair collides".
template <class RC>
__device__ inline bool arm_flush_held(const ArmK& V, const ArmLds& W, int mode, int& np, RC& rechecks) {
    wave_sync();
    bool hit = false;
    for (int k = wave_lane(); k < np; k += 64) {
        const int2 e = W.pend[k];
        const int i = e.y >> 16, j = e.y & 0xffff;
        if (e.x == 2) {
            double A[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
            const double* bc = V.boxch64() + 6 * j;
            A[9] = bc[0];
            A[10] = bc[1];
            A[11] = bc[2];
            hit |= obb3_overlap_f64(A, bc + 3, W.G64 + 12 * i, V.m_half() + 3 * (mode * V.nslot + i));
        } else if (e.x == 3) {
            double c[3];
            xform_point_f64(W.X + 12 * (W.sphLink[i] & 255), W.sphL[3 * i], W.sphL[3 * i + 1], W.sphL[3 * i + 2], c);
            hit |= sph_obb_f64(c, V.sph64()[4 * i + 3], W.G64 + 12 * j, V.m_half() + 3 * (mode * V.nslot + j));
        } else {
            hit |= obb3_overlap_f64(W.G64 + 12 * i, V.m_half() + 3 * (mode * V.nslot + i), W.G64 + 12 * j,
                                    V.m_half() + 3 * (mode * V.nslot + j));
        }
    }
    if (wave_lane() == 0) rechecks += np;
    np = 0;
    wave_sync();
    return __any(hit);
}

// active arms of mode md.  The LDS read goes through an LDS-qualified pointer so that the two
// loads cannot be merged into one flat load through a selected pointer.
typedef __attribute__((address_space(3))) const int lds_cint;
typedef __attribute__((address_space(3))) const unsigned long long lds_cu64;

// link pair pi (LDS table for up to ARM_PAIRS_LDS_MAX pairs)
__device__ inline int2 arm_pair(const ArmK& V, const ArmLds& W, int pi) {
    if (V.NP <= ARM_PAIRS_LDS_MAX) {
        const unsigned long long v = ((lds_cu64*)W.pairs)[pi];
        return make_int2((int)(unsigned)v, (int)(unsigned)(v >> 32));
    }
    return V.lpairs2()[pi];
}
__device__ inline int arm_mact(const ArmK& V, const ArmLds& W, int md) {
    if (V.nm <= MRCC_MODES_LDS_MAX) return ((lds_cint*)W.mact)[md];
    return V.m_active()[md];
}

// Per-lane register loaded once per kernel: the lane's link (first sphere, sphere count).
struct ArmRegs {
    int linkFC;
};

__device__ inline ArmRegs arm_regs(const ArmK& V, const ArmLds& W) {
    ArmRegs R;
    const int lane = wave_lane();
    R.linkFC = lane < V.L ? W.linkFC[lane] : 0;
    return R;
}

__device__ inline float box_d2(float4 c, float4 lo, float4 hi) {
    const float e0 = fminf(fmaxf(c.x, lo.x), hi.x) - c.x;
    const float e1 = fminf(fmaxf(c.y, lo.y), hi.y) - c.y;
    const float e2 = fminf(fmaxf(c.z, lo.z), hi.z) - c.z;
    return fmaf(e2, e2, fmaf(e1, e1, e0 * e0));
}

// Evaluates the state whose binary64 joint vector is distributed over the lanes (lane k < D
// holds coordinate k in `qv`).  All 64 lanes of the wave must call.  Returns the wave-uniform
// result.
// GRASP = false instantiates the evaluator for scenes without held boxes (fewer registers);
// CHAIN selects chain joints (V.chain = 1) over DH joints at compile time.
// `lazy` (wave-uniform): place the spheres beyond the first 64 only when a candidate pair needs
// them (pays when most states collide; the caller decides from its recent states).
// STATS = false compiles the test counters away.  A mode outside [0, nm) is reported as a
// collision (callers reject such modes before they get here).
// LOWLAT = true selects forms that shorten one state's latency at the cost of registers (for a
// kernel with one state per wave in flight: the per-call form): the kinematics load joint
// coefficients three joints ahead, and the first round of sphere centres shares one block of
// loads with the link spheres.
// `have_tq` (wave-uniform): `tq` already holds sincos_f64(qv) of the lane's coordinate (computed by
// the caller, for instance while the scene was being staged); otherwise it is computed here.
template <bool GRASP, bool CHAIN, bool STATS, bool LOWLAT = false>
__device__ inline bool arm_state_collides(const ArmK& V, const ArmLds& W, const ArmRegs& R, double qv,
                                          int mode, unsigned mask, RecheckT<STATS>& rechecks,
                                          ArmCountT<STATS>& cnt, bool lazy = false, bool have_tq = false,
                                          SinCos tq = SinCos{0.0, 1.0}) {
    const int lane = wave_lane();
    if (mode < 0 || mode >= V.nm) return true;
    const int na = arm_mact(V, W, mode);
    const int dof = V.dof;
    const int Dact = na * dof;
    // 1. joint sines / cosines (binary64), finiteness; lane l also forms rows 0 and 1 of joint l's
    //    step A = Rz(q) C, column by column: [A00, A10, A01, A11, A02, A12, A03, A13].  DH joints:
    //    dh_row_f64's products; chain joints: chain_coef_f64 (A0c = cs C0c - sn C1c,
    //    A1c = sn C0c + cs C1c), which gives the DH products exactly for C of a DH joint.
    //    A prismatic chain joint (bit l of V.prism) has A = Tz(q) C: cs = 1, sn = 0 and
    //    A23 = C23 + q; |q| beyond its travel counts as a collision, as a non-finite coordinate.
    bool bad = false;
    if (lane < Dact) {
        bad = !__builtin_isfinite(qv);
        const bool pr = CHAIN && ((V.prism >> lane) & 1ull);
        const SinCos t = have_tq ? tq : sincos_f64(qv);
        const double cs = pr ? 1.0 : t.c, sn = pr ? 0.0 : t.s;
        double2* f = (double2*)(W.fkc + 8 * lane);
        if (CHAIN) {
            const double* C = W.dh + 12 * lane;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const ChainCoef k = chain_coef_f64(cs, sn, C[c], C[4 + c]);
                f[c] = make_double2(k.a0, k.a1);
            }
            W.fkc[8 * V.D + lane] = chain_a23_f64(C, pr, qv);
            if (pr) bad = bad || fabs(qv) > V.travel()[lane];
        } else {
            const double* p = W.dh + 4 * lane;  // d a cos(alpha) sin(alpha)
            const double a = p[1], ca = p[2], sa = p[3];
            f[0] = make_double2(cs, sn);
            f[1] = make_double2(-(sn * ca), cs * ca);
            f[2] = make_double2(sn * sa, -(cs * sa));
            f[3] = make_double2(a * cs, a * sn);
        }
    }
    if (__any(bad)) return true;  // non-finite coordinates are reported as collisions
    wave_sync();
    // 2. forward kinematics (binary64).  Lane (arm a, row r, column c) computes entry c of row r
    //    of the arm's transform at every joint, with the operations of dh_row_f64:
    //      c < 3: (R0 * k0 + R1 * k1) + R2 * k2,   c = 3: ((R0 * k0 + R1 * k1) + R2 * k2) + t
    //    with (k0, k1) the joint's coefficient pair c and k2 = 0, sin(alpha), cos(alpha), d; the
    //    quad then shares the row's four new entries (DPP broadcasts within the quad).
    //    Arms are processed 5 at a time (60 lanes).  Grasped boxes are composed from the last row.
    {
        const int c = lane & 3, rr = (lane >> 2) % 3, ag = lane / 12;
        const int k2i = c == 1 ? 3 : (c == 2 ? 2 : 0);  // sa, ca, d in the DH table
        // k2 = A2c: 0, sin(alpha), cos(alpha), d of a DH joint; C2c of a chain joint
        const bool chain = CHAIN;
        for (int a0 = 0; a0 < na; a0 += 5) {
            const int a = a0 + ag;
            if (ag < 5 && a < na) {
                double R0 = W.base[12 * a + 3 * rr], R1 = W.base[12 * a + 3 * rr + 1],
                       R2 = W.base[12 * a + 3 * rr + 2], t = W.base[12 * a + 9 + rr];
                const int l0 = a * dof;
                // chain joints: A2c = C2c for c < 3, A23 from the kinematics coefficients; the
                // coefficients of later joints are loaded while a joint is stepped (an LDS round
                // trip is longer than a step)
                const double* a2 = c == 3 ? W.fkc + 8 * V.D : W.dh + 8 + c;
                const int a2s = c == 3 ? 1 : 12;
                // one joint step: entry c of row rr, then the row's four new entries from the quad
                auto step = [&](int l, double k0, double k1, double kz) {
                    double y;
                    {
#pragma clang fp contract(off)
                        y = (R0 * k0 + R1 * k1) + R2 * kz;
                        const double yt = y + t;
                        y = c == 3 ? yt : y;
                    }
                    W.X[12 * l + (c < 3 ? 3 * rr + c : 9 + rr)] = y;
                    R0 = quad_bcast_f64(y, 0);
                    R1 = quad_bcast_f64(y, 1);
                    R2 = quad_bcast_f64(y, 2);
                    t = quad_bcast_f64(y, 3);
                };
                auto coef = [&](int l) { return ((const double2*)(W.fkc + 8 * l))[c]; };
                // loaded by every lane (no lane-masked load: k2i = 0 reads d) and replaced by 0
                // for column 0 of a DH joint where it is used, so a load is waited for only there
                auto coef2 = [&](int l) { return chain ? a2[a2s * l] : W.dh[4 * l + k2i]; };
                const bool k2zero = !chain && c == 0;
                if constexpr (LOWLAT) {
                    constexpr int AH = 3;  // joints ahead
                    double2 KK[MRCC_MAX_DOF];
                    double K2[MRCC_MAX_DOF];
#pragma unroll
                    for (int k = 0; k < AH; ++k) {
                        if (k >= dof) break;  // wave-uniform
                        KK[k] = coef(l0 + k);
                        K2[k] = coef2(l0 + k);
                    }
#pragma unroll
                    for (int k = 0; k < MRCC_MAX_DOF; ++k) {
                        if (k >= dof) break;  // wave-uniform
                        if (k + AH < MRCC_MAX_DOF && k + AH < dof) {
                            KK[k + AH] = coef(l0 + k + AH);
                            K2[k + AH] = coef2(l0 + k + AH);
                        }
                        step(l0 + k, KK[k].x, KK[k].y, k2zero ? 0.0 : K2[k]);
                    }
                } else {
                    // joint k + 1's coefficients are loaded while joint k is stepped
                    double2 kk = coef(l0);
                    double k2 = coef2(l0);
                    for (int k = 0; k < dof; ++k) {
                        const double k0 = kk.x, k1 = kk.y, kz = k2zero ? 0.0 : k2;
                        if (k + 1 < dof) {
                            kk = coef(l0 + k + 1);
                            k2 = coef2(l0 + k + 1);
                        }
                        step(l0 + k, k0, k1, kz);
                    }
                }
                if (GRASP && c == 0 && V.m_grasp()[mode * V.nslot + a]) {
                    const double* Rel = V.m_rel() + 12 * (mode * V.nslot + a);
                    double g[4];
                    compose_row_f64(R0, R1, R2, t, Rel, g);
                    for (int cc = 0; cc < 3; ++cc) {
                        W.G64[12 * a + 3 * rr + cc] = g[cc];
                        W.G32[16 * a + 3 * rr + cc] = (float)g[cc];
                    }
                    W.G64[12 * a + 9 + rr] = g[3];
                    W.G32[16 * a + 9 + rr] = (float)g[3];
                    W.G32[16 * a + 12 + rr] = (float)V.m_half()[3 * (mode * V.nslot + a) + rr];
                }
            }
        }
    }
    if (GRASP && V.nslot > V.na) {
        // resting boxes of the mode: their world pose is the slot's transform
        for (int e = lane; e < 12 * (V.nslot - V.na); e += 64) {
            const int sl = V.na + e / 12, k = e % 12;
            if (!V.m_grasp()[mode * V.nslot + sl]) continue;
            const double v = V.m_rel()[12 * (mode * V.nslot + sl) + k];
            W.G64[12 * sl + k] = v;
            W.G32[16 * sl + k] = (float)v;
            if (k < 3) W.G32[16 * sl + 12 + k] = (float)V.m_half()[3 * (mode * V.nslot + sl) + k];
        }
    }
    wave_sync();
    // 3. sphere centres and link bounding spheres: binary64, rounded to binary32.  Lazily (link
    //    broad phase), only the first 64 spheres are placed here (they hold every sphere of arm 0:
    //    at most 8 links of 8); the others are placed when the first candidate pair that reads
    //    them comes up (arm 0's candidates come first, and most random states collide within one
    //    arm), or at the end of the self pairs.
    const int lastFC = __builtin_amdgcn_readlane(R.linkFC, Dact - 1);
    const int S_act = (lastFC & 0xffff) + (lastFC >> 16);
    const int S_first = lazy && V.broadphase && S_act > 64 ? 64 : S_act;
    float4 myLink = make_float4(0.f, 0.f, 0.f, 0.f);
    if (LOWLAT) {
        // the first round of spheres and the link spheres in one block: all their loads are
        // issued before any store
        const bool hs = lane < S_first, hl = V.broadphase && lane < Dact;
        const int s = hs ? lane : 0, ll = hl ? lane : 0;
        const double* l = W.sphL + 3 * s;
        const double* b = W.linkBC + 3 * ll;
        const double l0 = l[0], l1 = l[1], l2 = l[2], b0 = b[0], b1 = b[1], b2 = b[2];
        const float rs = W.sphR[s], rl = W.linkR[ll];
        double c[3], e[3];
        xform_point_f64(W.X + 12 * (W.sphLink[s] & 255), l0, l1, l2, c);
        xform_point_f64(W.X + 12 * ll, b0, b1, b2, e);
        if (hs) W.sph[s] = make_float4((float)c[0], (float)c[1], (float)c[2], rs);
        if (hl) {
            myLink = make_float4((float)e[0], (float)e[1], (float)e[2], rl);
            W.linkC[ll] = myLink;
        }
    }
    for (int s = (LOWLAT ? 64 : 0) + lane; s < S_first; s += 64) {
        double c[3];
        const double* l = W.sphL + 3 * s;
        xform_point_f64(W.X + 12 * (W.sphLink[s] & 255), l[0], l[1], l[2], c);
        W.sph[s] = make_float4((float)c[0], (float)c[1], (float)c[2], W.sphR[s]);
    }
    if (!LOWLAT && V.broadphase && lane < Dact) {
        const double* b = W.linkBC + 3 * lane;
        double c[3];
        xform_point_f64(W.X + 12 * lane, b[0], b[1], b[2], c);
        myLink = make_float4((float)c[0], (float)c[1], (float)c[2], W.linkR[lane]);
        W.linkC[lane] = myLink;
    }
    wave_sync();
    const float M = V.Msph;
    bool hit = false;
    int np = 0;  // queued band pairs
    const unsigned amask = na >= 32 ? ~0u : ((1u << na) - 1u);
    const bool all_arms = (mask & amask) == amask;  // no pair is masked out
    if (V.broadphase) {
        const float Mc = V.Mcull;
        // 4'. link pairs of the pair set, then the sphere pairs of candidate link pairs (8 x 8);
        //     first, since self and inter-arm contacts dominate collisions of random states.
        //     The self pairs of the active arms are walked first (most collisions of random states
        //     are self collisions); then the segment of an arm pair, only when the boxes around
        //     the two arms' link spheres are not separated by Marm.
        int p0 = 0, p_end = W.sets[na];
        bool placed = S_first == S_act;  // every sphere placed
        auto place_rest = [&]() {
            placed = true;
            for (int s = 64 + lane; s < S_act; s += 64) {
                double c[3];
                const double* l = W.sphL + 3 * s;
                xform_point_f64(W.X + 12 * (W.sphLink[s] & 255), l[0], l[1], l[2], c);
                W.sph[s] = make_float4((float)c[0], (float)c[1], (float)c[2], W.sphR[s]);
            }
            wave_sync();
        };
        int sg = 0;
        bool segs_known = false;
        unsigned long long segs = 0ull;
        for (;;) {
            if (p0 >= p_end) {
                if (!placed) place_rest();
                if (!segs_known) {
                    segs_known = true;
                    if (lane < na) {
                        float4 lo = make_float4(INFINITY, INFINITY, INFINITY, 0.f);
                        float4 hi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.f);
#pragma unroll
                        for (int k = 0; k < MRCC_MAX_DOF; ++k) {
                            if (k >= dof) break;  // wave-uniform
                            // selects, not a lane-masked block
                            const int l = lane * dof + k;
                            const float4 c = W.linkC[l];
                            const bool use = (W.linkFC[l] >> 16) > 0;
                            lo.x = use ? fminf(lo.x, c.x - c.w) : lo.x;
                            lo.y = use ? fminf(lo.y, c.y - c.w) : lo.y;
                            lo.z = use ? fminf(lo.z, c.z - c.w) : lo.z;
                            hi.x = use ? fmaxf(hi.x, c.x + c.w) : hi.x;
                            hi.y = use ? fmaxf(hi.y, c.y + c.w) : hi.y;
                            hi.z = use ? fmaxf(hi.z, c.z + c.w) : hi.z;
                        }
                        W.abox[2 * lane] = lo;
                        W.abox[2 * lane + 1] = hi;
                    }
                    wave_sync();
                    bool keep = false;
                    if (lane < (na * (na - 1)) >> 1) {
                        sg = W.sets[V.na + 1 + lane];
                        const int a = (sg >> 20) & 15, b = (sg >> 24) & 15;
                        if (((sg >> 12) & 255) > 0 && (all_arms || (((mask >> a) & 1u) && ((mask >> b) & 1u)))) {
                            const float4 la = W.abox[2 * a], ha = W.abox[2 * a + 1];
                            const float4 lb = W.abox[2 * b], hb = W.abox[2 * b + 1];
                            const float d = V.Marm;
                            keep = !(lb.x - ha.x > d || la.x - hb.x > d || lb.y - ha.y > d || la.y - hb.y > d ||
                                     lb.z - ha.z > d || la.z - hb.z > d);
                            ++cnt.lp;
                        }
                    }
                    segs = __ballot(keep);
                }
                if (segs == 0ull) break;
                const int k = __builtin_ctzll(segs);
                segs &= segs - 1ull;
                const int sk = __builtin_amdgcn_readlane(sg, k);
                p0 = sk & 4095;
                p_end = p0 + ((sk >> 12) & 255);
            }
            // branch-free: lanes past the round's pairs read its first pair, outcome masked
            const int pi = p0 + lane;
            const bool pin = pi < p_end;
            const int2 pr = arm_pair(V, W, pin ? pi : p0);  // p0 < p_end here
            p0 += 64;
            const int l = (pr.x >> 8) & 255, mm = pr.x & 255;
            const int y = pr.y;
            const bool inc =
                pin && (all_arms || (((mask >> ((pr.x >> 20) & 15)) & 1u) && ((mask >> ((pr.x >> 16) & 15)) & 1u)));
            const float4 A = W.linkC[l], Bc = W.linkC[mm];
            const float dx = A.x - Bc.x, dy = A.y - Bc.y, dz = A.z - Bc.z;
            const float Rs = A.w + Bc.w;
            const bool cand = inc && fmaf(dz, dz, fmaf(dy, dy, dx * dx)) <= Rs * Rs + Mc;
            cnt.lp += inc ? 1u : 0u;
            const unsigned long long m = __ballot(cand);
            if (m == 0ull) continue;
            if (cand) W.cand[lanes_below(m)] = y;  // the pair's sphere ranges
            wave_sync();
            const int n = __popcll(m);
            // spheres not placed yet: the candidates before the first one that reads spheres beyond
            // the first 64 are tested first, then those spheres are placed, then the others
            int c_hi = n;
            if (!placed) {
                const unsigned long long h = __ballot(cand && ((y >> 13) & 511) + ((y >> 22) & 15) > 64);
                if (h) c_hi = __popcll(m & ((h & (0ull - h)) - 1ull));
            }
            // the next candidate's entry is read from LDS while this one is tested
            int yn = W.cand[0];
            int c = 0;
            for (;;) {
                for (; c < c_hi; ++c) {
                    const int yc = yn;
                    if (c + 1 < n) yn = W.cand[c + 1];
                    const int fl = yc & 511, nl = (yc >> 9) & 15, fm = (yc >> 13) & 511, nm = (yc >> 22) & 15;
                    const int i = fl + (lane >> 3), j = fm + (lane & 7);
                    const bool act = (lane >> 3) < nl && (lane & 7) < nm;
                    // branch-free: lanes past the pair's spheres read the links' first spheres
                    // and their outcome is masked (no exec-mask regions in the hot loop)
                    const float4 ci = W.sph[act ? i : fl], cj = W.sph[act ? j : fm];
                    const float dx = ci.x - cj.x, dy = ci.y - cj.y, dz = ci.z - cj.z;
                    const float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                    const float Rs = ci.w + cj.w;
                    const float T = Rs * Rs;
                    cnt.sp += act ? 1u : 0u;
                    const bool near_c = act && d2 <= T + M;
                    const bool deep = d2 < T - M;
                    hit = hit || (near_c && deep);
                    const bool band = near_c && !deep;
                    arm_pend(W, band, 1, i, j, np);
                    if (__any(hit)) return true;
                    if (np > ARM_PEND_MAX - 64 && arm_flush(V, W, np, rechecks)) return true;
                }
                if (c_hi == n) break;
                place_rest();
                c_hi = n;
            }
            wave_sync();
            if (__any(hit)) return true;
        }
        // 5'. links vs boxes (boxes in registers, lane b holds box b), then the spheres of
        //     candidate (link, box) pairs: 8 candidates x 8 spheres per wave iteration.  A link
        //     whose bounding sphere clears the box around all obstacles cannot touch any box:
        //     lane l tests link l against that box first, and only the surviving links are walked.
        bool lsurv = false;
        if (lane < Dact && (all_arms || ((mask >> (lane / dof)) & 1u)) && (R.linkFC >> 16) > 0) {
            const float e0 = fminf(fmaxf(myLink.x, V.bb_lo[0]), V.bb_hi[0]) - myLink.x;
            const float e1 = fminf(fmaxf(myLink.y, V.bb_lo[1]), V.bb_hi[1]) - myLink.y;
            const float e2 = fminf(fmaxf(myLink.z, V.bb_lo[2]), V.bb_hi[2]) - myLink.z;
            lsurv = fmaf(e2, e2, fmaf(e1, e1, e0 * e0)) <= myLink.w * myLink.w + Mc;
        }
        const unsigned long long smask = __ballot(lsurv);
        if (V.nclu > 0 && smask) {
            // two levels: (surviving link, box cluster) pairs lane-parallel, then the boxes of the
            // candidate clusters (8 lanes per pair), then the spheres of the candidate (link, box)
            // pairs (8 lanes per pair)
            if (lsurv) W.lst[lanes_below(smask)] = lane;  // lane l carries link l
            wave_sync();
            const int nsv = __popcll(smask);
            const int K = V.nclu;
            const int lgKG = K <= 8 ? 3 : 4;  // lanes per link: 8 or 16
            for (int k0 = 0; k0 < nsv; k0 += 64 >> lgKG) {
                const int k = k0 + (lane >> lgKG), cc = lane & ((1 << lgKG) - 1);
                bool pc = false;
                int code = 0;
                if (k < nsv && cc < K) {
                    const int l = W.lst[k];
                    const float4 Lc = W.linkC[l];
                    pc = box_d2(Lc, W.clu[2 * cc], W.clu[2 * cc + 1]) <= Lc.w * Lc.w + Mc;
                    code = l | (cc << 8);
                    ++cnt.lb;
                }
                const unsigned long long pm = __ballot(pc);
                if (pm == 0ull) continue;
                if (pc) W.cand[lanes_below(pm)] = code;
                wave_sync();
                const int nA = __popcll(pm);
                for (int a0 = 0; a0 < nA; a0 += 8) {
                    const int g = lane >> 3, j = lane & 7;
                    // branch-free: idle lanes read valid entries (the group's first pair, box 0)
                    // and their outcome is masked
                    const bool ga = a0 + g < nA;
                    const int code2 = W.cand[ga ? a0 + g : a0];
                    const int l = code2 & 255, c2 = code2 >> 8;
                    const int first = __float_as_int(W.clu[2 * c2].w), count = __float_as_int(W.clu[2 * c2 + 1].w);
                    const bool act = ga && j < count;
                    const int b = act ? first + j : 0;
                    const float4 Lc = W.linkC[l];
                    const bool bcand = act && box_d2(Lc, W.box[2 * b], W.box[2 * b + 1]) <= Lc.w * Lc.w + Mc;
                    const int lb = l | (b << 8);
                    cnt.lb += act ? 1u : 0u;
                    const unsigned long long bm = __ballot(bcand);
                    if (bm == 0ull) continue;
                    if (bcand) W.lbc[lanes_below(bm)] = lb;
                    wave_sync();
                    const int nB = __popcll(bm);
                    for (int c0 = 0; c0 < nB; c0 += 8) {
                        const int g2 = lane >> 3, j2 = lane & 7;
                        const bool gb = c0 + g2 < nB;
                        const int v = W.lbc[gb ? c0 + g2 : c0];
                        const int fc = W.linkFC[v & 255];
                        const bool act = gb && j2 < (fc >> 16);
                        const int bb = v >> 8;
                        const int sph = (fc & 0xffff) + (act ? j2 : 0);
                        const float4 c = W.sph[sph];
                        const float d2 = box_d2(c, W.box[2 * bb], W.box[2 * bb + 1]);
                        const float T = c.w * c.w;
                        cnt.sb += act ? 1u : 0u;
                        const bool near_c = act && d2 <= T + M;
                        const bool deep = d2 < T - M;
                        hit = hit || (near_c && deep);
                        const bool band = near_c && !deep;
                        arm_pend(W, band, 0, sph, bb, np);
                        if (np > ARM_PEND_MAX - 64 && arm_flush(V, W, np, rechecks)) return true;
                    }
                    wave_sync();  // W.lbc is rewritten by the next group
                    if (__any(hit)) return true;
                }
                wave_sync();  // W.cand is rewritten by the next round
            }
        }
        for (unsigned long long lm = V.nclu > 0 ? 0ull : smask; lm; lm &= lm - 1) {
            const int l = __builtin_ctzll(lm);
            const int fc = __builtin_amdgcn_readlane(R.linkFC, l);
            const int f = fc & 0xffff, ns = fc >> 16;
            float4 Lc;
            Lc.x = readlane_f(myLink.x, l);
            Lc.y = readlane_f(myLink.y, l);
            Lc.z = readlane_f(myLink.z, l);
            Lc.w = readlane_f(myLink.w, l);
            const float Tl = Lc.w * Lc.w;
            for (int b0 = 0; b0 < V.nb; b0 += 64) {
                const int b = b0 + lane;
                const bool bin = b < V.nb;
                const int bc = bin ? b : 0;
                const float4 lo = W.boxG[2 * bc], hi = W.boxG[2 * bc + 1];
                const bool cand = bin && box_d2(Lc, lo, hi) <= Tl + Mc;
                cnt.lb += bin ? 1u : 0u;
                const unsigned long long m = __ballot(cand);
                if (m == 0ull) continue;
                if (cand) W.cand[lanes_below(m)] = b;
                wave_sync();
                const int n = __popcll(m);
                for (int c0 = 0; c0 < n; c0 += 8) {
                    const int k = c0 + (lane >> 3), j = lane & 7;
                    const bool act = k < n && j < ns;
                    const int bb = W.cand[act ? k : c0];
                    const int s = f + (act ? j : 0);
                    const float4 c = W.sph[s];
                    const float d2 = box_d2(c, W.boxG[2 * bb], W.boxG[2 * bb + 1]);
                    const float T = c.w * c.w;
                    cnt.sb += act ? 1u : 0u;
                    const bool near_c = act && d2 <= T + M;
                    const bool deep = d2 < T - M;
                    hit = hit || (near_c && deep);
                    const bool band = near_c && !deep;
                    arm_pend(W, band, 0, s, bb, np);
                    if (np > ARM_PEND_MAX - 64 && arm_flush(V, W, np, rechecks)) return true;
                }
                wave_sync();
                if (__any(hit)) return true;
            }
        }
    } else if (all_arms) {
        // Brute force with every arm active.  Blocks of 16 tests only compare against the bounds
        // (T - M: certain contact; T + M: near contact); a block in which some lane is near contact
        // without a certain one is tested again with the band bookkeeping (rare).
        // 4. spheres vs boxes
        const float4* gbox = (const float4*)V.box32();  // uniform index: scalar loads
        for (int base = 0; base < S_act; base += 64) {
            const int s = base + lane;
            const bool act = s < S_act;
            const float4 c = W.sph[act ? s : 0];
            const float T = c.w * c.w;
            const float Tp = act ? T + M : -1.0f, Tm = act ? T - M : -1.0f;  // -1: never (d2 >= 0)
            for (int b0 = 0; b0 < V.nb; b0 += 16) {
                const int b1 = b0 + 16 < V.nb ? b0 + 16 : V.nb;
                bool near = false;
#pragma unroll 4
                for (int b = b0; b < b1; ++b) {
                    const float d2 = box_d2(c, gbox[2 * b], gbox[2 * b + 1]);
                    near |= d2 <= Tp;
                    hit |= d2 < Tm;
                }
                if (act) cnt.sb += b1 - b0;
                if (__any(hit)) return true;
                if (__any(near)) {
                    for (int b = b0; b < b1; ++b) {
                        const float d2 = box_d2(c, gbox[2 * b], gbox[2 * b + 1]);
                        arm_pend(W, d2 <= Tp, 0, s, b, np);  // no certain contact in this block
                        if (np > ARM_PEND_MAX - 64 && arm_flush(V, W, np, rechecks)) return true;
                    }
                }
            }
        }
        // 5. sphere pairs: same arm links >= 2 apart, and different arms; lane i tests j >= lo(i)
        for (int base = 0; base < S_act; base += 64) {
            const int i = base + lane;
            const bool act = i < S_act;
            const float4 ci = W.sph[act ? i : 0];
            const int lo = act ? V.sph_lo()[i] : 0x7fffffff;
            const int jstart = V.sph_lo()[base];
            for (int j0 = jstart; j0 < S_act; j0 += 16) {
                const int j1 = j0 + 16 < S_act ? j0 + 16 : S_act;
                bool near = false;
#pragma unroll 4
                for (int j = j0; j < j1; ++j) {
                    const float4 cj = W.sph[j];
                    const float dx = ci.x - cj.x, dy = ci.y - cj.y, dz = ci.z - cj.z;
                    const float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                    const float Rs = ci.w + cj.w;
                    const float T = Rs * Rs;
                    const bool valid = j >= lo;
                    near |= valid && d2 <= T + M;
                    hit |= valid && d2 < T - M;
                }
                if (act && lo < j1) cnt.sp += j1 - (lo > j0 ? lo : j0);
                if (__any(hit)) return true;
                if (__any(near)) {
                    for (int j = j0; j < j1; ++j) {
                        const float4 cj = W.sph[j];
                        const float dx = ci.x - cj.x, dy = ci.y - cj.y, dz = ci.z - cj.z;
                        const float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                        const float Rs = ci.w + cj.w;
                        const float T = Rs * Rs;
                        arm_pend(W, j >= lo && d2 <= T + M, 1, i, j, np);  // no certain contact here
                        if (np > ARM_PEND_MAX - 64 && arm_flush(V, W, np, rechecks)) return true;
                    }
                }
            }
        }
    } else {
        // 4. spheres vs boxes (brute force, some arms masked out)
        for (int base = 0; base < S_act; base += 64) {
            const int s = base + lane;
            const bool act = s < S_act && ((mask >> (W.sphLink[s < S_act ? s : 0] >> 8)) & 1u);
            const float4 c = W.sph[act ? s : 0];
            const float T = c.w * c.w;
            const float4* gbox = (const float4*)V.box32();  // uniform index: scalar loads
            for (int b = 0; b < V.nb; ++b) {
                const float d2 = box_d2(c, gbox[2 * b], gbox[2 * b + 1]);
                cnt.sb += act;
                const bool near_c = act && d2 <= T + M;
                const bool deep = d2 < T - M;
                hit = hit || (near_c && deep);
                arm_pend(W, near_c && !deep, 0, s, b, np);
                if (np > ARM_PEND_MAX - 64 && arm_flush(V, W, np, rechecks)) return true;
            }
            if (__any(hit)) return true;
        }
        // 5. sphere pairs: same arm links >= 2 apart, and different arms (brute force, masked)
        for (int base = 0; base < S_act; base += 64) {
            const int i = base + lane;
            const bool act = i < S_act && ((mask >> (W.sphLink[i < S_act ? i : 0] >> 8)) & 1u);
            const float4 ci = W.sph[act ? i : 0];
            const int lo = act ? V.sph_lo()[i] : 0x7fffffff;
            const int jstart = V.sph_lo()[base];
            int nj = 0;
            for (int j = jstart; j < S_act; ++j) {
                const int aj = W.sphLink[j] >> 8;
                if (!((mask >> aj) & 1u)) continue;  // wave-uniform
                const float4 cj = W.sph[j];
                const float dx = ci.x - cj.x, dy = ci.y - cj.y, dz = ci.z - cj.z;
                const float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                const float Rs = ci.w + cj.w;
                const float T = Rs * Rs;
                cnt.sp += (j >= lo);
                const bool near_c = j >= lo && d2 <= T + M;
                const bool deep = d2 < T - M;
                hit = hit || (near_c && deep);
                arm_pend(W, near_c && !deep, 1, i, j, np);
                if (np > ARM_PEND_MAX - 64 && arm_flush(V, W, np, rechecks)) return true;
                if (++nj == 16) {
                    nj = 0;
                    if (__any(hit)) return true;
                }
            }
            if (__any(hit)) return true;
        }
    }
    if (np && arm_flush(V, W, np, rechecks)) return true;
    // 6. mode boxes: slot a < na holds the box held by arm a, slot na + j the j-th resting box
    if constexpr (!GRASP) return false;
    unsigned gmask = 0;  // held boxes of active arms in the mask, resting boxes
    for (int a = 0; a < V.nslot; ++a) {
        const bool held = a < V.na;
        if (held && (a >= na || !((mask >> a) & 1u))) continue;
        if (V.m_grasp()[mode * V.nslot + a]) gmask |= 1u << a;
    }
    if (gmask) {
        const float Mb = V.Mobb;
        // 6a. held box vs obstacle boxes (resting boxes do not move: they are not tested against
        //     the static obstacles)
        for (int a = 0; a < na; ++a) {
            if (!((gmask >> a) & 1u)) continue;
            const float* G = W.G32 + 16 * a;
            for (int b0 = 0; b0 < V.nb; b0 += 64) {
                const int b = b0 + lane;
                bool band = false;
                if (b < V.nb) {
                    float A[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
                    const float* bc = V.boxch32() + 6 * b;
                    A[9] = bc[0]; A[10] = bc[1]; A[11] = bc[2];
                    const int r = obb3_overlap_f32(A, bc + 3, G, G + 12, Mb);
                    if (r == 1) hit = true;
                    else if (r == 2) band = true;
                }
                arm_pend(W, band, 2, a, b, np);
                if (np > ARM_PEND_MAX - 64 && arm_flush_held(V, W, mode, np, rechecks)) return true;
            }
        }
        if (__any(hit)) return true;
        // 6b. mode box vs the spheres of every arm in the mask; a held box skips the spheres of
        //     its holding link (the last link of arm a): the attachment deactivates only the
        //     (object, holding link) pair (pinocchio_env.py:266-272; rai attach + setContact(-1),
        //     rai_base_env.py:801-805)
        for (int a = 0; a < V.nslot; ++a) {
            if (!((gmask >> a) & 1u)) continue;
            const float* G = W.G32 + 16 * a;
            const int hold = a < V.na ? a * dof + dof - 1 : 255;  // 255: no link
            for (int base = 0; base < S_act; base += 64) {
                const int s = base + lane;
                const int sl = s < S_act ? W.sphLink[s] : 255;
                const int as = sl >> 8;
                const bool act = s < S_act && (sl & 255) != hold && ((mask >> as) & 1u);
                bool band = false;
                if (act) {
                    const float4 c = W.sph[s];
                    const float w0 = c.x - G[9], w1 = c.y - G[10], w2 = c.z - G[11];
                    float d2 = 0.0f;
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const float l = G[k] * w0 + G[3 + k] * w1 + G[6 + k] * w2;
                        const float e = fminf(fmaxf(l, -G[12 + k]), G[12 + k]) - l;
                        d2 = fmaf(e, e, d2);
                    }
                    const float T = c.w * c.w;
                    if (d2 <= T + Mb) {
                        if (d2 < T - Mb) hit = true;
                        else band = true;
                    }
                }
                arm_pend(W, band, 3, s, a, np);
                if (np > ARM_PEND_MAX - 64 && arm_flush_held(V, W, mode, np, rechecks)) return true;
            }
        }
        if (__any(hit)) return true;
        // 6c. mode box pairs (a, b), a < b, a held: held vs held and held vs resting (resting
        //     boxes do not move relative to each other); one lane per pair
        const int ns = V.nslot;
        for (int p0 = 0; p0 < ns * ns; p0 += 64) {
            const int p = p0 + lane;
            const int a = p / ns, b = p - a * ns;
            bool band = false;
            if (p < ns * ns && a < b && a < V.na && ((gmask >> a) & 1u) && ((gmask >> b) & 1u)) {
                const float* Ga = W.G32 + 16 * a;
                const float* Gb = W.G32 + 16 * b;
                const int r = obb3_overlap_f32(Ga, Ga + 12, Gb, Gb + 12, Mb);
                if (r == 1) hit = true;
                else if (r == 2) band = true;
            }
            arm_pend(W, band, 4, a, b, np);
            if (np > ARM_PEND_MAX - 64 && arm_flush_held(V, W, mode, np, rechecks)) return true;
        }
        if (__any(hit)) return true;
        if (np && arm_flush_held(V, W, mode, np, rechecks)) return true;
    }
    return __any(hit);
}
