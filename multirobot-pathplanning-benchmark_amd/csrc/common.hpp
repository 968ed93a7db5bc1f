// common.hpp — shared device/host definitions of the collision engine (gfx950).
//
// Arithmetic contract
//   * Everything that decides a boolean near contact is binary64 with contraction disabled
//     (`#pragma clang fp contract(off)` inside the *_f64 functions): each product and sum rounds
//     separately, exactly like the CPU definition of the predicates.
//   * The bulk narrow phase runs in binary32 on coordinates rounded from the binary64 kinematics.
//     Every binary32 test returns one of three answers: sure-free, sure-hit, or "band" when the
//     value lies within a margin M of the contact threshold.  M bounds the binary32 error (see
//     margin_* below), so sure answers agree with binary64; band pairs are re-evaluated in binary64.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define MRCC_MAX_ARMS 8
#define MRCC_MAX_DOF 8
#define MRCC_MAX_LINKS (MRCC_MAX_ARMS * MRCC_MAX_DOF)
#define MRCC_PERSISTENT_GRASP_NW 10  // largest workgroup (waves) of the persistent kernel with held boxes
#define MRCC_MODES_LDS_MAX 1024  // per-mode arm counts of up to this many modes live in LDS
#define MRCC_MAX_SPHERES 512
#define MRCC_MAX_D 64
#define MRCC_DISK_MAX_RD 12
#define MRCC_DISK_MAX_ROBOTS 16
#define MRCC_PLANAR_MAX_ROBOTS 16
#define MRCC_MAX_PASSES 16384
#define MRCC_MAX_EDGE_STATES (1 << 24)  // longer edges are rejected (error), not iterated

enum { KIND_DISKS = 1, KIND_PLANAR = 2, KIND_ARMS = 3 };

// --------------------------------------------------------------------------------------------
// binary64 sincos (Cody-Waite reduction by pi/2 in three parts, Taylor polynomials on
// [-pi/4, pi/4], Horner, no FMA).  The same sequence of operations on host and device gives the
// same bits on both.
namespace mrcc {

struct SinCos {
    double s, c;
};

__host__ __device__ inline SinCos sincos_f64(double x) {
#pragma clang fp contract(off)
    const double TWO_OVER_PI = 0x1.45f306dc9c883p-1;
    const double PIO2_1 = 0x1.921fb54400000p+0;
    const double PIO2_2 = 0x1.0b4611a600000p-34;
    const double PIO2_3 = 0x1.3198a2e037073p-69;
    double kd = __builtin_rint(x * TWO_OVER_PI);
    double r = ((x - kd * PIO2_1) - kd * PIO2_2) - kd * PIO2_3;
    double r2 = r * r;
    double ps = 0x1.952c77030ad4ap-49;
    ps = -0x1.ae7f3e733b81fp-41 + r2 * ps;
    ps = 0x1.6124613a86d09p-33 + r2 * ps;
    ps = -0x1.ae64567f544e4p-26 + r2 * ps;
    ps = 0x1.71de3a556c734p-19 + r2 * ps;
    ps = -0x1.a01a01a01a01ap-13 + r2 * ps;
    ps = 0x1.1111111111111p-7 + r2 * ps;
    ps = -0x1.5555555555555p-3 + r2 * ps;
    double sn = r + (r * r2) * ps;
    double pc = -0x1.6827863b97d97p-53;
    pc = 0x1.ae7f3e733b81fp-45 + r2 * pc;
    pc = -0x1.93974a8c07c9dp-37 + r2 * pc;
    pc = 0x1.1eed8eff8d898p-29 + r2 * pc;
    pc = -0x1.27e4fb7789f5cp-22 + r2 * pc;
    pc = 0x1.a01a01a01a01ap-16 + r2 * pc;
    pc = -0x1.6c16c16c16c17p-10 + r2 * pc;
    pc = 0x1.5555555555555p-5 + r2 * pc;
    double cs = (1.0 - 0.5 * r2) + (r2 * r2) * pc;
    double qd = kd - 4.0 * __builtin_floor(kd * 0.25);
    int q = (int)qd & 3;
    // quadrant q: (s, c) = (sn, cs), (cs, -sn), (-sn, -cs), (-cs, sn), as selects (q differs
    // between lanes: branches would run every quadrant's path under lane masks)
    const bool odd = (q & 1) != 0;
    const double s0 = odd ? cs : sn, c0 = odd ? sn : cs;
    SinCos o;
    o.s = (q & 2) ? -s0 : s0;
    o.c = ((q + 1) & 2) ? -c0 : c0;
    return o;
}

// X = [R row-major | t]; X <- X * A(q) for one DH joint (d, a, cos alpha, sin alpha)
__host__ __device__ inline void dh_step_f64(double* X, double d, double a, double ca, double sa,
                                            SinCos j) {
#pragma clang fp contract(off)
    const double sn = j.s, cs = j.c;
    double A[3][4];
    A[0][0] = cs; A[0][1] = -(sn * ca); A[0][2] = sn * sa; A[0][3] = a * cs;
    A[1][0] = sn; A[1][1] = cs * ca;   A[1][2] = -(cs * sa); A[1][3] = a * sn;
    A[2][0] = 0.0; A[2][1] = sa;       A[2][2] = ca;         A[2][3] = d;
    double Y[12];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const double R0 = X[3 * r], R1 = X[3 * r + 1], R2 = X[3 * r + 2];
#pragma unroll
        for (int c = 0; c < 3; ++c) Y[3 * r + c] = (R0 * A[0][c] + R1 * A[1][c]) + R2 * A[2][c];
        Y[9 + r] = ((R0 * A[0][3] + R1 * A[1][3]) + R2 * A[2][3]) + X[9 + r];
    }
#pragma unroll
    for (int k = 0; k < 12; ++k) X[k] = Y[k];
}

// one row (R0 R1 R2 | t) of X <- X * A(q): the row-r entries of dh_step_f64, same operations
__host__ __device__ inline void dh_row_f64(double& R0, double& R1, double& R2, double& t,
                                           const double* p, double cs, double sn) {
#pragma clang fp contract(off)
    const double d = p[0], a = p[1], ca = p[2], sa = p[3];
    const double A01 = -(sn * ca), A02 = sn * sa, A03 = a * cs;
    const double A11 = cs * ca, A12 = -(cs * sa), A13 = a * sn;
    const double y0 = (R0 * cs + R1 * sn) + R2 * 0.0;
    const double y1 = (R0 * A01 + R1 * A11) + R2 * sa;
    const double y2 = (R0 * A02 + R1 * A12) + R2 * ca;
    const double ty = ((R0 * A03 + R1 * A13) + R2 * d) + t;
    R0 = y0;
    R1 = y1;
    R2 = y2;
    t = ty;
}

// Chain joints: X <- X * Rz(q) C, C a fixed 3 x 4 transform [R | t] (rows C0, C1, C2).  The step
// A = Rz(q) C has rows A0c = cs C0c - sn C1c, A1c = sn C0c + cs C1c, A2c = C2c; each entry of
// X's row r is (R0 A0c + R1 A1c) + R2 A2c (+ t for c = 3), as dh_step_f64.  For C of a DH joint,
// [[1, 0, 0, a], [0, ca, -sa, 0], [0, sa, ca, d]], the zeros make every A entry equal to
// dh_step_f64's (up to the sign of a zero).
struct ChainCoef {
    double a0, a1;
};
__host__ __device__ inline ChainCoef chain_coef_f64(double cs, double sn, double c0, double c1) {
#pragma clang fp contract(off)
    ChainCoef k;
    k.a0 = cs * c0 - sn * c1;
    k.a1 = sn * c0 + cs * c1;
    return k;
}

// Row 2, column 3 of a chain joint's step: C23 for a revolute joint (A = Rz(q) C), C23 + q for a
// prismatic joint (A = Tz(q) C, whose rows 0 and 1 are those of Rz(0) C).
__host__ __device__ inline double chain_a23_f64(const double* C, bool prismatic, double q) {
    return prismatic ? C[11] + q : C[11];
}

// X <- X * A for a chain joint with joint value q: A = Rz(q) C (revolute, j = sincos(q)) or
// A = Tz(q) C (prismatic, j = {cos 0, sin 0} = {1, 0}).
__host__ __device__ inline void chain_step_f64(double* X, const double* C, SinCos j, bool prismatic = false,
                                               double q = 0.0) {
#pragma clang fp contract(off)
    double A[3][4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const ChainCoef k = chain_coef_f64(j.c, j.s, C[c], C[4 + c]);
        A[0][c] = k.a0;
        A[1][c] = k.a1;
        A[2][c] = c == 3 ? chain_a23_f64(C, prismatic, q) : C[8 + c];
    }
    double Y[12];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const double R0 = X[3 * r], R1 = X[3 * r + 1], R2 = X[3 * r + 2];
#pragma unroll
        for (int c = 0; c < 3; ++c) Y[3 * r + c] = (R0 * A[0][c] + R1 * A[1][c]) + R2 * A[2][c];
        Y[9 + r] = ((R0 * A[0][3] + R1 * A[1][3]) + R2 * A[2][3]) + X[9 + r];
    }
#pragma unroll
    for (int k = 0; k < 12; ++k) X[k] = Y[k];
}

// one row of G = X * Rel (compose_f64)
__host__ __device__ inline void compose_row_f64(double R0, double R1, double R2, double t,
                                                const double* Rel, double* g) {
#pragma clang fp contract(off)
    for (int c = 0; c < 3; ++c) g[c] = (R0 * Rel[c] + R1 * Rel[3 + c]) + R2 * Rel[6 + c];
    g[3] = ((R0 * Rel[9] + R1 * Rel[10]) + R2 * Rel[11]) + t;
}

__host__ __device__ inline void xform_point_f64(const double* X, double l0, double l1, double l2,
                                                double* c) {
#pragma clang fp contract(off)
#pragma unroll
    for (int r = 0; r < 3; ++r) c[r] = ((X[3 * r] * l0 + X[3 * r + 1] * l1) + X[3 * r + 2] * l2) + X[9 + r];
}

__host__ __device__ inline void compose_f64(const double* X, const double* Rel, double* G) {
#pragma clang fp contract(off)
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
            G[3 * r + c] = (X[3 * r] * Rel[c] + X[3 * r + 1] * Rel[3 + c]) + X[3 * r + 2] * Rel[6 + c];
        G[9 + r] = ((X[3 * r] * Rel[9] + X[3 * r + 1] * Rel[10]) + X[3 * r + 2] * Rel[11]) + X[9 + r];
    }
}

// closed sphere vs axis-aligned box [lo, hi]
__host__ __device__ inline bool sph_box_f64(const double* c, double r, const double* lo,
                                            const double* hi) {
#pragma clang fp contract(off)
    double e[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double m = c[k] >= lo[k] ? c[k] : lo[k];
        m = m <= hi[k] ? m : hi[k];
        e[k] = m - c[k];
    }
    return ((e[0] * e[0] + e[1] * e[1]) + e[2] * e[2]) <= r * r;
}

__host__ __device__ inline bool sph_sph_f64(const double* a, double ra, const double* b, double rb) {
#pragma clang fp contract(off)
    double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
    double R = ra + rb;
    return ((d0 * d0 + d1 * d1) + d2 * d2) <= R * R;
}

// closed sphere vs oriented box G = [R | t] with half extents h
__host__ __device__ inline bool sph_obb_f64(const double* c, double r, const double* G,
                                            const double* h) {
#pragma clang fp contract(off)
    double w0 = c[0] - G[9], w1 = c[1] - G[10], w2 = c[2] - G[11];
    double e[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double l = (G[k] * w0 + G[3 + k] * w1) + G[6 + k] * w2;
        double m = l >= -h[k] ? l : -h[k];
        m = m <= h[k] ? m : h[k];
        e[k] = m - l;
    }
    return ((e[0] * e[0] + e[1] * e[1]) + e[2] * e[2]) <= r * r;
}

// closed 15-axis separating-axis overlap of boxes A = [R | t] (half a) and B (half b)
__host__ __device__ inline bool obb3_overlap_f64(const double* A, const double* a, const double* B,
                                                 const double* b) {
#pragma clang fp contract(off)
    double Rm[3][3], AR[3][3], t[3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            Rm[i][j] = (A[i] * B[j] + A[3 + i] * B[3 + j]) + A[6 + i] * B[6 + j];
            AR[i][j] = __builtin_fabs(Rm[i][j]);
        }
    double D0 = B[9] - A[9], D1 = B[10] - A[10], D2 = B[11] - A[11];
    for (int i = 0; i < 3; ++i) t[i] = (A[i] * D0 + A[3 + i] * D1) + A[6 + i] * D2;
    for (int i = 0; i < 3; ++i) {
        double rb = (b[0] * AR[i][0] + b[1] * AR[i][1]) + b[2] * AR[i][2];
        if (__builtin_fabs(t[i]) - (a[i] + rb) > 0.0) return false;
    }
    for (int j = 0; j < 3; ++j) {
        double ra = (a[0] * AR[0][j] + a[1] * AR[1][j]) + a[2] * AR[2][j];
        double tp = (t[0] * Rm[0][j] + t[1] * Rm[1][j]) + t[2] * Rm[2][j];
        if (__builtin_fabs(tp) - (ra + b[j]) > 0.0) return false;
    }
    for (int i = 0; i < 3; ++i) {
        int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
        for (int j = 0; j < 3; ++j) {
            int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            double tp = t[i2] * Rm[i1][j] - t[i1] * Rm[i2][j];
            double rr = ((a[i1] * AR[i2][j] + a[i2] * AR[i1][j]) + b[j1] * AR[i][j2]) + b[j2] * AR[i][j1];
            if (__builtin_fabs(tp) - rr > 0.0) return false;
        }
    }
    return true;
}

// binary32 15-axis SAT: returns 1 sure-overlap, 0 sure-separated, 2 band (|value| <= M on an
// axis, no axis surely separating)
__device__ inline int obb3_overlap_f32(const float* A, const float* a, const float* B,
                                       const float* b, float M) {
    float Rm[3][3], AR[3][3], t[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            Rm[i][j] = A[i] * B[j] + A[3 + i] * B[3 + j] + A[6 + i] * B[6 + j];
            AR[i][j] = fabsf(Rm[i][j]);
        }
    float D0 = B[9] - A[9], D1 = B[10] - A[10], D2 = B[11] - A[11];
#pragma unroll
    for (int i = 0; i < 3; ++i) t[i] = A[i] * D0 + A[3 + i] * D1 + A[6 + i] * D2;
    float mx = -3.0e38f;
#pragma unroll
    for (int i = 0; i < 3; ++i)
        mx = fmaxf(mx, fabsf(t[i]) - (a[i] + b[0] * AR[i][0] + b[1] * AR[i][1] + b[2] * AR[i][2]));
#pragma unroll
    for (int j = 0; j < 3; ++j)
        mx = fmaxf(mx, fabsf(t[0] * Rm[0][j] + t[1] * Rm[1][j] + t[2] * Rm[2][j]) -
                           (a[0] * AR[0][j] + a[1] * AR[1][j] + a[2] * AR[2][j] + b[j]));
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            mx = fmaxf(mx, fabsf(t[i2] * Rm[i1][j] - t[i1] * Rm[i2][j]) -
                               (a[i1] * AR[i2][j] + a[i2] * AR[i1][j] + b[j1] * AR[i][j2] + b[j2] * AR[i][j1]));
        }
    }
    if (mx > M) return 0;
    if (mx < -M) return 1;
    return 2;
}

// --------------------------------------------------------------------------------------------
// breadth-first midpoint order of planning_env.py:34-51 without tables: the k-th element.
// Levels of the interval tree fill completely except the last; the last level holds the
// length-1 intervals, found by descending with the count of non-empty nodes per subtree.
__host__ __device__ inline int ilog2_floor(unsigned v) { return 31 - __builtin_clz(v); }

// number of non-empty nodes at depth dd in the tree of an interval of length n
__host__ __device__ inline int bfs_level_count(int n, int dd) {
    if (n <= 0) return 0;
    int D = ilog2_floor((unsigned)n + 1u);  // levels 0..D-1 are full
    if (dd < D) return 1 << dd;
    if (dd == D) return n - ((1 << D) - 1);
    return 0;
}

__host__ __device__ inline int bfs_index(int N, int k) {
    int D = ilog2_floor((unsigned)N + 1u);
    int s = 0, e = N - 1;
    if (k < (1 << D) - 1) {
        int L = ilog2_floor((unsigned)k + 1u);
        int o = k - ((1 << L) - 1);
        for (int b = L - 1; b >= 0; --b) {
            int mid = (s + e) >> 1;
            if ((o >> b) & 1) s = mid + 1;
            else e = mid - 1;
        }
        return (s + e) >> 1;
    }
    int o = k - ((1 << D) - 1);
    for (int dd = D; dd > 0; --dd) {
        int mid = (s + e) >> 1;
        int nl = bfs_level_count(mid - s, dd - 1);
        if (o < nl) e = mid - 1;
        else { o -= nl; s = mid + 1; }
    }
    return (s + e) >> 1;
}

// --------------------------------------------------------------------------------------------
// edge window bookkeeping (abstract_env.py:321-352)
struct EdgeWin {
    int N;        // state count
    int w0;       // first BFS position of the window
    int M;        // slots (BFS positions or interior states) to visit
    int full;     // 1: whole edge (visiting order is free)
    int first;    // full edges: first state index (0 with endpoints, else 1)
    int err;      // 1: N_start > N
};

__host__ __device__ inline int edge_N_from_max(double m, double res) {
#pragma clang fp contract(off)
    double t = m / res;
    if (!(t < 1.0e9)) return 1000000001;  // saturate (rejected by the planner of the batch)
    int N = (int)t + 1;
    return N < 2 ? 2 : N;
}

__host__ __device__ inline EdgeWin edge_window(int N, int N_start, int N_max, int incl) {
    EdgeWin w;
    w.N = N;
    w.err = (N_start < 0 || N_start > N || N > MRCC_MAX_EDGE_STATES) ? 1 : 0;
    int hi = (N_max < 0 || N_max > N) ? N : N_max;
    w.full = (N_start == 0 && hi == N) ? 1 : 0;
    w.first = incl ? 0 : 1;
    if (w.err) { w.M = 0; w.w0 = 0; return w; }
    if (w.full) {
        w.w0 = 0;
        w.M = incl ? N : N - 2;
        if (w.M < 0) w.M = 0;
    } else {
        w.w0 = N_start;
        w.M = hi > N_start ? hi - N_start : 0;
    }
    return w;
}

// state index of slot `local` (0 <= local < M) of an edge, or -1 when the slot is an endpoint
// that the call excludes
__host__ __device__ inline int edge_state_index(const EdgeWin& w, int local, int incl) {
    if (w.full) return w.first + local;
    int i = bfs_index(w.N, w.w0 + local);
    if (!incl && (i == 0 || i == w.N - 1)) return -1;
    return i;
}

}  // namespace mrcc

// --------------------------------------------------------------------------------------------
// device views of the scenes (plain pointers into one library-owned device allocation)
struct DiskView {
    int nr, rd, D, ns, nrect;
    float Bscene;
    const float* r32;                       // [nr]
    const float* rrT32; const double* rrT64;  // [nr*nr] thresholds of sqrt(d2) < r_i + r_j
    const float* sc32; const double* sc64;    // [ns*rd]
    const float* srT32; const double* srT64;  // [nr*ns] thresholds of sqrt(d2) < r_o + r_i
    const float* lo32; const float* hi32;     // [nrect*rd]
    const double* lo64; const double* hi64;
    const float* rp2_32; const double* rp2_64;  // [nr] r**2
};

struct PlanarView {
    int nr, no, D;
    float Bscene;
    const float* rh32; const double* rh64;  // [nr*2]
    const float* ob32; const double* ob64;  // [no*8] x y ux uy vx vy h0 h1
};

struct ArmView {
    int na, dof, D, L, S, nb, nm;
    int nslot;    // mode-box slots per mode (na held + resting)
    float Msph;   // binary32 margin of sphere tests (squared-distance units)
    float Mobb;   // binary32 margin of box SAT axis values
    const double* base;      // [na*12]
    int chain;               // 1: joints are chain joints (dh holds C, 12 per joint)
    const double* dh;        // [L*4] d a cos(alpha) sin(alpha), or [L*12] C of each chain joint
    unsigned long long prism;  // bit l: chain joint l is prismatic (Tz(q) C)
    const double* travel;    // [L] |q| bound of each prismatic joint
    const int* link_first;   // [L+1]
    const int* sph_link;     // [S] global link index
    const int* sph_lo;       // [S] first allowed partner sphere (same arm, links >= link+2)
    const double* sph64;     // [S*4] local xyz r
    const float* sph_r32;    // [S]
    const float* box32;      // [nb*8] lo xyz, pad, hi xyz, pad
    const double* box64;     // [nb*6] lo xyz hi xyz
    const double* boxch64;   // [nb*6] centre xyz half xyz
    const float* boxch32;    // [nb*6]
    const int* m_active;     // [nm]
    const unsigned char* m_grasp;  // [nm*na]
    const double* m_rel;     // [nm*na*12]
    const double* m_half;    // [nm*na*3]
    // broad phase: link bounding spheres and the link pairs to test
    int broadphase;          // 1: cull sphere tests with link bounding spheres
    float Mcull;             // binary32 margin of the bounding-sphere tests
    const double* link_bc64; // [L*3] bounding-sphere centre in the link frame
    const float* link_R32;   // [L] bounding radius (covers every sphere of the link, inflated)
    int grasp;               // 1: some mode holds a box (the wave LDS region has held-box space)
    int NP;                  // link pairs in the pair list
    const int2* lpairs2;     // [NP] (m | l << 8 | arm(m) << 16 | arm(l) << 20,
                             //  fl | nl << 9 | fm << 13 | nm << 22): link pair l < m and the sphere
                             // ranges [fl, fl+nl), [fm, fm+nm); the self pairs of every
                             // arm, then one segment per arm pair (a, b), a < b
    const int* lpair_sets;   // [na+1 + na(na-1)/2]: [n] self pairs of arms < n; [na+1+k]
                             // segment k = b(b-1)/2 + a: first | count << 12 | a << 20 | b << 24
    float Marm;              // binary32 margin of the arm-box separation test
    float bb_lo[3], bb_hi[3];   // box containing every obstacle box (binary32, rounded outward)
    int nclu;                // obstacle box clusters (0: the link-box test has one level)
    const float* clu32;      // [nclu*8] cluster bounds: lo xyz, first box (int bits), hi xyz, count
};

// Arm scene as the edge and configuration kernels see it: scalars, and one base pointer into the
// scene's device allocation with 32-bit byte offsets of the arrays read from global memory (fewer
// scalar registers than one 64-bit pointer per array).  The tables the evaluator reads from LDS are
// one contiguous image in the allocation, laid out as the static LDS region (arm_static_layout),
// which a workgroup copies with 16-byte loads.
struct ArmK {
    int na, dof, D, L, S, nb, nm, NP, nclu;
    int nslot;       // mode-box slots: na held boxes (slot a: arm a's end-effector) + resting boxes
    int chain;       // 1: chain joints
    int broadphase;  // 1: link bounding spheres cull sphere tests
    int grasp;       // 1: some mode holds a box
    int img_bytes;   // size of the static LDS image (multiple of 16)
    unsigned long long prism;  // bit l: chain joint l is prismatic (Tz(q) C)
    float Msph, Mobb, Mcull, Marm;
    float bb_lo[3], bb_hi[3];
    const unsigned char* g;
    int o_img, o_box32, o_box64, o_boxch32, o_boxch64, o_sph64, o_sphlo, o_pairs, o_mact, o_mgrasp,
        o_mrel, o_mhalf, o_travel;
    __device__ const float4* img() const { return (const float4*)(g + o_img); }
    __device__ const float* box32() const { return (const float*)(g + o_box32); }       // [nb*8]
    __device__ const double* box64() const { return (const double*)(g + o_box64); }     // [nb*6] lo hi
    __device__ const float* boxch32() const { return (const float*)(g + o_boxch32); }   // [nb*6]
    __device__ const double* boxch64() const { return (const double*)(g + o_boxch64); } // [nb*6]
    __device__ const double* sph64() const { return (const double*)(g + o_sph64); }     // [S*4]
    __device__ const int* sph_lo() const { return (const int*)(g + o_sphlo); }          // [S]
    __device__ const int2* lpairs2() const { return (const int2*)(g + o_pairs); }       // [NP]
    __device__ const int* m_active() const { return (const int*)(g + o_mact); }         // [nm]
    __device__ const unsigned char* m_grasp() const { return g + o_mgrasp; }            // [nm*nslot]
    __device__ const double* m_rel() const { return (const double*)(g + o_mrel); }      // [nm*nslot*12]
    __device__ const double* m_half() const { return (const double*)(g + o_mhalf); }    // [nm*nslot*3]
    __device__ const double* travel() const { return (const double*)(g + o_travel); }   // [L]
};

// batch of edges as seen by kernels.  The scalars N0, Ns0, Nm0, md0 stand for the arrays N, N_start,
// N_max, modes when those are null (N0 <= 0: N from the resolution), so a one-edge call passes its
// arguments by value.
struct EdgeBatchView {
    const double* q1;
    const double* q2;
    long long rs, ds;  // row / dim strides
    const int* idx1;
    const int* idx2;
    long long n;
    double res;
    const int* N;
    const int* N_start;
    const int* N_max;
    int incl;
    const int* modes;
    unsigned mask;
    unsigned char* out;
    int N0, Ns0, Nm0, md0;
};

__host__ __device__ inline int batch_N_start(const EdgeBatchView& B, long long e) {
    return B.N_start ? B.N_start[e] : B.Ns0;
}
__host__ __device__ inline int batch_N_max(const EdgeBatchView& B, long long e) {
    return B.N_max ? B.N_max[e] : B.Nm0;
}
__host__ __device__ inline int batch_mode(const EdgeBatchView& B, long long e) {
    return B.modes ? B.modes[e] : B.md0;
}

// What the persistent arm kernels read of a batch and its plan, in one argument: the fields of
// EdgeBatchView and PlanWork they use, under the same names (fewer scalar registers than both).
struct ArmQueue {
    const double* q1;
    const int* idx1;
    const int* modes;
    unsigned char* out;
    const int4* erec;
    const int* order;
    const long long* pass_start;
    unsigned long long* ticket;
    const double* dir;
    unsigned long long* stats;
    const long long* meta;
    long long rs;
    int ds, n, md0, incl;
    unsigned mask;
};

__host__ __device__ inline int batch_mode(const ArmQueue& B, long long e) {
    return B.modes ? B.modes[e] : B.md0;
}

// Per-edge record of a plan: N, slot count M, passes S, window start | full << 31.
__host__ __device__ inline int erec_w0(int4 r) { return r.w & 0x7fffffff; }
__host__ __device__ inline int erec_full(int4 r) { return (int)((unsigned)r.w >> 31); }

// workspace of a batched plan
struct PlanWork {
    int4* erec;        // [n] per-edge record (N, M, S, w0 | full << 31)
    int* order;        // [n] edges longest-first
    int* hist;         // [MAX_PASSES+1]
    int* cursor;       // [MAX_PASSES+1]
    long long* pass_start;  // [MAX_PASSES+1] slot offset of each pass (G * sum E)
    long long* meta;   // [8]: 0 total slots, 1 n_pass, 2 error count, 3 long edges, 5 largest pass count
    int* long_list;    // [n] edges with more passes than MAX_PASSES
    unsigned long long* ticket;  // [8 * 32]: one chunk counter per XCD, 256 bytes apart
    unsigned long long* stats;  // [8]
    double* dir;       // [n*D] (lane-per-state kinds) or null
};
