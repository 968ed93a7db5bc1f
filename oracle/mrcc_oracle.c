/*
 * mrcc_oracle.c — CPU oracle (FP64).  TEST INFRASTRUCTURE ONLY: the checker for the GPU engine and
 * the CPU baseline of bench.py; never part of the product path.  See mrcc_oracle.h for the list of
 * reference functions restated here.  Build: oracle/Makefile (-O2 -ffp-contract=off -fopenmp).
 */
#include "mrcc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

enum { K_DISKS = 1, K_PLANAR = 2, K_ARMS = 3 };
#define MAX_ARMS 8
#define MAX_DOF 8

struct or_scene {
    int kind;
    int D;
    /* disks */
    int nr, rd;
    double* rr;             /* robot radii */
    double* rr_pow2;        /* pow(r, 2.0): Python's r**2 */
    int ns;
    double *sc, *sr;        /* sphere obstacles */
    int nrect;
    double *rlo, *rhi;      /* rectangle bounds c -/+ b/2 */
    /* planar */
    double* rh;             /* robot half extents [nr,2] */
    int no;
    double* ob;             /* obstacles [no,8]: x y ux uy vx vy h0 h1 */
    /* arms */
    int na, dof, S, nb;
    double base[MAX_ARMS][12];
    double dh[MAX_ARMS][MAX_DOF][5]; /* d a cos(alpha) sin(alpha) alpha */
    int is_chain;           /* 1: joints are T <- T Rz(q) C with C below */
    double chain[MAX_ARMS][MAX_DOF][12]; /* C rows: R00 R01 R02 t0, R10 .. t1, R20 .. t2 */
    uint8_t prism[MAX_ARMS * MAX_DOF];    /* 1: prismatic chain joint, T <- T Tz(q) C */
    double travel[MAX_ARMS * MAX_DOF];    /* prismatic joints: |q| <= travel, else a collision */
    int* sph_link;          /* [S] global link index arm*dof+link */
    double* sph;            /* [S,4] local xyz r */
    int link_first[MAX_ARMS * MAX_DOF + 1];
    double* box;            /* [nb,6] lo xyz hi xyz */
    double* boxch;          /* [nb,6] centre xyz half xyz */
    int nm;
    int nslot;              /* mode-box slots: na held (slot a: arm a's end-effector) + resting */
    int32_t* m_active;      /* [nm] */
    uint8_t* m_grasp;       /* [nm*nslot] slot present in the mode */
    double* m_rel;          /* [nm*nslot*12] held: pose in the end-effector frame; resting: world */
    double* m_half;         /* [nm*nslot*3] */
};
#define MAX_SLOTS 32

/* ------------------------------------------------------------------------------------------ */
/* sincos: Cody-Waite reduction by pi/2 (three-part constant, 33+33+53 bits) and Taylor series of
 * sin (to r^17) and cos (to r^18) on [-pi/4, pi/4], evaluated by Horner, no FMA. */
static const double TWO_OVER_PI = 0x1.45f306dc9c883p-1;
static const double PIO2_1 = 0x1.921fb54400000p+0;
static const double PIO2_2 = 0x1.0b4611a600000p-34;
static const double PIO2_3 = 0x1.3198a2e037073p-69;
static const double S3 = -0x1.5555555555555p-3, S5 = 0x1.1111111111111p-7,
                    S7 = -0x1.a01a01a01a01ap-13, S9 = 0x1.71de3a556c734p-19,
                    S11 = -0x1.ae64567f544e4p-26, S13 = 0x1.6124613a86d09p-33,
                    S15 = -0x1.ae7f3e733b81fp-41, S17 = 0x1.952c77030ad4ap-49;
static const double C4 = 0x1.5555555555555p-5, C6 = -0x1.6c16c16c16c17p-10,
                    C8 = 0x1.a01a01a01a01ap-16, C10 = -0x1.27e4fb7789f5cp-22,
                    C12 = 0x1.1eed8eff8d898p-29, C14 = -0x1.93974a8c07c9dp-37,
                    C16 = 0x1.ae7f3e733b81fp-45, C18 = -0x1.6827863b97d97p-53;

void or_sincos(double x, double* s_out, double* c_out) {
    double kd = rint(x * TWO_OVER_PI);
    double r = ((x - kd * PIO2_1) - kd * PIO2_2) - kd * PIO2_3;
    double r2 = r * r;
    double ps = S17;
    ps = S15 + r2 * ps;
    ps = S13 + r2 * ps;
    ps = S11 + r2 * ps;
    ps = S9 + r2 * ps;
    ps = S7 + r2 * ps;
    ps = S5 + r2 * ps;
    ps = S3 + r2 * ps;
    double sn = r + (r * r2) * ps;
    double pc = C18;
    pc = C16 + r2 * pc;
    pc = C14 + r2 * pc;
    pc = C12 + r2 * pc;
    pc = C10 + r2 * pc;
    pc = C8 + r2 * pc;
    pc = C6 + r2 * pc;
    pc = C4 + r2 * pc;
    double cs = (1.0 - 0.5 * r2) + (r2 * r2) * pc;
    double qd = kd - 4.0 * floor(kd * 0.25);
    int q = (int)qd;
    switch (q & 3) {
        case 0: *s_out = sn; *c_out = cs; break;
        case 1: *s_out = cs; *c_out = -sn; break;
        case 2: *s_out = -sn; *c_out = -cs; break;
        default: *s_out = -cs; *c_out = sn; break;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* planning_env.py:34-51: breadth-first midpoint order over [0, N-1]. */
int or_bfs_indices(int N, int32_t* out) {
    if (N < 1) return -1;
    int32_t* qs = (int32_t*)malloc(sizeof(int32_t) * 2 * (size_t)N);
    int32_t* qe = qs + N;
    int head = 0, tail = 0, idx = 0;
    qs[tail] = 0; qe[tail] = N - 1; tail++;
    while (head < tail) {
        int s = qs[head], e = qe[head]; head++;
        int mid = (s + e) / 2;            /* s, e >= 0: floor division */
        out[idx++] = mid;
        if (s <= mid - 1) { qs[tail] = s; qe[tail] = mid - 1; tail++; }
        if (mid + 1 <= e) { qs[tail] = mid + 1; qe[tail] = e; tail++; }
    }
    free(qs);
    return 0;
}

/* abstract_env.py:321-323 with config_dist(q1, q2, "max") = max_k |q1_k - q2_k|
 * (configuration.py:328-329 -> compute_abs_max_reduction :174-197); int() truncates. */
int or_edge_N(const double* q1, const double* q2, int D, double res) {
    double m = fabs(q1[0] - q2[0]);
    for (int k = 1; k < D; ++k) {
        double a = fabs(q1[k] - q2[k]);
        if (a > m) m = a;
    }
    double t = m / res;
    int N = (int)t + 1;
    return N < 2 ? 2 : N;
}

/* Number of states abstract_env.py:343-352 hands to is_collision_free. */
int64_t or_edge_num_states(int N, int N_start, int N_max, int include_endpoints) {
    if (N < 1 || N_start < 0 || N_start > N) return -1;
    int hi = (N_max < 0 || N_max > N) ? N : N_max;
    if (hi <= N_start) return 0;
    if (include_endpoints) return hi - N_start;
    int32_t* idx = (int32_t*)malloc(sizeof(int32_t) * (size_t)N);
    or_bfs_indices(N, idx);
    int64_t c = 0;
    for (int t = N_start; t < hi; ++t)
        if (idx[t] != 0 && idx[t] != N - 1) c++;
    free(idx);
    return c;
}

/* ------------------------------------------------------------------------------------------ */
/* numpy helpers */
static double (*volatile pow_fn)(double, double) = pow; /* keep the libm call (Python r**2) */

/* np.dot(v, v) for n < 16 with OpenBLAS' scalar tail loop: a sequential fused chain. */
static double np_sqnorm(const double* v, int n) {
    double s = 0.0;
    for (int k = 0; k < n; ++k) s = fma(v[k], v[k], s);
    return s;
}

/* np.sum (add.reduce, identity 0 + numpy pairwise_sum) for n <= 128 */
static double np_sum(const double* a, int n) {
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

/* NaN-propagating clip (numpy _NPY_MIN(_NPY_MAX(x, lo), hi)) */
static double np_clip(double x, double lo, double hi) {
    double m = isnan(x) ? x : (x > lo ? x : lo);
    return isnan(m) ? m : (m < hi ? m : hi);
}

/* ------------------------------------------------------------------------------------------ */
or_scene* or_scene_disks(int n_robots, int robot_dim, const double* robot_radius, int n_spheres,
                         const double* sphere_center, const double* sphere_radius, int n_rects,
                         const double* rect_center, const double* rect_size) {
    if (n_robots < 1 || robot_dim < 1 || robot_dim > 12) return NULL;
    or_scene* s = (or_scene*)calloc(1, sizeof(or_scene));
    s->kind = K_DISKS;
    s->nr = n_robots;
    s->rd = robot_dim;
    s->D = n_robots * robot_dim;
    s->rr = (double*)malloc(sizeof(double) * n_robots);
    s->rr_pow2 = (double*)malloc(sizeof(double) * n_robots);
    for (int i = 0; i < n_robots; ++i) {
        s->rr[i] = robot_radius[i];
        s->rr_pow2[i] = pow_fn(robot_radius[i], 2.0);
    }
    s->ns = n_spheres;
    s->sc = (double*)malloc(sizeof(double) * (size_t)(n_spheres * robot_dim + 1));
    s->sr = (double*)malloc(sizeof(double) * (size_t)(n_spheres + 1));
    if (n_spheres) {
        memcpy(s->sc, sphere_center, sizeof(double) * n_spheres * robot_dim);
        memcpy(s->sr, sphere_radius, sizeof(double) * n_spheres);
    }
    s->nrect = n_rects;
    s->rlo = (double*)malloc(sizeof(double) * (size_t)(n_rects * robot_dim + 1));
    s->rhi = (double*)malloc(sizeof(double) * (size_t)(n_rects * robot_dim + 1));
    for (int o = 0; o < n_rects; ++o)
        for (int k = 0; k < robot_dim; ++k) {
            double c = rect_center[o * robot_dim + k], h = rect_size[o * robot_dim + k] / 2;
            s->rlo[o * robot_dim + k] = c - h; /* abstract_env.py:66 */
            s->rhi[o * robot_dim + k] = c + h; /* abstract_env.py:67 */
        }
    return s;
}

/* abstract_env.py:255-276 (robot pairs :261-268, obstacles :270-274); mask rule of
 * rai_base_env.py:560-586: a pair counts iff every robot in it is in the mask. */
static int disk_free(const or_scene* s, const double* q, uint32_t mask) {
    const int rd = s->rd;
    double d[12];
    for (int i = 0; i < s->nr; ++i)
        for (int j = i + 1; j < s->nr; ++j) {
            if (!((mask >> i) & 1u) || !((mask >> j) & 1u)) continue;
            for (int k = 0; k < rd; ++k) d[k] = q[i * rd + k] - q[j * rd + k];
            if (sqrt(np_sqnorm(d, rd)) < s->rr[i] + s->rr[j]) return 0;
        }
    for (int i = 0; i < s->nr; ++i) {
        if (!((mask >> i) & 1u)) continue;
        const double* p = q + i * rd;
        for (int o = 0; o < s->ns; ++o) { /* Sphere.collides_with_sphere :46-49 */
            for (int k = 0; k < rd; ++k) d[k] = s->sc[o * rd + k] - p[k];
            if (sqrt(np_sqnorm(d, rd)) < s->sr[o] + s->rr[i]) return 0;
        }
        for (int o = 0; o < s->nrect; ++o) { /* Rectangle.collides_with_sphere :69-84 */
            for (int k = 0; k < rd; ++k) {
                double cp = np_clip(p[k], s->rlo[o * rd + k], s->rhi[o * rd + k]);
                double e = cp - p[k];
                d[k] = e * e;
            }
            if (np_sum(d, rd) <= s->rr_pow2[i]) return 0;
        }
    }
    return 1;
}

/* ------------------------------------------------------------------------------------------ */
/* planar oriented boxes */
or_scene* or_scene_planar(int n_robots, const double* robot_half, int n_obs,
                          const double* obs_pose, const double* obs_half) {
    if (n_robots < 1 || n_robots > 16) return NULL;
    or_scene* s = (or_scene*)calloc(1, sizeof(or_scene));
    s->kind = K_PLANAR;
    s->nr = n_robots;
    s->D = 3 * n_robots;
    s->rh = (double*)malloc(sizeof(double) * 2 * n_robots);
    memcpy(s->rh, robot_half, sizeof(double) * 2 * n_robots);
    s->no = n_obs;
    s->ob = (double*)malloc(sizeof(double) * 8 * (size_t)(n_obs + 1));
    for (int o = 0; o < n_obs; ++o) {
        double sn, cs;
        or_sincos(obs_pose[3 * o + 2], &sn, &cs);
        double* b = s->ob + 8 * o;
        b[0] = obs_pose[3 * o];
        b[1] = obs_pose[3 * o + 1];
        b[2] = cs; b[3] = sn;      /* u */
        b[4] = -sn; b[5] = cs;     /* v */
        b[6] = obs_half[2 * o];
        b[7] = obs_half[2 * o + 1];
    }
    return s;
}

/* closed 2D separating-axis test; boxes as [x y ux uy vx vy h0 h1] */
static int obb2_overlap(const double* A, const double* B) {
    double tx = B[0] - A[0], ty = B[1] - A[1];
    double C = A[2] * B[2] + A[3] * B[3]; /* ua.ub */
    double S = A[4] * B[2] + A[5] * B[3]; /* va.ub */
    double aC = fabs(C), aS = fabs(S);
    double tA0 = tx * A[2] + ty * A[3];
    double tA1 = tx * A[4] + ty * A[5];
    double tB0 = tx * B[2] + ty * B[3];
    double tB1 = tx * B[4] + ty * B[5];
    double s0 = fabs(tA0) - ((A[6] + B[6] * aC) + B[7] * aS);
    double s1 = fabs(tA1) - ((A[7] + B[6] * aS) + B[7] * aC);
    double s2 = fabs(tB0) - ((B[6] + A[6] * aC) + A[7] * aS);
    double s3 = fabs(tB1) - ((B[7] + A[6] * aS) + A[7] * aC);
    return s0 <= 0.0 && s1 <= 0.0 && s2 <= 0.0 && s3 <= 0.0;
}

static int all_finite(const double* q, int D) {
    for (int k = 0; k < D; ++k)
        if (!isfinite(q[k])) return 0;
    return 1;
}

static int planar_free(const or_scene* s, const double* q, uint32_t mask) {
    double R[16][8];
    if (!all_finite(q, s->D)) return 0;
    for (int i = 0; i < s->nr; ++i) {
        double sn, cs;
        or_sincos(q[3 * i + 2], &sn, &cs);
        R[i][0] = q[3 * i];
        R[i][1] = q[3 * i + 1];
        R[i][2] = cs; R[i][3] = sn;
        R[i][4] = -sn; R[i][5] = cs;
        R[i][6] = s->rh[2 * i];
        R[i][7] = s->rh[2 * i + 1];
    }
    for (int i = 0; i < s->nr; ++i)
        for (int j = i + 1; j < s->nr; ++j) {
            if (!((mask >> i) & 1u) || !((mask >> j) & 1u)) continue;
            if (obb2_overlap(R[i], R[j])) return 0;
        }
    for (int i = 0; i < s->nr; ++i) {
        if (!((mask >> i) & 1u)) continue;
        for (int o = 0; o < s->no; ++o)
            if (obb2_overlap(R[i], s->ob + 8 * o)) return 0;
    }
    return 1;
}

/* ------------------------------------------------------------------------------------------ */
/* arms */
or_scene* or_scene_arms(int n_arms, int dof, const double* base, const double* dh,
                        const int32_t* link_sphere_count, const double* sphere_local,
                        int n_boxes, const double* boxes) {
    if (n_arms < 1 || n_arms > MAX_ARMS || dof < 1 || dof > MAX_DOF) return NULL;
    or_scene* s = (or_scene*)calloc(1, sizeof(or_scene));
    s->kind = K_ARMS;
    s->na = n_arms;
    s->dof = dof;
    s->D = n_arms * dof;
    int S = 0;
    for (int l = 0; l < n_arms * dof; ++l) {
        s->link_first[l] = S;
        S += link_sphere_count[l];
    }
    s->link_first[n_arms * dof] = S;
    s->S = S;
    s->sph = (double*)malloc(sizeof(double) * 4 * (size_t)(S + 1));
    s->sph_link = (int*)malloc(sizeof(int) * (size_t)(S + 1));
    memcpy(s->sph, sphere_local, sizeof(double) * 4 * S);
    for (int l = 0; l < n_arms * dof; ++l)
        for (int t = s->link_first[l]; t < s->link_first[l + 1]; ++t) s->sph_link[t] = l;
    for (int a = 0; a < n_arms; ++a) {
        memcpy(s->base[a], base + 12 * a, sizeof(double) * 12);
        for (int k = 0; k < dof; ++k) {
            const double* p = dh + 3 * (a * dof + k);
            double sa, ca;
            or_sincos(p[2], &sa, &ca);
            s->dh[a][k][0] = p[0];
            s->dh[a][k][1] = p[1];
            s->dh[a][k][2] = ca;
            s->dh[a][k][3] = sa;
            s->dh[a][k][4] = p[2];
        }
    }
    s->nb = n_boxes;
    s->box = (double*)malloc(sizeof(double) * 6 * (size_t)(n_boxes + 1));
    s->boxch = (double*)malloc(sizeof(double) * 6 * (size_t)(n_boxes + 1));
    if (n_boxes) memcpy(s->boxch, boxes, sizeof(double) * 6 * n_boxes);
    for (int b = 0; b < n_boxes; ++b)
        for (int k = 0; k < 3; ++k) {
            s->box[6 * b + k] = boxes[6 * b + k] - boxes[6 * b + 3 + k];
            s->box[6 * b + 3 + k] = boxes[6 * b + k] + boxes[6 * b + 3 + k];
        }
    int32_t na = n_arms;
    uint8_t g[MAX_ARMS] = {0};
    double rel[MAX_ARMS * 12] = {0}, half[MAX_ARMS * 3] = {0};
    or_scene_set_modes(s, 1, &na, 0, g, rel, half);
    return s;
}

or_scene* or_scene_arm_chains(int n_arms, int dof, const double* base, const double* joint_fixed,
                              const int32_t* link_sphere_count, const double* sphere_local,
                              int n_boxes, const double* boxes) {
    if (n_arms < 1 || n_arms > MAX_ARMS || dof < 1 || dof > MAX_DOF) return NULL;
    /* the DH slots are unused: build with alpha = 0, then switch the joints to C */
    double* dh0 = (double*)calloc(3 * (size_t)n_arms * dof, sizeof(double));
    or_scene* s = or_scene_arms(n_arms, dof, base, dh0, link_sphere_count, sphere_local, n_boxes, boxes);
    free(dh0);
    if (!s) return NULL;
    s->is_chain = 1;
    for (int a = 0; a < n_arms; ++a)
        for (int k = 0; k < dof; ++k) memcpy(s->chain[a][k], joint_fixed + 12 * (a * dof + k), sizeof(double) * 12);
    return s;
}

int or_scene_set_joint_types(or_scene* s, const uint8_t* types, const double* travel) {
    if (s->kind != K_ARMS) return -1;
    const int L = s->na * s->dof;
    for (int l = 0; l < L; ++l) {
        if (types[l] > 1 || (types[l] == 1 && (!s->is_chain || !(travel[l] >= 0.0) || !isfinite(travel[l]))))
            return -1;
    }
    for (int l = 0; l < L; ++l) {
        s->prism[l] = types[l];
        s->travel[l] = types[l] ? travel[l] : 0.0;
    }
    return 0;
}

int or_scene_set_modes(or_scene* s, int n_modes, const int32_t* n_active, int n_rest,
                       const uint8_t* has_grasp, const double* grasp_rel,
                       const double* grasp_half) {
    if (s->kind != K_ARMS || n_modes < 1 || n_rest < 0 || s->na + n_rest > MAX_SLOTS) return -1;
    free(s->m_active); free(s->m_grasp); free(s->m_rel); free(s->m_half);
    s->nm = n_modes;
    s->nslot = s->na + n_rest;
    size_t nma = (size_t)n_modes * s->nslot;
    s->m_active = (int32_t*)malloc(sizeof(int32_t) * n_modes);
    s->m_grasp = (uint8_t*)malloc(nma);
    s->m_rel = (double*)malloc(sizeof(double) * 12 * nma);
    s->m_half = (double*)malloc(sizeof(double) * 3 * nma);
    memcpy(s->m_active, n_active, sizeof(int32_t) * n_modes);
    memcpy(s->m_grasp, has_grasp, nma);
    memcpy(s->m_rel, grasp_rel, sizeof(double) * 12 * nma);
    memcpy(s->m_half, grasp_half, sizeof(double) * 3 * nma);
    return 0;
}

/* T = T * Rz(q) C for one revolute chain joint: A = Rz(q) C, rows A0c = cs C0c - sn C1c,
 * A1c = sn C0c + cs C1c, A2c = C2c, then the entries of T * A as dh_step.  A prismatic joint is
 * T = T * Tz(q) C: the same rows with cs = 1, sn = 0, and A23 = C23 + q. */
static void chain_step(double X[12], const double C[12], double q, int prismatic) {
    double sn = 0.0, cs = 1.0;
    if (!prismatic) or_sincos(q, &sn, &cs);
    double A[3][4];
    for (int c = 0; c < 4; ++c) {
        A[0][c] = cs * C[c] - sn * C[4 + c];
        A[1][c] = sn * C[c] + cs * C[4 + c];
        A[2][c] = C[8 + c];
    }
    if (prismatic) A[2][3] = C[11] + q;
    double Y[12];
    for (int r = 0; r < 3; ++r) {
        const double* R = X + 3 * r;
        for (int c = 0; c < 3; ++c)
            Y[3 * r + c] = (R[0] * A[0][c] + R[1] * A[1][c]) + R[2] * A[2][c];
        Y[9 + r] = ((R[0] * A[0][3] + R[1] * A[1][3]) + R[2] * A[2][3]) + X[9 + r];
    }
    memcpy(X, Y, sizeof(Y));
}

/* T = T * A(q) for one DH joint; X = [R row-major | t] */
static void dh_step(double X[12], const double p[5], double q) {
    double sn, cs;
    or_sincos(q, &sn, &cs);
    const double d = p[0], a = p[1], ca = p[2], sa = p[3];
    double A[3][4];
    A[0][0] = cs; A[0][1] = -(sn * ca); A[0][2] = sn * sa; A[0][3] = a * cs;
    A[1][0] = sn; A[1][1] = cs * ca; A[1][2] = -(cs * sa); A[1][3] = a * sn;
    A[2][0] = 0.0; A[2][1] = sa; A[2][2] = ca; A[2][3] = d;
    double Y[12];
    for (int r = 0; r < 3; ++r) {
        const double* R = X + 3 * r;
        for (int c = 0; c < 3; ++c)
            Y[3 * r + c] = (R[0] * A[0][c] + R[1] * A[1][c]) + R[2] * A[2][c];
        Y[9 + r] = ((R[0] * A[0][3] + R[1] * A[1][3]) + R[2] * A[2][3]) + X[9 + r];
    }
    memcpy(X, Y, sizeof(Y));
}

int or_arm_fk(const or_scene* s, const double* q, double* out) {
    if (s->kind != K_ARMS) return -1;
    for (int a = 0; a < s->na; ++a) {
        double X[12];
        memcpy(X, s->base[a], sizeof(X));
        for (int k = 0; k < s->dof; ++k) {
            if (s->is_chain) chain_step(X, s->chain[a][k], q[a * s->dof + k], s->prism[a * s->dof + k]);
            else dh_step(X, s->dh[a][k], q[a * s->dof + k]);
            memcpy(out + 12 * (a * s->dof + k), X, sizeof(X));
        }
    }
    return 0;
}

static void xform_point(const double X[12], const double l[3], double c[3]) {
    for (int r = 0; r < 3; ++r)
        c[r] = ((X[3 * r] * l[0] + X[3 * r + 1] * l[1]) + X[3 * r + 2] * l[2]) + X[9 + r];
}

static void compose(const double X[12], const double Rel[12], double G[12]) {
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c)
            G[3 * r + c] = (X[3 * r] * Rel[c] + X[3 * r + 1] * Rel[3 + c]) + X[3 * r + 2] * Rel[6 + c];
        G[9 + r] = ((X[3 * r] * Rel[9] + X[3 * r + 1] * Rel[10]) + X[3 * r + 2] * Rel[11]) + X[9 + r];
    }
}

static int sph_box(const double c[3], double r, const double* bx) {
    double e[3];
    for (int k = 0; k < 3; ++k) {
        double m = c[k] >= bx[k] ? c[k] : bx[k];
        m = m <= bx[3 + k] ? m : bx[3 + k];
        e[k] = m - c[k];
    }
    return ((e[0] * e[0] + e[1] * e[1]) + e[2] * e[2]) <= r * r;
}

static int sph_sph(const double a[3], double ra, const double b[3], double rb) {
    double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
    double R = ra + rb;
    return ((d0 * d0 + d1 * d1) + d2 * d2) <= R * R;
}

/* sphere vs oriented box G = [R | t], half h */
static int sph_obb(const double c[3], double r, const double G[12], const double h[3]) {
    double w0 = c[0] - G[9], w1 = c[1] - G[10], w2 = c[2] - G[11];
    double e[3];
    for (int k = 0; k < 3; ++k) {
        double l = (G[k] * w0 + G[3 + k] * w1) + G[6 + k] * w2;
        double m = l >= -h[k] ? l : -h[k];
        m = m <= h[k] ? m : h[k];
        e[k] = m - l;
    }
    return ((e[0] * e[0] + e[1] * e[1]) + e[2] * e[2]) <= r * r;
}

/* closed 15-axis separating-axis test of boxes A = [R | t] half a and B half b */
static int obb3_overlap(const double A[12], const double a[3], const double B[12], const double b[3]) {
    double Rm[3][3], AR[3][3], t[3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            Rm[i][j] = (A[i] * B[j] + A[3 + i] * B[3 + j]) + A[6 + i] * B[6 + j];
            AR[i][j] = fabs(Rm[i][j]);
        }
    double D0 = B[9] - A[9], D1 = B[10] - A[10], D2 = B[11] - A[11];
    for (int i = 0; i < 3; ++i) t[i] = (A[i] * D0 + A[3 + i] * D1) + A[6 + i] * D2;
    for (int i = 0; i < 3; ++i) {
        double rb = (b[0] * AR[i][0] + b[1] * AR[i][1]) + b[2] * AR[i][2];
        if (fabs(t[i]) - (a[i] + rb) > 0.0) return 0;
    }
    for (int j = 0; j < 3; ++j) {
        double ra = (a[0] * AR[0][j] + a[1] * AR[1][j]) + a[2] * AR[2][j];
        double tp = (t[0] * Rm[0][j] + t[1] * Rm[1][j]) + t[2] * Rm[2][j];
        if (fabs(tp) - (ra + b[j]) > 0.0) return 0;
    }
    for (int i = 0; i < 3; ++i) {
        int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
        for (int j = 0; j < 3; ++j) {
            int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            double tp = t[i2] * Rm[i1][j] - t[i1] * Rm[i2][j];
            double rr = ((a[i1] * AR[i2][j] + a[i2] * AR[i1][j]) + b[j1] * AR[i][j2]) + b[j2] * AR[i][j1];
            if (fabs(tp) - rr > 0.0) return 0;
        }
    }
    return 1;
}

static const double IDENT12[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};

static int arm_free(const or_scene* s, const double* q, int mode, uint32_t mask) {
    if (mode < 0 || mode >= s->nm) return -1;
    const int na = s->m_active[mode], dof = s->dof, L = s->na * dof, ns = s->nslot;
    /* the coordinates of the mode's arms: non-finite, or a prismatic joint beyond its travel, is a
     * collision */
    if (!all_finite(q, na * dof)) return 0;
    for (int l = 0; l < na * dof; ++l)
        if (s->prism[l] && fabs(q[l]) > s->travel[l]) return 0;
    double* X = (double*)malloc(sizeof(double) * 12 * (size_t)L);
    double* C = (double*)malloc(sizeof(double) * 3 * (size_t)(s->S + 1));
    double G[MAX_SLOTS][12];
    int present[MAX_SLOTS];
    int free_ = 1;
    or_arm_fk(s, q, X);
    for (int t = 0; t < s->S; ++t) xform_point(X + 12 * s->sph_link[t], s->sph + 4 * t, C + 3 * t);
    /* mode boxes: slot a < na is held by arm a (pose T_ee(q) T_rel), slot na + j rests at a world
     * pose.  A held box counts when its arm is active and in the mask; a resting box always. */
    for (int a = 0; a < ns; ++a) {
        const int held = a < s->na;
        present[a] = s->m_grasp[mode * ns + a] && (!held || (a < na && ((mask >> a) & 1u)));
        if (!present[a]) continue;
        const double* rel = s->m_rel + 12 * (mode * ns + a);
        if (held) compose(X + 12 * (a * dof + dof - 1), rel, G[a]);
        else memcpy(G[a], rel, sizeof(double) * 12);
    }
    for (int a = 0; a < na && free_; ++a) {
        if (!((mask >> a) & 1u)) continue;
        const int s0 = s->link_first[a * dof], s1 = s->link_first[(a + 1) * dof];
        /* spheres vs boxes */
        for (int t = s0; t < s1 && free_; ++t)
            for (int b = 0; b < s->nb; ++b)
                if (sph_box(C + 3 * t, s->sph[4 * t + 3], s->box + 6 * b)) { free_ = 0; break; }
        /* self: links at least two apart */
        for (int t = s0; t < s1 && free_; ++t) {
            int lt = s->sph_link[t] - a * dof;
            int u0 = (lt + 2 < dof) ? s->link_first[a * dof + lt + 2] : s1;
            for (int u = u0; u < s1; ++u)
                if (sph_sph(C + 3 * t, s->sph[4 * t + 3], C + 3 * u, s->sph[4 * u + 3])) { free_ = 0; break; }
        }
    }
    /* spheres of two different arms */
    for (int a = 0; a < na && free_; ++a) {
        if (!((mask >> a) & 1u)) continue;
        for (int b = a + 1; b < na && free_; ++b) {
            if (!((mask >> b) & 1u)) continue;
            const int a0 = s->link_first[a * dof], a1 = s->link_first[(a + 1) * dof];
            const int b0 = s->link_first[b * dof], b1 = s->link_first[(b + 1) * dof];
            for (int t = a0; t < a1 && free_; ++t)
                for (int u = b0; u < b1; ++u)
                    if (sph_sph(C + 3 * t, s->sph[4 * t + 3], C + 3 * u, s->sph[4 * u + 3])) { free_ = 0; break; }
        }
    }
    /* held boxes vs obstacle boxes (resting boxes do not move) */
    for (int a = 0; a < s->na && a < ns && free_; ++a) {
        if (!present[a]) continue;
        const double* h = s->m_half + 3 * (mode * ns + a);
        for (int b = 0; b < s->nb; ++b) {
            double Bx[12];
            memcpy(Bx, IDENT12, sizeof(Bx));
            for (int k = 0; k < 3; ++k) Bx[9 + k] = s->boxch[6 * b + k];
            if (obb3_overlap(Bx, s->boxch + 6 * b + 3, G[a], h)) { free_ = 0; break; }
        }
    }
    /* mode boxes vs the spheres of the active arms in the mask, except, for a held box, those of
     * its holding link: the attachment deactivates only the (object, holding link) pair
     * (pinocchio_env.py:266-272; rai attach + setContact(-1), rai_base_env.py:801-805) */
    for (int a = 0; a < ns && free_; ++a) {
        if (!present[a]) continue;
        const double* h = s->m_half + 3 * (mode * ns + a);
        for (int b = 0; b < na && free_; ++b) {
            if (!((mask >> b) & 1u)) continue;
            const int t0 = s->link_first[b * dof];
            const int t1 = (a == b) ? s->link_first[b * dof + dof - 1] : s->link_first[(b + 1) * dof];
            for (int t = t0; t < t1; ++t)
                if (sph_obb(C + 3 * t, s->sph[4 * t + 3], G[a], h)) { free_ = 0; break; }
        }
    }
    /* pairs of mode boxes with at least one held box (resting boxes do not move) */
    for (int a = 0; a < s->na && a < ns && free_; ++a) {
        if (!present[a]) continue;
        for (int b = a + 1; b < ns; ++b) {
            if (!present[b]) continue;
            if (obb3_overlap(G[a], s->m_half + 3 * (mode * ns + a), G[b], s->m_half + 3 * (mode * ns + b))) {
                free_ = 0;
                break;
            }
        }
    }
    free(X);
    free(C);
    return free_;
}

/* ------------------------------------------------------------------------------------------ */
void or_scene_free(or_scene* s) {
    if (!s) return;
    free(s->rr); free(s->rr_pow2); free(s->sc); free(s->sr); free(s->rlo); free(s->rhi);
    free(s->rh); free(s->ob); free(s->sph_link); free(s->sph); free(s->box); free(s->boxch);
    free(s->m_active); free(s->m_grasp); free(s->m_rel); free(s->m_half);
    free(s);
}

int or_scene_dim(const or_scene* s) { return s->D; }

int or_config_free(const or_scene* s, const double* q, int mode, uint32_t mask) {
    switch (s->kind) {
        case K_DISKS: return disk_free(s, q, mask);
        case K_PLANAR: return planar_free(s, q, mask);
        case K_ARMS: return arm_free(s, q, mode, mask);
    }
    return -1;
}

/* abstract_env.py:301-354 */
int or_edge_free(const or_scene* s, const double* q1, const double* q2, double res, int N,
                 int N_start, int N_max, int include_endpoints, int mode, uint32_t mask,
                 int64_t* n_checked) {
    const int D = s->D;
    if (N <= 0) N = or_edge_N(q1, q2, D, res);
    if (N_start < 0 || N_start > N) return -1; /* assert False (:325-326) */
    int hi = (N_max < 0 || N_max > N) ? N : N_max;
    int32_t* idx = (int32_t*)malloc(sizeof(int32_t) * (size_t)N);
    or_bfs_indices(N, idx);
    double dir[64], q[64];
    for (int k = 0; k < D; ++k) dir[k] = (q2[k] - q1[k]) / (double)(N - 1);
    int result = 1;
    for (int t = N_start; t < hi; ++t) {
        int i = idx[t];
        if (!include_endpoints && (i == 0 || i == N - 1)) continue;
        for (int k = 0; k < D; ++k) q[k] = q1[k] + dir[k] * (double)i;
        if (n_checked) (*n_checked)++;
        int f = or_config_free(s, q, mode, mask);
        if (f < 0) { result = f; break; }
        if (!f) { result = 0; break; }
    }
    free(idx);
    return result;
}

int or_check_configs(const or_scene* s, const double* q, int64_t n, int64_t row_stride,
                     const int32_t* modes, uint32_t mask, uint8_t* out, int threads) {
    int err = 0;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads) reduction(| : err)
#endif
    for (int64_t r = 0; r < n; ++r) {
        int f = or_config_free(s, q + r * row_stride, modes ? modes[r] : 0, mask);
        if (f < 0) err = 1;
        out[r] = (uint8_t)(f > 0);
    }
    (void)threads;
    return err ? -1 : 0;
}

int or_check_edges(const or_scene* s, const double* q1, const double* q2, int64_t n,
                   int64_t row_stride, double res, const int32_t* N, const int32_t* N_start,
                   const int32_t* N_max, int include_endpoints, const int32_t* modes,
                   uint32_t mask, uint8_t* out, int threads, int64_t* n_checked) {
    int err = 0;
    int64_t checked = 0;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads) reduction(| : err) reduction(+ : checked)
#endif
    for (int64_t e = 0; e < n; ++e) {
        int64_t c = 0;
        int f = or_edge_free(s, q1 + e * row_stride, q2 + e * row_stride, res, N ? N[e] : 0,
                             N_start ? N_start[e] : 0, N_max ? N_max[e] : -1, include_endpoints,
                             modes ? modes[e] : 0, mask, &c);
        if (f < 0) err = 1;
        out[e] = (uint8_t)(f > 0);
        checked += c;
    }
    (void)threads;
    if (n_checked) *n_checked = checked;
    return err ? -1 : 0;
}

/* ------------------------------------------------------------------------------------------ */
/* clearance cost: sum of max(0, m + r_i + r_j - d)^2 over the full pair set of mode 0 */
static double pen_sph_box(const double c[3], double r, const double* bx, double m, double g[3]) {
    double e[3];
    for (int k = 0; k < 3; ++k) {
        double v = c[k] >= bx[k] ? c[k] : bx[k];
        v = v <= bx[3 + k] ? v : bx[3 + k];
        e[k] = c[k] - v;
    }
    double d = sqrt((e[0] * e[0] + e[1] * e[1]) + e[2] * e[2]);
    double pen = m + r - d;
    if (pen <= 0.0) return 0.0;
    if (g && d > 0.0)
        for (int k = 0; k < 3; ++k) g[k] += -2.0 * pen * e[k] / d;
    return pen * pen;
}

static double pen_sph_sph(const double a[3], double ra, const double b[3], double rb, double m,
                          double ga[3], double gb[3]) {
    double e[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
    double d = sqrt((e[0] * e[0] + e[1] * e[1]) + e[2] * e[2]);
    double pen = m + ra + rb - d;
    if (pen <= 0.0) return 0.0;
    if (ga && d > 0.0)
        for (int k = 0; k < 3; ++k) {
            double f = -2.0 * pen * e[k] / d;
            ga[k] += f;
            gb[k] -= f;
        }
    return pen * pen;
}

static double clearance(const or_scene* s, const double* q, double m, double* grad) {
    const int dof = s->dof, na = s->na, L = na * dof;
    double* X = (double*)malloc(sizeof(double) * 12 * (size_t)L);
    double* C = (double*)malloc(sizeof(double) * 3 * (size_t)(s->S + 1));
    double* F = (double*)calloc(3 * (size_t)(s->S + 1), sizeof(double));
    or_arm_fk(s, q, X);
    for (int t = 0; t < s->S; ++t) xform_point(X + 12 * s->sph_link[t], s->sph + 4 * t, C + 3 * t);
    double cost = 0.0;
    for (int a = 0; a < na; ++a) {
        const int s0 = s->link_first[a * dof], s1 = s->link_first[(a + 1) * dof];
        for (int t = s0; t < s1; ++t) {
            for (int b = 0; b < s->nb; ++b)
                cost += pen_sph_box(C + 3 * t, s->sph[4 * t + 3], s->box + 6 * b, m, grad ? F + 3 * t : NULL);
            int lt = s->sph_link[t] - a * dof;
            int u0 = (lt + 2 < dof) ? s->link_first[a * dof + lt + 2] : s1;
            for (int u = u0; u < s1; ++u)
                cost += pen_sph_sph(C + 3 * t, s->sph[4 * t + 3], C + 3 * u, s->sph[4 * u + 3], m,
                                    grad ? F + 3 * t : NULL, grad ? F + 3 * u : NULL);
            for (int u = s1; u < s->S; ++u)
                cost += pen_sph_sph(C + 3 * t, s->sph[4 * t + 3], C + 3 * u, s->sph[4 * u + 3], m,
                                    grad ? F + 3 * t : NULL, grad ? F + 3 * u : NULL);
        }
    }
    if (grad) {
        /* d c / d q_k = z_{k-1} x (c - o_{k-1}) (revolute) or z_{k-1} (prismatic) for spheres on
         * links >= k of the same arm */
        for (int a = 0; a < na; ++a)
            for (int k = 0; k < dof; ++k) {
                const double* P = (k == 0) ? s->base[a] : X + 12 * (a * dof + k - 1);
                double z[3] = {P[2], P[5], P[8]}, o[3] = {P[9], P[10], P[11]};
                double g = 0.0;
                for (int t = s->link_first[a * dof + k]; t < s->link_first[(a + 1) * dof]; ++t) {
                    if (s->prism[a * dof + k]) {
                        g += z[0] * F[3 * t] + z[1] * F[3 * t + 1] + z[2] * F[3 * t + 2];
                        continue;
                    }
                    double w[3] = {C[3 * t] - o[0], C[3 * t + 1] - o[1], C[3 * t + 2] - o[2]};
                    double j0 = z[1] * w[2] - z[2] * w[1];
                    double j1 = z[2] * w[0] - z[0] * w[2];
                    double j2 = z[0] * w[1] - z[1] * w[0];
                    g += j0 * F[3 * t] + j1 * F[3 * t + 1] + j2 * F[3 * t + 2];
                }
                grad[a * dof + k] = g;
            }
    }
    free(X); free(C); free(F);
    return cost;
}

double or_clearance_cost(const or_scene* s, const double* q, double margin) {
    if (s->kind != K_ARMS) return NAN;
    return clearance(s, q, margin, NULL);
}

int or_clearance_grad(const or_scene* s, const double* q, double margin, double* grad) {
    if (s->kind != K_ARMS) return -1;
    clearance(s, q, margin, grad);
    return 0;
}

int or_clearance_cost_batch(const or_scene* s, const double* q, int64_t n, double margin,
                            double* cost, int threads) {
    if (s->kind != K_ARMS) return -1;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
#endif
    for (int64_t r = 0; r < n; ++r) cost[r] = clearance(s, q + r * s->D, margin, NULL);
    (void)threads;
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* configuration-pair costs (core/configuration.py) */

/* compute_sliced_euclidean_dists (:101-126): sum_squared = 0.0; sum_squared += d*d in coordinate
 * order (numba fastmath contracts the step); sqrt.  This is also np.linalg.norm of a short vector
 * (the fused chain), which test_config.py:110-126 compares for equality. */
static double cost_euclid(const double* d, int s, int e) {
    double acc = 0.0;
    for (int k = s; k < e; ++k) acc = fma(d[k], d[k], acc);
    return sqrt(acc);
}

/* np.max(np.abs(diff[:, s:e]), axis=1) (:503): the maximum; a NaN propagates */
static double cost_np_absmax(const double* d, int s, int e) {
    double m = fabs(d[s]);
    for (int k = s + 1; k < e; ++k) {
        const double u = fabs(d[k]);
        if (isnan(m)) break;
        if (isnan(u) || u > m) m = u;
    }
    return m;
}

/* compute_abs_max_reduction (:174-197): max_val = abs(d[0]); abs_val > max_val replaces it */
static double cost_seq_absmax(const double* d, int s, int e) {
    double m = fabs(d[s]);
    for (int k = s + 1; k < e; ++k) {
        const double u = fabs(d[k]);
        if (u > m) m = u;
    }
    return m;
}

int or_config_cost(const double* p, int64_t ps, const double* q, int64_t qs, int64_t n, int dim,
                   const int32_t* sl, int n_slices, int metric, int reduction, double w, double* out,
                   int threads) {
    if (dim < 1 || dim > 1024 || n_slices < 1 || n_slices > 64 || metric < 0 || metric > 2 ||
        reduction < 0 || reduction > 2)
        return -1;
    for (int i = 0; i < n_slices; ++i)
        if (sl[2 * i] < 0 || sl[2 * i + 1] > dim || sl[2 * i] > sl[2 * i + 1] ||
            (sl[2 * i] == sl[2 * i + 1] && metric != 1))
            return -1;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(threads)
#endif
    for (int64_t r = 0; r < n; ++r) {
        double d[1024];
        double v[64] = {0.0};
        for (int k = 0; k < dim; ++k) d[k] = p[r * ps + k] - q[r * qs + k];
        for (int i = 0; i < n_slices; ++i) {
            const int s = sl[2 * i], e = sl[2 * i + 1];
            v[i] = metric == 1 ? cost_euclid(d, s, e) : (metric == 0 ? cost_np_absmax(d, s, e)
                                                                      : cost_seq_absmax(d, s, e));
        }
        double res;
        if (reduction == 1) {  /* compute_sum_reduction (:129-147) */
            double sum_val = 0.0;
            for (int i = 0; i < n_slices; ++i) sum_val += v[i];
            res = sum_val;
        } else {
            /* compute_max_reduction (:200-219) / compute_max_sum_reduction (:150-171):
             * max_val = sum_val = dists[0]; dists[i] > max_val replaces it; sum_val += dists[i];
             * max_val + w * sum_val, contracted */
            double max_val = v[0], sum_val = v[0];
            for (int i = 1; i < n_slices; ++i) {
                if (v[i] > max_val) max_val = v[i];
                sum_val += v[i];
            }
            res = reduction == 2 ? max_val : fma(w, sum_val, max_val);
        }
        out[r] = res;
    }
    (void)threads;
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
int or_test_obb2(const double* A8, const double* B8) { return obb2_overlap(A8, B8); }
int or_test_obb3(const double* A12, const double* a3, const double* B12, const double* b3) {
    return obb3_overlap(A12, a3, B12, b3);
}
int or_test_sph_obb(const double* c3, double r, const double* G12, const double* h3) {
    return sph_obb(c3, r, G12, h3);
}
