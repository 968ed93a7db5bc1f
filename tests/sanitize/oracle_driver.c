/* oracle_driver.c — runs every entry point of the CPU oracle (oracle/mrcc_oracle.c) on small
 * seeded scenes of every kind, under AddressSanitizer and UndefinedBehaviorSanitizer (test
 * infrastructure; built and run by tests/test_sanitizers.py).  Exit status 0 and a line
 * "oracle driver ok" when every call returned a valid status; the sanitizers abort on the first
 * finding. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../oracle/mrcc_oracle.h"

static uint64_t g_state = 0x9e3779b97f4a7c15ull;
static double uni(double lo, double hi) {
    g_state = g_state * 6364136223846793005ull + 1442695040888963407ull;
    return lo + (hi - lo) * (double)(g_state >> 11) * (1.0 / 9007199254740992.0);
}

#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            return 1;                                                   \
        }                                                               \
    } while (0)

static int run_edges(or_scene* s, int D, int n, double res, const int32_t* modes) {
    double* q1 = malloc(sizeof(double) * D * n);
    double* q2 = malloc(sizeof(double) * D * n);
    uint8_t* out = malloc(n);
    int32_t* N = malloc(sizeof(int32_t) * n);
    int32_t* Ns = malloc(sizeof(int32_t) * n);
    int32_t* Nm = malloc(sizeof(int32_t) * n);
    for (int i = 0; i < D * n; ++i) {
        q1[i] = uni(-1.5, 1.5);
        q2[i] = q1[i] + uni(-0.3, 0.3);
    }
    for (int i = 0; i < n; ++i) {
        N[i] = (int32_t)uni(2, 40);
        Ns[i] = (int32_t)uni(0, N[i]);
        Nm[i] = Ns[i] + (int32_t)uni(0, 10);
    }
    int64_t checked = 0;
    int rc = or_check_edges(s, q1, q2, n, D, res, NULL, NULL, NULL, 0, modes, 0xffffffffu, out, 2, &checked);
    rc |= or_check_edges(s, q1, q2, n, D, res, N, Ns, Nm, 1, modes, 0x5u, out, 1, NULL);
    for (int i = 0; i < 8 && i < n; ++i) {
        int r = or_edge_free(s, q1 + D * i, q2 + D * i, res, 0, 0, -1, 0, modes ? modes[i] : 0, 0xffffffffu, &checked);
        if (r < 0) rc = 1;
    }
    free(q1); free(q2); free(out); free(N); free(Ns); free(Nm);
    return rc;
}

static int run_configs(or_scene* s, int D, int n, const int32_t* modes) {
    double* q = malloc(sizeof(double) * D * n);
    uint8_t* out = malloc(n);
    for (int i = 0; i < D * n; ++i) q[i] = uni(-2.5, 2.5);
    q[0] = NAN;  /* a non-finite coordinate */
    int rc = or_check_configs(s, q, n, D, modes, 0xffffffffu, out, 2);
    rc |= or_check_configs(s, q, n, D, modes, 0x2u, out, 1);
    free(q); free(out);
    return rc;
}

int main(void) {
    /* sincos and the index helpers */
    const double xs[] = {0.0, -0.0, 1e-300, 0.7853981633974483, 3.14159, -1e6, 1e15, INFINITY, NAN};
    for (size_t i = 0; i < sizeof(xs) / sizeof(xs[0]); ++i) {
        double s, c;
        or_sincos(xs[i], &s, &c);
    }
    for (int N = 1; N <= 700; ++N) {
        int32_t* idx = malloc(sizeof(int32_t) * N);
        CHECK(or_bfs_indices(N, idx) == 0);
        free(idx);
        for (int s = 0; s <= N; s += 7) (void)or_edge_num_states(N, s, s + 3, N & 1);
    }
    {
        double a[4] = {0, 0, 0, 0}, b[4] = {1, -2, 0.5, 3};
        CHECK(or_edge_N(a, b, 4, 0.01) > 0);
    }

    /* disks: abstract.test (abstract_env.py:363-421) */
    {
        const double rr[2] = {0.1, 0.1}, sc[2] = {0, 0}, sr[1] = {0.2}, rc[2] = {0, 0.4}, rs[2] = {0.5, 0.5};
        or_scene* s = or_scene_disks(2, 2, rr, 1, sc, sr, 1, rc, rs);
        CHECK(s && or_scene_dim(s) == 4);
        CHECK(run_configs(s, 4, 500, NULL) == 0);
        CHECK(run_edges(s, 4, 300, 0.01, NULL) == 0);
        or_scene_free(s);
    }
    /* planar oriented boxes */
    {
        double rh[6], op[30], oh[20];
        for (int i = 0; i < 6; ++i) rh[i] = uni(0.05, 0.2);
        for (int i = 0; i < 10; ++i) {
            op[3 * i] = uni(-2, 2); op[3 * i + 1] = uni(-2, 2); op[3 * i + 2] = uni(-3, 3);
            oh[2 * i] = uni(0.05, 0.3); oh[2 * i + 1] = uni(0.05, 0.3);
        }
        or_scene* s = or_scene_planar(3, rh, 10, op, oh);
        CHECK(s && or_scene_dim(s) == 9);
        CHECK(run_configs(s, 9, 500, NULL) == 0);
        CHECK(run_edges(s, 9, 300, 0.02, NULL) == 0);
        or_scene_free(s);
    }
    /* arms: DH and chain joints (prismatic among them), modes with held and resting boxes */
    for (int chain = 0; chain < 2; ++chain) {
        const int na = 3, dof = 4, L = na * dof, nb = 4, nm = 5, nrest = 2, ns = na + nrest;
        double base[36] = {0}, dh[36], jf[12 * 12], boxes[24];
        int32_t cnt[12];
        double sph[12 * 3 * 4];
        int S = 0;
        for (int a = 0; a < na; ++a) {
            double* B = base + 12 * a;
            B[0] = B[4] = B[8] = 1.0;
            B[9] = 0.4 * cos(2.1 * a); B[10] = 0.4 * sin(2.1 * a); B[11] = 0.05;
        }
        for (int l = 0; l < L; ++l) {
            dh[3 * l] = uni(0, 0.2); dh[3 * l + 1] = uni(0.05, 0.2); dh[3 * l + 2] = uni(-3, 3);
            double* C = jf + 12 * l;
            memset(C, 0, sizeof(double) * 12);
            C[0] = 1; C[5] = cos(dh[3 * l + 2]); C[6] = -sin(dh[3 * l + 2]);
            C[9] = sin(dh[3 * l + 2]); C[10] = cos(dh[3 * l + 2]);
            C[3] = dh[3 * l + 1]; C[11] = dh[3 * l];
            cnt[l] = 1 + (l % 3);
            for (int k = 0; k < cnt[l]; ++k, ++S) {
                sph[4 * S] = -uni(0, 0.15); sph[4 * S + 1] = uni(-0.02, 0.02);
                sph[4 * S + 2] = uni(-0.02, 0.02); sph[4 * S + 3] = uni(0.02, 0.06);
            }
        }
        for (int b = 0; b < nb; ++b) {
            boxes[6 * b] = uni(-0.4, 0.4); boxes[6 * b + 1] = uni(-0.4, 0.4); boxes[6 * b + 2] = uni(0, 0.3);
            boxes[6 * b + 3] = uni(0.03, 0.1); boxes[6 * b + 4] = uni(0.03, 0.1); boxes[6 * b + 5] = uni(0.03, 0.1);
        }
        or_scene* s = chain ? or_scene_arm_chains(na, dof, base, jf, cnt, sph, nb, boxes)
                            : or_scene_arms(na, dof, base, dh, cnt, sph, nb, boxes);
        CHECK(s && or_scene_dim(s) == L);
        if (chain) {
            uint8_t jt[12];
            double tr[12];
            for (int l = 0; l < L; ++l) { jt[l] = (l % 4) == 1; tr[l] = 0.3; }
            CHECK(or_scene_set_joint_types(s, jt, tr) == 0);
            jt[0] = 2;
            CHECK(or_scene_set_joint_types(s, jt, tr) != 0);
        }
        int32_t nact[5] = {3, 2, 1, 3, 3};
        uint8_t has[5 * 5];
        double rel[5 * 5 * 12], half[5 * 5 * 3];
        for (int i = 0; i < nm * ns; ++i) {
            has[i] = (i % 3) != 0 && (i % ns) < nact[i / ns] + nrest;
            memset(rel + 12 * i, 0, sizeof(double) * 12);
            rel[12 * i] = rel[12 * i + 4] = rel[12 * i + 8] = 1.0;
            rel[12 * i + 9] = uni(-0.3, 0.3); rel[12 * i + 10] = uni(-0.3, 0.3); rel[12 * i + 11] = uni(0, 0.3);
            for (int k = 0; k < 3; ++k) half[3 * i + k] = uni(0.02, 0.08);
        }
        CHECK(or_scene_set_modes(s, nm, nact, nrest, has, rel, half) == 0);
        int32_t modes[500];
        for (int i = 0; i < 500; ++i) modes[i] = i % nm;
        CHECK(run_configs(s, L, 500, modes) == 0);
        CHECK(run_edges(s, L, 300, 0.05, modes) == 0);
        double q[12], X[12 * 12], g[12];
        for (int l = 0; l < L; ++l) q[l] = uni(-0.3, 0.3);
        CHECK(or_arm_fk(s, q, X) == 0);
        CHECK(or_config_free(s, q, nm, 0xffffffffu) < 0);  /* mode out of range: an error */
        (void)or_clearance_cost(s, q, 0.02);
        CHECK(or_clearance_grad(s, q, 0.02, g) == 0);
        double qb[12 * 16], cost[16];
        for (int i = 0; i < 12 * 16; ++i) qb[i] = uni(-1, 1);
        CHECK(or_clearance_cost_batch(s, qb, 16, 0.02, cost, 2) == 0);
        or_scene_free(s);
    }
    /* configuration costs: every metric and reduction */
    {
        const int n = 64, dim = 7;
        double p[64 * 7], q[64 * 7], out[64];
        for (int i = 0; i < n * dim; ++i) { p[i] = uni(-2, 2); q[i] = uni(-2, 2); }
        const int32_t sl[6] = {0, 3, 3, 5, 5, 7};
        for (int m = 0; m < 3; ++m)
            for (int r = 0; r < 3; ++r) CHECK(or_config_cost(p, dim, q, dim, n, dim, sl, 3, m, r, 0.1, out, 2) == 0);
        CHECK(or_config_cost(p, dim, q, dim, n, dim, sl, 3, 7, 0, 0.1, out, 1) != 0);
    }
    /* single predicates */
    {
        double A[8] = {0, 0, 1, 0, 0, 1, 0.5, 0.2}, B[8] = {0.6, 0.1, 0.8, 0.6, -0.6, 0.8, 0.3, 0.1};
        (void)or_test_obb2(A, B);
        double A12[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0}, a3[3] = {0.1, 0.2, 0.3};
        double B12[12] = {0, -1, 0, 1, 0, 0, 0, 0, 1, 0.2, 0.1, 0.0}, b3[3] = {0.1, 0.1, 0.1};
        (void)or_test_obb3(A12, a3, B12, b3);
        double c3[3] = {0.3, 0, 0};
        (void)or_test_sph_obb(c3, 0.1, A12, a3);
    }
    printf("oracle driver ok\n");
    return 0;
}
