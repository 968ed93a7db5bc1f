// clearance.hip — differentiable clearance cost of arm scenes (binary32).
//
//   cost(q) = sum over the checked pair set of max(0, m + r_i + r_j - |c_i - c_j|)^2
//           + sum over spheres x boxes of max(0, m + r_i - dist(c_i, box))^2
// with the pair set of the boolean check (links >= 2 apart in one arm, all links of different
// arms), culled exactly with the links' bounding spheres.  One waypoint per wavefront.  The
// backward recomputes the forward kinematics, gathers the force on every sphere centre
// (d cost / d c) in LDS and maps it to the joints with
// d c / d q_k = z_{k-1} x (c - o_{k-1}) (revolute joints) or z_{k-1} (prismatic chain joints) for
// spheres on links >= k of the same arm.
#include "engine.hpp"

namespace {

struct ClrLds {
    float* cs;     // [D]
    float* sn;     // [D]
    float* X;      // [L*12]
    float4* sph;   // [S]
    float* F;      // [S*3] forces (backward)
    float* red;    // [64]
    float4* lk;    // [L] link bounding spheres (x, y, z, R)
    int* cand;     // [64] candidates of one culling round
    float* tf;     // [L*6] per link: sum of c x f and sum of f over its spheres (backward)
    int* cand2;    // [64] second-level candidates (link, box) of the clustered box test
};

__device__ inline ClrLds clr_lds(void* base, const ArmView& V) {
    ClrLds W;
    float* f = (float*)base;
    W.cs = f;
    W.sn = f + V.D;
    W.X = f + 2 * V.D;
    W.sph = (float4*)(((uintptr_t)(W.X + 12 * V.L) + 15) & ~uintptr_t(15));
    W.F = (float*)(W.sph + V.S);
    W.red = W.F + 3 * V.S;
    W.lk = (float4*)(((uintptr_t)(W.red + 64) + 15) & ~uintptr_t(15));
    W.cand = (int*)(W.lk + V.L);
    W.tf = (float*)(W.cand + 64);
    W.cand2 = (int*)(W.tf + 6 * V.L);
    return W;
}

// value of lane (lane & ~3) + J of the lane's quad (DPP quad_perm broadcast; the quad's lanes active)
template <int J>
__device__ inline float quad_bcast_f32(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), J * 0x55, 0xf, 0xf, false));
}

// forward kinematics in binary32 and sphere centres into LDS
__device__ inline void clr_fk(const ArmView& V, const ClrLds& W, const float* q) {
    const int lane = threadIdx.x;
    if (lane < V.D) {
        float s, c;
        sincosf(q[lane], &s, &c);
        const bool pr = (V.prism >> lane) & 1ull;  // prismatic: Tz(q) C = Rz(0) C + q e_z
        W.cs[lane] = pr ? 1.f : c;
        W.sn[lane] = pr ? 0.f : s;
    }
    __syncthreads();
    // lane (arm a, row r, column c) carries entry (r, c) of its arm's transform through the chain,
    // arms 5 at a time (60 lanes); the quad of a row shares the row's new entries (DPP broadcasts)
    {
        const int c = lane & 3, rr = (lane >> 2) % 3, ag = lane / 12;
        for (int a0 = 0; a0 < V.na; a0 += 5) {
            const int a = a0 + ag;
            if (ag < 5 && a < V.na) {
                const double* B = V.base + 12 * a;
                float R0 = (float)B[3 * rr], R1 = (float)B[3 * rr + 1], R2 = (float)B[3 * rr + 2], t = (float)B[9 + rr];
                for (int k = 0; k < V.dof; ++k) {
                    const int l = a * V.dof + k;
                    const float cs = W.cs[l], sn = W.sn[l];
                    float A0, A1, A2;  // column c of the joint's step A (3 x 4)
                    if (V.chain) {  // A = Rz(q) C (revolute), Tz(q) C (prismatic)
                        const double* C = V.dh + 12 * l;
                        const float c0 = (float)C[c], c1 = (float)C[4 + c];
                        A0 = cs * c0 - sn * c1;
                        A1 = sn * c0 + cs * c1;
                        A2 = (float)C[8 + c];
                        if (c == 3 && ((V.prism >> l) & 1ull)) A2 += q[l];
                    } else {
                        const double* pd = V.dh + 4 * l;
                        const float d = (float)pd[0], al = (float)pd[1], ca = (float)pd[2], sa = (float)pd[3];
                        A0 = c == 0 ? cs : (c == 1 ? -sn * ca : (c == 2 ? sn * sa : al * cs));
                        A1 = c == 0 ? sn : (c == 1 ? cs * ca : (c == 2 ? -cs * sa : al * sn));
                        A2 = c == 0 ? 0.f : (c == 1 ? sa : (c == 2 ? ca : d));
                    }
                    const float y = R0 * A0 + R1 * A1 + R2 * A2 + (c == 3 ? t : 0.f);
                    W.X[12 * l + (c < 3 ? 3 * rr + c : 9 + rr)] = y;
                    R0 = quad_bcast_f32<0>(y);
                    R1 = quad_bcast_f32<1>(y);
                    R2 = quad_bcast_f32<2>(y);
                    t = quad_bcast_f32<3>(y);
                }
            }
        }
    }
    __syncthreads();
    for (int s = lane; s < V.S; s += 64) {
        const float* X = W.X + 12 * V.sph_link[s];
        const double* l = V.sph64 + 4 * s;
        const float l0 = (float)l[0], l1 = (float)l[1], l2 = (float)l[2];
        W.sph[s] = make_float4(X[0] * l0 + X[1] * l1 + X[2] * l2 + X[9], X[3] * l0 + X[4] * l1 + X[5] * l2 + X[10],
                               X[6] * l0 + X[7] * l1 + X[8] * l2 + X[11], V.sph_r32[s]);
    }
    __syncthreads();
}

// Slack of the link-level cull (metres): covers the binary32 rounding of the link and sphere
// centres (micrometres) with a wide margin.
#define CLR_CULL_EPS 1e-4f

// One penetration term: cost += pen^2 for pen = R - d > 0, and the force on the first sphere
// (fx, fy, fz += g e, g = -2 pen / d) when `grad`.  The square root only for pairs within reach:
// d < R implies e2 < R^2 (1 + 2^-21), so the bound passes every pair with pen > 0.
__device__ inline bool clr_term(float ex, float ey, float ez, float R, bool grad, float& cost, float& gx,
                                float& gy, float& gz) {
    const float e2 = ex * ex + ey * ey + ez * ez;
    if (!(e2 <= R * R * 1.0001f)) return false;
    const float d = sqrtf(e2);
    const float pen = R - d;
    if (!(pen > 0.f)) return false;
    cost += pen * pen;
    if (!grad || !(d > 0.f)) return false;
    const float g = -2.f * pen / d;
    gx = g * ex;
    gy = g * ey;
    gz = g * ez;
    return true;
}

// The full pair set without culling (the scene's broad phase off): lane = sphere i, every partner
// j >= sph_lo[i] and every box in turn.  Accumulates the cost; when `grad`, also the forces in W.F
// (zeroed here).
__device__ inline float clr_pairs_all(const ArmView& V, const ClrLds& W, float m, bool grad) {
    const int lane = threadIdx.x;
    const int S = V.S;
    if (grad) {
        for (int t = lane; t < 3 * S; t += 64) W.F[t] = 0.f;
        __syncthreads();
    }
    float cost = 0.f;
    for (int base = 0; base < S; base += 64) {
        const int i = base + lane;
        const bool act = i < S;
        const float4 ci = W.sph[act ? i : 0];
        float fx = 0.f, fy = 0.f, fz = 0.f;
        // spheres vs boxes
        for (int b = 0; b < V.nb; ++b) {
            const float* bx = V.box32 + 8 * b;
            const float ex = ci.x - fminf(fmaxf(ci.x, bx[0]), bx[4]);
            const float ey = ci.y - fminf(fmaxf(ci.y, bx[1]), bx[5]);
            const float ez = ci.z - fminf(fmaxf(ci.z, bx[2]), bx[6]);
            const float e2 = ex * ex + ey * ey + ez * ez;
            const float R = m + ci.w;
            // the square root only for pairs within reach: d < R implies e2 < R^2 (1 + 2^-21),
            // so the bound below passes every pair with pen > 0, whose terms are unchanged
            if (!(act && e2 <= R * R * 1.0001f)) continue;
            const float d = sqrtf(e2);
            const float pen = R - d;
            if (pen > 0.f) {
                cost += pen * pen;
                if (grad && d > 0.f) {
                    const float g = -2.f * pen / d;
                    fx += g * ex;
                    fy += g * ey;
                    fz += g * ez;
                }
            }
        }
        // sphere pairs
        const int lo = act ? V.sph_lo[i] : 0x7fffffff;
        for (int j = V.sph_lo[base]; j < S; ++j) {
            const float4 cj = W.sph[j];
            const float ex = ci.x - cj.x, ey = ci.y - cj.y, ez = ci.z - cj.z;
            const float e2 = ex * ex + ey * ey + ez * ez;
            const float R = m + ci.w + cj.w;
            if (!(j >= lo && e2 <= R * R * 1.0001f)) continue;  // as above
            const float d = sqrtf(e2);
            const float pen = R - d;
            if (pen > 0.f) {
                cost += pen * pen;
                if (grad && d > 0.f) {
                    const float g = -2.f * pen / d;
                    fx += g * ex;
                    fy += g * ey;
                    fz += g * ez;
                    atomicAdd(&W.F[3 * j], -g * ex);
                    atomicAdd(&W.F[3 * j + 1], -g * ey);
                    atomicAdd(&W.F[3 * j + 2], -g * ez);
                }
            }
        }
        if (grad && act) {
            atomicAdd(&W.F[3 * i], fx);
            atomicAdd(&W.F[3 * i + 1], fy);
            atomicAdd(&W.F[3 * i + 2], fz);
        }
    }
    // wave sum
    W.red[lane] = cost;
    __syncthreads();
    for (int off = 32; off > 0; off >>= 1) {
        if (lane < off) W.red[lane] += W.red[lane + off];
        __syncthreads();
    }
    const float total = W.red[0];
    __syncthreads();
    return total;
}

// Accumulates the cost of the pair set; when `grad`, also the forces in W.F (zeroed here).
// Terms are nonzero only for pairs within reach, so both parts are culled with the links'
// bounding spheres first (the triangle inequality: a sphere of link A lies within R_A - r of A's
// centre, so no pair of two links farther apart than m + R_A + R_B, and no sphere of a link
// farther than m + R_A from a box, penetrates); the surviving (link, box) and (link, link)
// candidates are expanded to their spheres, 8 x 8 lanes at a time.  With the scene's broad phase
// off (V.broadphase = 0) the full pair set is walked instead (clr_pairs_all).
__device__ inline float clr_pairs(const ArmView& V, const ClrLds& W, float m, bool grad) {
    if (!V.broadphase) return clr_pairs_all(V, W, m, grad);
    const int lane = threadIdx.x;
    const int S = V.S, L = V.L;
    if (grad)
        for (int t = lane; t < 3 * S; t += 64) W.F[t] = 0.f;
    for (int l = lane; l < L; l += 64) {
        const float* X = W.X + 12 * l;
        const double* b = V.link_bc64 + 3 * l;
        const float b0 = (float)b[0], b1 = (float)b[1], b2 = (float)b[2];
        W.lk[l] = make_float4(X[0] * b0 + X[1] * b1 + X[2] * b2 + X[9], X[3] * b0 + X[4] * b1 + X[5] * b2 + X[10],
                              X[6] * b0 + X[7] * b1 + X[8] * b2 + X[11], V.link_R32[l]);
    }
    __syncthreads();
    float cost = 0.f;
    const unsigned long long below = (1ull << lane) - 1ull;
    // the spheres of (link, box) candidate `cd` (lane k of its group of 8 takes the link's sphere k)
    auto sphere_box = [&](int cd, int k) {
        const int l = cd & 255, b = cd >> 8;
        const int s = V.link_first[l] + k;
        if (s >= V.link_first[l + 1]) return;
        const float4 ci = W.sph[s];
        const float* bx = V.box32 + 8 * b;
        float gx, gy, gz;
        if (clr_term(ci.x - fminf(fmaxf(ci.x, bx[0]), bx[4]), ci.y - fminf(fmaxf(ci.y, bx[1]), bx[5]),
                     ci.z - fminf(fmaxf(ci.z, bx[2]), bx[6]), m + ci.w, grad, cost, gx, gy, gz)) {
            atomicAdd(&W.F[3 * s], gx);
            atomicAdd(&W.F[3 * s + 1], gy);
            atomicAdd(&W.F[3 * s + 2], gz);
        }
    };
    // squared distance from the link sphere's centre to a box [lo, hi]
    auto box_e2 = [](float4 c, const float* lo, const float* hi) {
        const float ex = c.x - fminf(fmaxf(c.x, lo[0]), hi[0]);
        const float ey = c.y - fminf(fmaxf(c.y, lo[1]), hi[1]);
        const float ez = c.z - fminf(fmaxf(c.z, lo[2]), hi[2]);
        return ex * ex + ey * ey + ez * ez;
    };
    // 1. spheres vs boxes.  Scenes with box clusters (at most 8 boxes each): (link, cluster) pairs
    //    lane-parallel, the boxes of the candidate clusters 8 lanes per pair, then the spheres of
    //    the candidate (link, box) pairs 8 lanes per pair; otherwise (link, box) pairs directly.
    if (V.nclu > 0) {
        const int nlc = L * V.nclu;
        for (int t0 = 0; t0 < nlc; t0 += 64) {
            const int t = t0 + lane;
            bool cand = false;
            int code = 0;
            if (t < nlc) {
                const int l = t / V.nclu, c = t - l * V.nclu;
                const float4 lk = W.lk[l];
                const float R = m + lk.w + CLR_CULL_EPS;
                cand = V.link_first[l + 1] > V.link_first[l] &&
                       box_e2(lk, V.clu32 + 8 * c, V.clu32 + 8 * c + 4) <= R * R;
                code = l | (c << 8);
            }
            const unsigned long long mk = __ballot(cand);
            if (mk == 0ull) continue;  // wave-uniform
            if (cand) W.cand[__popcll(mk & below)] = code;
            __syncthreads();
            const int nA = __popcll(mk);
            for (int a0 = 0; a0 < nA; a0 += 8) {  // 8 (link, cluster) candidates x 8 boxes
                const int g = lane >> 3, j = lane & 7;
                const bool ga = a0 + g < nA;
                const int c2 = W.cand[ga ? a0 + g : a0];
                const int l = c2 & 255, cl = c2 >> 8;
                const int first = __float_as_int(V.clu32[8 * cl + 3]), count = __float_as_int(V.clu32[8 * cl + 7]);
                const bool act = ga && j < count;
                const int b = first + (act ? j : 0);
                const float4 lk = W.lk[l];
                const float R = m + lk.w + CLR_CULL_EPS;
                const bool bc = act && box_e2(lk, V.box32 + 8 * b, V.box32 + 8 * b + 4) <= R * R;
                const unsigned long long bm = __ballot(bc);
                if (bm == 0ull) continue;  // wave-uniform
                if (bc) W.cand2[__popcll(bm & below)] = l | (b << 8);
                __syncthreads();
                const int nB = __popcll(bm);
                for (int c0 = 0; c0 < nB; c0 += 8)  // 8 (link, box) candidates x 8 spheres
                    if (c0 + (lane >> 3) < nB) sphere_box(W.cand2[c0 + (lane >> 3)], lane & 7);
                __syncthreads();  // W.cand2 is rewritten by the next group
            }
            __syncthreads();  // W.cand is rewritten by the next round
        }
    } else {
        const int nlb = L * V.nb;
        for (int t0 = 0; t0 < nlb; t0 += 64) {
            const int t = t0 + lane;
            bool cand = false;
            int code = 0;
            if (t < nlb) {
                const int l = t / V.nb, b = t - l * V.nb;
                const float4 lk = W.lk[l];
                const float R = m + lk.w + CLR_CULL_EPS;
                cand = V.link_first[l + 1] > V.link_first[l] && box_e2(lk, V.box32 + 8 * b, V.box32 + 8 * b + 4) <= R * R;
                code = l | (b << 8);
            }
            const unsigned long long mk = __ballot(cand);
            if (mk == 0ull) continue;  // wave-uniform
            if (cand) W.cand[__popcll(mk & below)] = code;
            __syncthreads();
            const int n = __popcll(mk);
            for (int c0 = 0; c0 < n; c0 += 8)  // 8 candidates x 8 spheres per round
                if (c0 + (lane >> 3) < n) sphere_box(W.cand[c0 + (lane >> 3)], lane & 7);
            __syncthreads();  // W.cand is rewritten by the next round
        }
    }
    // 2. sphere pairs of the link pairs in the pair set
    for (int p0 = 0; p0 < V.NP; p0 += 64) {
        const int pi = p0 + lane;
        bool cand = false;
        int y = 0;
        if (pi < V.NP) {
            const int2 pr = V.lpairs2[pi];
            const float4 A = W.lk[(pr.x >> 8) & 255], B = W.lk[pr.x & 255];
            const float dx = A.x - B.x, dy = A.y - B.y, dz = A.z - B.z;
            const float R = m + A.w + B.w + CLR_CULL_EPS;
            cand = dx * dx + dy * dy + dz * dz <= R * R;
            y = pr.y;
        }
        const unsigned long long mk = __ballot(cand);
        if (mk == 0ull) continue;  // wave-uniform
        if (cand) W.cand[__popcll(mk & below)] = y;
        __syncthreads();
        const int n = __popcll(mk);
        for (int c = 0; c < n; ++c) {  // one link pair per round: lane (i, j) of its 8 x 8 spheres
            const int yc = W.cand[c];
            const int fl = yc & 511, nl = (yc >> 9) & 15, fm = (yc >> 13) & 511, nm = (yc >> 22) & 15;
            if ((lane >> 3) >= nl || (lane & 7) >= nm) continue;
            const int i = fl + (lane >> 3), j = fm + (lane & 7);
            const float4 ci = W.sph[i], cj = W.sph[j];
            float gx, gy, gz;
            if (clr_term(ci.x - cj.x, ci.y - cj.y, ci.z - cj.z, m + ci.w + cj.w, grad, cost, gx, gy, gz)) {
                atomicAdd(&W.F[3 * i], gx);
                atomicAdd(&W.F[3 * i + 1], gy);
                atomicAdd(&W.F[3 * i + 2], gz);
                atomicAdd(&W.F[3 * j], -gx);
                atomicAdd(&W.F[3 * j + 1], -gy);
                atomicAdd(&W.F[3 * j + 2], -gz);
            }
        }
        __syncthreads();
    }
    // wave sum
    W.red[lane] = cost;
    __syncthreads();
    for (int off = 32; off > 0; off >>= 1) {
        if (lane < off) W.red[lane] += W.red[lane + off];
        __syncthreads();
    }
    const float total = W.red[0];
    __syncthreads();
    return total;
}

__global__ void __launch_bounds__(64) k_clearance(ArmView V, const float* q, long long n, float m,
                                                  const float* grad_cost, float* cost, float* grad_q) {
    extern __shared__ float lds_clr[];
    ClrLds W = clr_lds(lds_clr, V);
    const bool grad = grad_q != nullptr;
    for (long long r = blockIdx.x; r < n; r += gridDim.x) {
        const float* qr = q + r * V.D;
        clr_fk(V, W, qr);
        const float c = clr_pairs(V, W, m, grad);
        if (!grad) {
            if (threadIdx.x == 0) cost[r] = c;
        } else {
            __syncthreads();
            // lane (arm a, joint k): g = z_{k-1} . sum_{s on links >= k} (c_s - o_{k-1}) x f_s
            //   = z_{k-1} . (T_k - o_{k-1} x F_k), T_k and F_k the sums of c_s x f_s and of f_s
            //   over the arm's links >= k: per-link sums first (lane per link), then suffix sums
            const int lane = threadIdx.x;
            for (int l = lane; l < V.L; l += 64) {
                float t0 = 0.f, t1 = 0.f, t2 = 0.f, f0 = 0.f, f1 = 0.f, f2 = 0.f;
                for (int s = V.link_first[l]; s < V.link_first[l + 1]; ++s) {
                    const float4 c = W.sph[s];
                    const float a0 = W.F[3 * s], a1 = W.F[3 * s + 1], a2 = W.F[3 * s + 2];
                    t0 += c.y * a2 - c.z * a1;
                    t1 += c.z * a0 - c.x * a2;
                    t2 += c.x * a1 - c.y * a0;
                    f0 += a0;
                    f1 += a1;
                    f2 += a2;
                }
                float* tf = W.tf + 6 * l;
                tf[0] = t0; tf[1] = t1; tf[2] = t2;
                tf[3] = f0; tf[4] = f1; tf[5] = f2;
            }
            __syncthreads();
            if (lane < V.D) {
                const int a = lane / V.dof, k = lane % V.dof;
                float z0, z1, z2, o0, o1, o2;
                if (k == 0) {
                    const double* P = V.base + 12 * a;
                    z0 = (float)P[2]; z1 = (float)P[5]; z2 = (float)P[8];
                    o0 = (float)P[9]; o1 = (float)P[10]; o2 = (float)P[11];
                } else {
                    const float* P = W.X + 12 * (a * V.dof + k - 1);
                    z0 = P[2]; z1 = P[5]; z2 = P[8];
                    o0 = P[9]; o1 = P[10]; o2 = P[11];
                }
                float T0 = 0.f, T1 = 0.f, T2 = 0.f, F0 = 0.f, F1 = 0.f, F2 = 0.f;
                for (int l = (a + 1) * V.dof - 1; l >= a * V.dof + k; --l) {
                    const float* tf = W.tf + 6 * l;
                    T0 += tf[0]; T1 += tf[1]; T2 += tf[2];
                    F0 += tf[3]; F1 += tf[4]; F2 += tf[5];
                }
                const bool pr = (V.prism >> lane) & 1ull;  // prismatic: d c / d q = z, g = z . F
                const float g = pr ? z0 * F0 + z1 * F1 + z2 * F2
                                   : z0 * (T0 - (o1 * F2 - o2 * F1)) + z1 * (T1 - (o2 * F0 - o0 * F2)) +
                                         z2 * (T2 - (o0 * F1 - o1 * F0));
                grad_q[r * V.D + lane] = grad_cost[r] * g;
            }
        }
        __syncthreads();
    }
}

}  // namespace

size_t clearance_lds_bytes(const ArmView& V) {
    return sizeof(float) * (2 * V.D + 12 * V.L) + 16 + sizeof(float4) * V.S + sizeof(float) * 3 * V.S +
           sizeof(float) * 64 + 16 + sizeof(float4) * V.L + sizeof(int) * 64 + sizeof(float) * 6 * V.L +
           sizeof(int) * 64;
}

void launch_clearance(const SceneRec& sc, const float* q, long long n, float margin,
                      const float* grad_cost, float* cost, float* grad_q, hipStream_t st) {
    long long g = n < 65536 ? n : 65536;
    k_clearance<<<(int)g, 64, clearance_lds_bytes(sc.arm), st>>>(sc.arm, q, n, margin, grad_cost, cost, grad_q);
}
