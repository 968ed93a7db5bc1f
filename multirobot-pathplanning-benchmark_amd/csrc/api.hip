// api.hip — C ABI of libmrcc.so (declared in include/mrcc.h).
//
// Host side of the engine: scene construction (exact binary64 thresholds, binary32 copies and
// margins), edge plans (workspace + hipGraph), the per-call entry points and the error contract
// (status codes, thread-local message, injectable failures).
#include <math.h>

#include <cmath>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mrcc.h"
#include "engine.hpp"

namespace {

thread_local std::string g_err;
thread_local int g_inject_code = 0;
thread_local int g_inject_count = 0;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int injected() {
    if (g_inject_count > 0) {
        --g_inject_count;
        return fail(g_inject_code ? g_inject_code : MRCC_ERR_INJECTED, "injected failure");
    }
    return MRCC_OK;
}

#define HIP_OK(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(MRCC_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));      \
    } while (0)

#define CHECK_INJECT()                 \
    do {                               \
        int r_ = injected();           \
        if (r_ != MRCC_OK) return r_;  \
    } while (0)

int device_cus() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        return 256;
    return cus;
}

// smallest binary64 x >= 0 with sqrt(x) >= R (so sqrt(d2) < R <=> d2 < T)
double sqrt_threshold(double R) {
    if (!(R > 0.0)) return 0.0;  // sqrt(d2) < R never holds for d2 >= 0
    double x = R * R;
    while (x > 0.0 && sqrt(x) >= R) x = nextafter(x, 0.0);
    while (sqrt(x) < R) x = nextafter(x, INFINITY);
    return x;
}

double (*volatile pow_fn)(double, double) = pow;  // a libm call, as Python's r**2

// packs host arrays into one device allocation
struct Packer {
    std::vector<unsigned char> buf;
    template <class T>
    size_t add(const T* p, size_t n) {
        size_t off = (buf.size() + 255) & ~size_t(255);
        buf.resize(off + sizeof(T) * n + 1);
        if (n) memcpy(buf.data() + off, p, sizeof(T) * n);
        return off;
    }
    template <class T>
    size_t add(const std::vector<T>& v) { return add(v.data(), v.size()); }
};

template <class T>
const T* at(void* base, size_t off) { return (const T*)((unsigned char*)base + off); }

int upload(SceneRec* sc, Packer& P) {
    HIP_OK(hipMalloc(&sc->dev_mem, P.buf.size()));
    HIP_OK(hipMemcpy(sc->dev_mem, P.buf.data(), P.buf.size(), hipMemcpyHostToDevice));
    sc->dev_bytes = P.buf.size();
    return MRCC_OK;
}

// host description of an arm scene, kept so modes can be replaced
struct ArmHost {
    int na, dof;
    int is_chain;                  // 1: chain joints (C per joint), 0: DH joints
    std::vector<double> base, dh;  // dh: [L*3] d a alpha (DH joints)
    std::vector<double> chain;     // [L*12] C = [R | t] rows of each chain joint
    std::vector<double> travel;    // [L] |q| bound of each prismatic chain joint (0: revolute)
    unsigned long long prism = 0;  // bit l: chain joint l is prismatic
    std::vector<int> counts;
    std::vector<double> sph;       // [S*4]
    std::vector<double> boxes;     // [nb*6] centre half
    int nm;
    int nslot;                     // mode-box slots per mode: na held + resting
    std::vector<int> m_active;
    std::vector<unsigned char> m_grasp;
    std::vector<double> m_rel, m_half;
};

int build_arm_device(SceneRec* sc, const ArmHost& H) {
    const int na = H.na, dof = H.dof, L = na * dof;
    std::vector<int> link_first(L + 1, 0);
    for (int l = 0; l < L; ++l) link_first[l + 1] = link_first[l] + H.counts[l];
    const int S = link_first[L];
    if (S > MRCC_MAX_SPHERES) return fail(MRCC_ERR_INVALID, "at most %d spheres", MRCC_MAX_SPHERES);
    std::vector<int> sph_link(S), sph_lo(S);
    for (int l = 0; l < L; ++l)
        for (int t = link_first[l]; t < link_first[l + 1]; ++t) {
            sph_link[t] = l;
            const int a = l / dof, k = l % dof;
            sph_lo[t] = (k + 2 < dof) ? link_first[a * dof + k + 2] : link_first[(a + 1) * dof];
        }
    // joint table of the device: DH joints d a cos(alpha) sin(alpha); chain joints C as given
    std::vector<double> dh4(H.is_chain ? 12 * (size_t)L : 4 * (size_t)L);
    double B = 0.0;
    for (int a = 0; a < na; ++a) {
        double reach = sqrt(H.base[12 * a + 9] * H.base[12 * a + 9] + H.base[12 * a + 10] * H.base[12 * a + 10] +
                            H.base[12 * a + 11] * H.base[12 * a + 11]);
        for (int k = 0; k < dof; ++k) {
            const int l = a * dof + k;
            if (H.is_chain) {
                const double* C = H.chain.data() + 12 * l;
                for (int i = 0; i < 12; ++i) dh4[12 * l + i] = C[i];
                reach += sqrt(C[3] * C[3] + C[7] * C[7] + C[11] * C[11]);
                if ((H.prism >> l) & 1ull) reach += H.travel[l];  // translation |q| <= travel
                continue;
            }
            mrcc::SinCos t = mrcc::sincos_f64(H.dh[3 * l + 2]);
            dh4[4 * l] = H.dh[3 * l];
            dh4[4 * l + 1] = H.dh[3 * l + 1];
            dh4[4 * l + 2] = t.c;
            dh4[4 * l + 3] = t.s;
            reach += fabs(H.dh[3 * l]) + fabs(H.dh[3 * l + 1]);
        }
        B = std::max(B, reach);
    }
    double rmax = 0.0, lmax = 0.0;
    std::vector<float> sph_r32(S);
    for (int t = 0; t < S; ++t) {
        sph_r32[t] = (float)H.sph[4 * t + 3];
        rmax = std::max(rmax, H.sph[4 * t + 3]);
        lmax = std::max(lmax, fabs(H.sph[4 * t]) + fabs(H.sph[4 * t + 1]) + fabs(H.sph[4 * t + 2]));
    }
    B += lmax + rmax;
    const int nb = (int)(H.boxes.size() / 6);
    // Obstacle boxes are grouped into clusters of at most 8 for the two-level link-box test:
    // median splits of the box centres along their widest axis.  The device arrays list the
    // boxes cluster by cluster (the order of boxes does not change any result).
    std::vector<int> perm(nb);
    for (int b = 0; b < nb; ++b) perm[b] = b;
    std::vector<std::pair<int, int>> leaves;
    {
        std::vector<std::pair<int, int>> todo;
        if (nb > 0) todo.push_back({0, nb});
        while (!todo.empty()) {
            std::pair<int, int> r = todo.back();
            todo.pop_back();
            if (r.second - r.first <= 8) {
                leaves.push_back(r);
                continue;
            }
            double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (int i = r.first; i < r.second; ++i)
                for (int k = 0; k < 3; ++k) {
                    lo[k] = std::min(lo[k], H.boxes[6 * perm[i] + k]);
                    hi[k] = std::max(hi[k], H.boxes[6 * perm[i] + k]);
                }
            int ax = 0;
            for (int k = 1; k < 3; ++k)
                if (hi[k] - lo[k] > hi[ax] - lo[ax]) ax = k;
            const int mid = (r.first + r.second) / 2;
            std::stable_sort(perm.begin() + r.first, perm.begin() + r.second, [&](int a, int b) {
                return H.boxes[6 * a + ax] < H.boxes[6 * b + ax];
            });
            todo.push_back({mid, r.second});
            todo.push_back({r.first, mid});
        }
        std::sort(leaves.begin(), leaves.end());
    }
    const int nclu = (int)leaves.size() <= 16 ? (int)leaves.size() : 0;  // more: one-level test
    std::vector<double> bx(6 * (size_t)nb);
    for (int b = 0; b < nb; ++b)
        for (int k = 0; k < 6; ++k) bx[6 * b + k] = H.boxes[6 * perm[b] + k];
    std::vector<double> box64(6 * (size_t)nb);
    std::vector<float> box32(8 * (size_t)nb), boxch32(6 * (size_t)nb);
    double hmax = 0.0;
    for (int b = 0; b < nb; ++b) {
        for (int k = 0; k < 3; ++k) {
            const double c = bx[6 * b + k], h = bx[6 * b + 3 + k];
            box64[6 * b + k] = c - h;
            box64[6 * b + 3 + k] = c + h;
            box32[8 * b + k] = (float)(c - h);
            box32[8 * b + 4 + k] = (float)(c + h);
            B = std::max(B, fabs(c) + fabs(h));
            hmax = std::max(hmax, fabs(h));
        }
        for (int k = 0; k < 6; ++k) boxch32[6 * b + k] = (float)bx[6 * b + k];
    }
    // cluster bounds in binary32, rounded outward: lo xyz, first box, hi xyz, box count
    std::vector<float> clu32(8 * (size_t)(nclu > 0 ? nclu : 1), 0.0f);
    for (int c = 0; c < nclu; ++c) {
        const int f = leaves[c].first, e = leaves[c].second;
        for (int k = 0; k < 3; ++k) {
            double lo = INFINITY, hi = -INFINITY;
            for (int b = f; b < e; ++b) {
                lo = std::min(lo, box64[6 * b + k]);
                hi = std::max(hi, box64[6 * b + 3 + k]);
            }
            float l32 = (float)lo, h32 = (float)hi;
            if ((double)l32 > lo) l32 = nextafterf(l32, -INFINITY);
            if ((double)h32 < hi) h32 = nextafterf(h32, INFINITY);
            clu32[8 * c + k] = l32;
            clu32[8 * c + 4 + k] = h32;
        }
        int fi = f, ni = e - f;
        memcpy(&clu32[8 * c + 3], &fi, 4);
        memcpy(&clu32[8 * c + 7], &ni, 4);
    }
    // held boxes: reach of the end-effector + grasp offset + box extent
    double grasp_reach = 0.0;
    for (size_t i = 0; i < H.m_half.size(); ++i) hmax = std::max(hmax, fabs(H.m_half[i]));
    for (size_t i = 0; i < H.m_rel.size(); i += 12)  // held offsets, resting positions
        grasp_reach = std::max(grasp_reach, fabs(H.m_rel[i + 9]) + fabs(H.m_rel[i + 10]) + fabs(H.m_rel[i + 11]));
    B += grasp_reach + 2.0 * hmax;
    // broad phase: per link a bounding sphere (centre of the local sphere centres' box, radius
    // covering every sphere, inflated by 1e-9 relative + 1e-9 absolute against rounding), and the
    // list of link pairs the narrow phase tests: same arm with links >= 2 apart, or different arms.
    std::vector<double> link_bc(3 * (size_t)L, 0.0);
    std::vector<float> link_R(L, 0.0f);
    double Rl_max = 0.0;
    for (int l = 0; l < L; ++l) {
        const int f = link_first[l], e = link_first[l + 1];
        if (f == e) continue;
        double lo[3], hi[3];
        for (int k = 0; k < 3; ++k) lo[k] = hi[k] = H.sph[4 * f + k];
        for (int t = f; t < e; ++t)
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], H.sph[4 * t + k]);
                hi[k] = std::max(hi[k], H.sph[4 * t + k]);
            }
        double R = 0.0;
        for (int k = 0; k < 3; ++k) link_bc[3 * l + k] = 0.5 * (lo[k] + hi[k]);
        for (int t = f; t < e; ++t) {
            double d = 0.0;
            for (int k = 0; k < 3; ++k) d += (H.sph[4 * t + k] - link_bc[3 * l + k]) * (H.sph[4 * t + k] - link_bc[3 * l + k]);
            R = std::max(R, sqrt(d) + H.sph[4 * t + 3]);
        }
        R = R * (1.0 + 1e-9) + 1e-9;
        link_R[l] = (float)R;
        if ((double)link_R[l] < R) link_R[l] = nextafterf(link_R[l], INFINITY);
        Rl_max = std::max(Rl_max, (double)link_R[l]);
    }
    // pair list: first the self pairs of every arm (arm by arm), then one segment per arm pair
    // (a, b), a < b, in the order b = 1..na-1, a = 0..b-1, so that the pairs of the first n arms
    // form a prefix of each part.  lpair_sets[n] = self pairs of arms < n (n = 0..na); entry
    // na + 1 + k describes segment k = b (b - 1) / 2 + a: first pair | pairs << 12 | a << 20 | b << 24.
    std::vector<int> lpairs2, lpair_sets(na + 1 + na * (na - 1) / 2, 0);
    auto add_pair = [&](int l, int m) {
        if (link_first[l] == link_first[l + 1] || link_first[m] == link_first[m + 1]) return false;
        lpairs2.push_back(m | (l << 8) | ((m / dof) << 16) | ((l / dof) << 20));
        lpairs2.push_back(link_first[l] | (H.counts[l] << 9) | (link_first[m] << 13) | (H.counts[m] << 22));
        return true;
    };
    int n_pairs = 0;
    // self pairs by link gap (2, 3, ...): links two apart touch most often, and the wave stops at
    // the first contact it finds
    for (int a = 0; a < na; ++a) {
        for (int gap = 2; gap < dof; ++gap)
            for (int l = a * dof; l + gap < (a + 1) * dof; ++l) n_pairs += add_pair(l, l + gap);
        lpair_sets[a + 1] = n_pairs;
    }
    for (int b = 1; b < na; ++b)
        for (int a = 0; a < b; ++a) {
            const int first = n_pairs;
            for (int l = a * dof; l < (a + 1) * dof; ++l)
                for (int m = b * dof; m < (b + 1) * dof; ++m) n_pairs += add_pair(l, m);
            lpair_sets[na + 1 + b * (b - 1) / 2 + a] = first | ((n_pairs - first) << 12) | (a << 20) | (b << 24);
        }
    // the static LDS region of the arm kernels, laid out as arm_static_layout: a workgroup stages it
    // with one contiguous copy
    const int NP = n_pairs, nm = H.nm;
    const ArmStaticLayout lay = arm_static_layout(L, na, nb, S, NP, nm, nclu, H.is_chain);
    std::vector<unsigned char> img(lay.bytes, 0);
    auto put = [&](int off, const void* src, size_t bytes) {
        if (bytes) memcpy(img.data() + off, src, bytes);
    };
    put(lay.o_dh, dh4.data(), 8 * dh4.size());
    put(lay.o_base, H.base.data(), 8 * 12 * (size_t)na);
    for (int t = 0; t < S; ++t) put(lay.o_sphL + 24 * t, &H.sph[4 * t], 24);
    put(lay.o_linkBC, link_bc.data(), 8 * link_bc.size());
    if (nclu > 0) put(lay.o_box, box32.data(), 4 * box32.size());
    put(lay.o_clu, clu32.data(), 32 * (size_t)nclu);
    put(lay.o_sphR, sph_r32.data(), 4 * (size_t)S);
    for (int t = 0; t < S; ++t) {
        const int v = sph_link[t] | ((sph_link[t] / dof) << 8);  // link | arm << 8
        put(lay.o_sphLink + 4 * t, &v, 4);
    }
    put(lay.o_linkR, link_R.data(), 4 * (size_t)L);
    for (int l = 0; l < L; ++l) {
        const int v = link_first[l] | ((link_first[l + 1] - link_first[l]) << 16);  // first | count << 16
        put(lay.o_linkFC + 4 * l, &v, 4);
    }
    if (nm <= MRCC_MODES_LDS_MAX) put(lay.o_mact, H.m_active.data(), 4 * (size_t)nm);
    put(lay.o_sets, lpair_sets.data(), 4 * lpair_sets.size());
    if (NP <= ARM_PAIRS_LDS_MAX) put(lay.o_pairs, lpairs2.data(), 4 * lpairs2.size());
    Packer P;
    size_t o_img = P.add(img);
    size_t o_lbc = P.add(link_bc), o_lR = P.add(link_R), o_lp = P.add(lpairs2), o_lpc = P.add(lpair_sets);
    size_t o_base = P.add(H.base), o_dh = P.add(dh4), o_lf = P.add(link_first), o_sl = P.add(sph_link),
           o_slo = P.add(sph_lo), o_s64 = P.add(H.sph), o_sr = P.add(sph_r32), o_b32 = P.add(box32),
           o_b64 = P.add(box64), o_bch64 = P.add(bx), o_bch32 = P.add(boxch32), o_clu = P.add(clu32),
           o_ma = P.add(H.m_active), o_mg = P.add(H.m_grasp), o_mr = P.add(H.m_rel),
           o_mh = P.add(H.m_half);
    std::vector<double> travel(H.travel);
    travel.resize(L, 0.0);
    const size_t o_tr = P.add(travel);
    const int bp = sc->dev_mem ? sc->arm.broadphase : 1;  // keep the setting across mode updates
    void* old = sc->dev_mem;
    int rc = upload(sc, P);
    if (rc != MRCC_OK) return rc;
    if (old) (void)hipFree(old);
    void* d = sc->dev_mem;
    ArmView& V = sc->arm;
    V.na = na; V.dof = dof; V.D = na * dof; V.L = L; V.S = S; V.nb = nb; V.nm = H.nm; V.nslot = H.nslot;
    V.chain = H.is_chain;
    V.prism = H.prism;
    V.broadphase = bp;
    // binary32 margins: squared distances near contact are off by <= ~2^-21 sqrt(3) R B; box SAT
    // axis values by <= ~2^-21 (B + 3 H)
    V.Msph = (float)(ldexp(1.0, -18) * (1.0 + B) * (1.0 + 2.0 * rmax));
    V.Mobb = (float)(ldexp(1.0, -17) * (1.0 + B) * (1.0 + 3.0 * hmax));
    V.base = at<double>(d, o_base);
    V.dh = at<double>(d, o_dh);
    V.link_first = at<int>(d, o_lf);
    V.sph_link = at<int>(d, o_sl);
    V.sph_lo = at<int>(d, o_slo);
    V.sph64 = at<double>(d, o_s64);
    V.sph_r32 = at<float>(d, o_sr);
    V.box32 = at<float>(d, o_b32);
    V.box64 = at<double>(d, o_b64);
    V.boxch64 = at<double>(d, o_bch64);
    V.boxch32 = at<float>(d, o_bch32);
    V.m_active = at<int>(d, o_ma);
    V.m_grasp = at<unsigned char>(d, o_mg);
    V.m_rel = at<double>(d, o_mr);
    V.m_half = at<double>(d, o_mh);
    V.travel = at<double>(d, o_tr);
    V.Mcull = (float)(ldexp(1.0, -18) * (1.0 + B) * (1.0 + 2.0 * Rl_max));
    for (int k = 0; k < 3; ++k) {  // union of the obstacle boxes, rounded outward
        double lo = INFINITY, hi = -INFINITY;
        for (int b = 0; b < nb; ++b) {
            lo = std::min(lo, box64[6 * b + k]);
            hi = std::max(hi, box64[6 * b + 3 + k]);
        }
        if (nb == 0) lo = hi = 0.0;
        float l32 = (float)lo, h32 = (float)hi;
        if ((double)l32 > lo) l32 = nextafterf(l32, -INFINITY);
        if ((double)h32 < hi) h32 = nextafterf(h32, INFINITY);
        V.bb_lo[k] = l32;
        V.bb_hi[k] = h32;
    }
    V.nclu = nclu;
    V.clu32 = at<float>(d, o_clu);
    V.link_bc64 = at<double>(d, o_lbc);
    V.link_R32 = at<float>(d, o_lR);
    V.NP = n_pairs;
    V.lpairs2 = at<int2>(d, o_lp);
    V.lpair_sets = at<int>(d, o_lpc);
    // arm boxes separated by more than Marm on some axis cannot hold a link pair that passes the
    // link test: 2 sqrt(Mcull) covers the test's margin, 2^-20 (1 + B + R) the rounding of the boxes
    {
        const double m = 2.0 * sqrt((double)V.Mcull) + ldexp(1.0, -20) * (1.0 + B + Rl_max);
        V.Marm = (float)m;
        if ((double)V.Marm < m) V.Marm = nextafterf(V.Marm, INFINITY);
    }
    sc->D = V.D;
    sc->n_robots = na;
    V.grasp = 0;
    for (unsigned char g : H.m_grasp) V.grasp |= (g != 0);
    {
        ArmK& K = sc->armk;
        memset(&K, 0, sizeof(K));
        K.na = na; K.dof = dof; K.D = V.D; K.L = L; K.S = S; K.nb = nb; K.nm = nm; K.NP = NP; K.nclu = nclu;
        K.nslot = H.nslot;
        K.chain = V.chain;
        K.prism = V.prism;
        K.broadphase = bp;
        K.grasp = V.grasp;
        K.img_bytes = lay.bytes;
        K.Msph = V.Msph; K.Mobb = V.Mobb; K.Mcull = V.Mcull; K.Marm = V.Marm;
        for (int k = 0; k < 3; ++k) {
            K.bb_lo[k] = V.bb_lo[k];
            K.bb_hi[k] = V.bb_hi[k];
        }
        K.g = (const unsigned char*)d;
        K.o_img = (int)o_img; K.o_box32 = (int)o_b32; K.o_box64 = (int)o_b64; K.o_boxch32 = (int)o_bch32;
        K.o_boxch64 = (int)o_bch64; K.o_sph64 = (int)o_s64; K.o_sphlo = (int)o_slo; K.o_pairs = (int)o_lp;
        K.o_mact = (int)o_ma; K.o_mgrasp = (int)o_mg; K.o_mrel = (int)o_mr; K.o_mhalf = (int)o_mh;
        K.o_travel = (int)o_tr;
    }
    // waves per workgroup: the most resident waves per CU (160 KiB of LDS) within 64 KiB per
    // workgroup; the scene is staged once per workgroup
    sc->arm_nw = 0;
    int best = 0;
    for (int nw = 4; nw >= 1; nw >>= 1) {
        const size_t b = arm_lds_bytes(V.D, L, na, nb, S, V.NP, V.nm, V.nclu, V.grasp, V.chain, nw, H.nslot);
        if (b > 65536) continue;
        const int resident = (int)(163840 / b) * nw;
        if (resident > best) {
            best = resident;
            sc->arm_nw = nw;
        }
    }
    if (sc->arm_nw == 0) return fail(MRCC_ERR_INVALID, "arm scene needs more than 64 KiB of LDS per wave");
    // the per-call kernel: one workgroup of as many waves as 64 KiB of LDS holds, at most one per
    // SIMD of the CU (4): a launch is latency-bound, and a second wave on a SIMD slows the first
    // (the states of one edge usually all collide, so the first round decides most edges)
    sc->arm_nw_edge = 1;
    for (int nw = 4; nw >= 1; --nw)
        if (arm_lds_bytes(V.D, L, na, nb, S, V.NP, V.nm, V.nclu, V.grasp, V.chain, nw, H.nslot) <= 65536) {
            sc->arm_nw_edge = nw;
            break;
        }
    sc->arm_lds_edge = arm_lds_bytes(V.D, L, na, nb, S, V.NP, V.nm, V.nclu, V.grasp, V.chain, sc->arm_nw_edge, H.nslot);
    sc->arm_lds = arm_lds_bytes(V.D, L, na, nb, S, V.NP, V.nm, V.nclu, V.grasp, V.chain, sc->arm_nw, H.nslot);
    sc->arm_lds1 = arm_lds_bytes(V.D, L, na, nb, S, V.NP, V.nm, V.nclu, V.grasp, V.chain, 1, H.nslot);
    // the persistent kernel: the most resident waves per CU, counting both budgets: its registers
    // allow 4 waves per SIMD (16 per CU), 3 with held boxes (12), and the LDS holds what fits in
    // 160 KiB, one staged scene per workgroup.  gfx950 gives a workgroup up to 160 KiB, so scenes
    // with held boxes, whose waves need the most LDS (an 8-arm scene: 14 KB per wave), take
    // workgroups of up to MRCC_PERSISTENT_GRASP_NW waves and share one scene among more of them.
    {
        const int reg_waves = V.grasp ? 12 : 16;
        const int nw_max = V.grasp ? MRCC_PERSISTENT_GRASP_NW : 4;
        int bestw = 0;
        for (int nw = 1; nw <= nw_max; ++nw) {
            const size_t b = arm_lds_bytes(V.D, L, na, nb, S, V.NP, V.nm, V.nclu, V.grasp, V.chain, nw, H.nslot);
            if (b > 163840) break;
            const int wgs = std::min((int)(163840 / b), reg_waves / nw);
            if (wgs * nw > bestw) {
                bestw = wgs * nw;
                sc->arm_nw_pers = nw;
                sc->arm_wg_pers = wgs;
            }
        }
        if (bestw == 0) return fail(MRCC_ERR_INVALID, "arm scene needs more than 160 KiB of LDS per wave");
        sc->arm_lds_pers = arm_lds_bytes(V.D, L, na, nb, S, V.NP, V.nm, V.nclu, V.grasp, V.chain, sc->arm_nw_pers, H.nslot);
        sc->arm_blocks_pers = sc->arm_wg_pers * device_cus();
    }
    sc->arm_na = na;
    sc->arm_dof = dof;
    sc->arm_S = S;
    sc->arm_nm = H.nm;
    sc->arm_has_grasp = V.grasp;
    return MRCC_OK;
}

struct SceneBox {
    SceneRec rec;
    ArmHost arm_host;
};

SceneRec* rec_of(const mrcc_scene* s) { return &((SceneBox*)s)->rec; }

EdgeBatchView view_of(const mrcc_edge_batch* b) {
    EdgeBatchView v;
    v.q1 = b->q1;
    v.q2 = b->q2;
    v.rs = b->row_stride;
    v.ds = b->dim_stride;
    v.idx1 = b->idx1;
    v.idx2 = b->idx2;
    v.n = b->n_edges;
    v.res = b->resolution;
    v.N = b->N;
    v.N_start = b->N_start;
    v.N_max = b->N_max;
    v.incl = b->include_endpoints;
    v.modes = b->mode_ids;
    v.mask = b->robot_mask;
    v.out = b->out;
    v.N0 = 0;
    v.Ns0 = 0;
    v.Nm0 = -1;
    v.md0 = 0;
    return v;
}

int check_batch(const SceneRec* sc, const mrcc_edge_batch* b) {
    if (!b) return fail(MRCC_ERR_INVALID, "null edge batch");
    if (b->n_edges < 0) return fail(MRCC_ERR_INVALID, "n_edges < 0");
    if (b->n_edges == 0) return MRCC_OK;  // nothing is read or written
    if (b->n_edges > 0x7fffffffLL) return fail(MRCC_ERR_INVALID, "at most 2^31 - 1 edges per batch");
    if (!b->q1 || !b->q2 || !b->out) return fail(MRCC_ERR_INVALID, "null edge batch buffer");
    if (!b->N && !(b->resolution > 0.0)) return fail(MRCC_ERR_INVALID, "resolution must be > 0");
    if (sc->D > MRCC_MAX_D) return fail(MRCC_ERR_INVALID, "D > %d", MRCC_MAX_D);
    return MRCC_OK;
}

}  // namespace

// ---- shared with the no-allocation entry point (noalloc.hip) ----------------------------------
namespace mrcc_internal {
int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}
int injected() { return ::injected(); }
int device_cus() { return ::device_cus(); }
const SceneRec* scene_rec(const mrcc_scene* s) { return rec_of(s); }
int check_batch(const SceneRec* sc, const mrcc_edge_batch* b) { return ::check_batch(sc, b); }
EdgeBatchView batch_view(const mrcc_edge_batch* b) { return view_of(b); }

int tune_defaults(const SceneRec* sc, const mrcc_tune* tune, mrcc_tune* t) {
    memset(t, 0, sizeof(*t));
    if (tune) *t = *tune;
    const bool arms = sc->kind == KIND_ARMS;
    if (t->slots_per_pass <= 0) t->slots_per_pass = arms ? 4 : 8;
    if (t->slots_per_ticket <= 0) t->slots_per_ticket = arms ? 16 : 4;
    if (t->waves_per_cu <= 0) t->waves_per_cu = 16;
    if (!tune) {
        t->sorted = 1;
        t->persistent = 1;
        t->use_graph = 1;
    }
    if ((t->slots_per_pass & (t->slots_per_pass - 1)) || t->slots_per_pass > 64)
        return ::fail(MRCC_ERR_INVALID, "slots_per_pass must be a power of two <= 64");
    if (arms && t->slots_per_ticket > 64)  // one lane resolves one slot of a claimed chunk
        return ::fail(MRCC_ERR_INVALID, "arm scenes: slots_per_ticket <= 64");
    return MRCC_OK;
}

int plan_grid(const SceneRec* sc, const mrcc_tune& t) {
    const int cus = ::device_cus();
    return sc->kind == KIND_ARMS ? t.waves_per_cu * cus : std::max(1, t.waves_per_cu * cus / 4);
}

size_t plan_layout(long long n_edges, int D, unsigned char* base, PlanWork* W) {
    const long long n = std::max<long long>(n_edges, 1);
    size_t top = 0;
    auto sz = [&](size_t bytes) {
        const size_t off = (top + 255) & ~size_t(255);
        top = off + bytes;
        return off;
    };
    const size_t o_rec = sz(16 * n), o_order = sz(4 * n), o_hist = sz(4 * (MRCC_MAX_PASSES + 1)),
                 o_cur = sz(4 * (MRCC_MAX_PASSES + 1)), o_ps = sz(8 * (MRCC_MAX_PASSES + 1)), o_meta = sz(8 * 8),
                 o_long = sz(4 * n), o_tick = sz(8 * 32 * 8), o_stats = sz(8 * 8), o_dir = sz(8 * (size_t)n * D);
    if (W) {
        W->erec = (int4*)(base + o_rec);
        W->order = (int*)(base + o_order);
        W->hist = (int*)(base + o_hist);
        W->cursor = (int*)(base + o_cur);
        W->pass_start = (long long*)(base + o_ps);
        W->meta = (long long*)(base + o_meta);
        W->long_list = (int*)(base + o_long);
        W->ticket = (unsigned long long*)(base + o_tick);
        W->stats = (unsigned long long*)(base + o_stats);
        W->dir = (double*)(base + o_dir);
    }
    return top;
}

void launch_plan_sequence(const SceneRec& sc, const EdgeBatchView& B, const PlanWork& W, int lgG, int spt,
                          int grid, hipStream_t st) {
    launch_plan_build(sc, B, W, lgG, st);
    launch_plan_run(sc, B, W, lgG, spt, grid, st);
    // edges with more passes than the queue holds: one workgroup per edge
    launch_direct(sc, B, W.long_list, 0, W.meta + 3, 1, 0, 4 * ::device_cus(), W.stats, st);
}
}  // namespace mrcc_internal

struct mrcc_edge_plan {
    SceneRec* sc;
    EdgeBatchView B;
    mrcc_tune tune;
    int lgG;
    int grid;
    void* work_mem;
    PlanWork W;
    hipGraph_t graph;
    hipGraphExec_t exec;
    hipStream_t cap_stream;
    hipEvent_t ev[4];  // time_kernels: start, queue built, main kernel done, end
    int timed_runs;
};

static void plan_launch_sequence(mrcc_edge_plan* p, hipStream_t st) {
    const bool ev = p->ev[0] != nullptr;
    if (ev) (void)hipEventRecord(p->ev[0], st);
    if (p->tune.persistent && !ev) {
        mrcc_internal::launch_plan_sequence(*p->sc, p->B, p->W, p->lgG, p->tune.slots_per_ticket, p->grid, st);
    } else if (p->tune.persistent) {
        launch_plan_build(*p->sc, p->B, p->W, p->lgG, st);
        (void)hipEventRecord(p->ev[1], st);
        launch_plan_run(*p->sc, p->B, p->W, p->lgG, p->tune.slots_per_ticket, p->grid, st);
        (void)hipEventRecord(p->ev[2], st);
        // edges with more passes than the queue holds: one workgroup per edge
        launch_direct(*p->sc, p->B, p->W.long_list, 0, p->W.meta + 3, 1, 0, 4 * device_cus(),
                      p->W.stats, st);
    } else {
        if (ev) (void)hipEventRecord(p->ev[1], st);
        launch_direct(*p->sc, p->B, nullptr, p->B.n, nullptr, 1, 1, 1 << 20, p->W.stats, st);
        if (ev) (void)hipEventRecord(p->ev[2], st);
    }
    if (ev) {
        (void)hipEventRecord(p->ev[3], st);
        ++p->timed_runs;
    }
}

extern "C" {

int mrcc_abi_version(void) { return MRCC_ABI_VERSION; }

const char* mrcc_last_error(void) { return g_err.c_str(); }

int mrcc_init(int device) {
    CHECK_INJECT();
    int n = 0;
    HIP_OK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(MRCC_ERR_INVALID, "device %d of %d", device, n);
    HIP_OK(hipSetDevice(device));
    return mrcc_noalloc_setup();  // per-device table of the no-allocation entry point
}

int mrcc_debug_inject_failure(int code, int count) {
    g_inject_code = code;
    g_inject_count = count;
    return MRCC_OK;
}

int mrcc_scene_create_disks(int n_robots, int robot_dim, const double* robot_radius,
                            int n_spheres, const double* sphere_center, const double* sphere_radius,
                            int n_rects, const double* rect_center, const double* rect_size,
                            mrcc_scene** out) {
    CHECK_INJECT();
    if (!out) return fail(MRCC_ERR_INVALID, "out is null");
    if (n_robots < 1 || n_robots > MRCC_DISK_MAX_ROBOTS || robot_dim < 1 || robot_dim > MRCC_DISK_MAX_RD)
        return fail(MRCC_ERR_INVALID, "disks: need 1..%d robots of dimension 1..%d", MRCC_DISK_MAX_ROBOTS,
                    MRCC_DISK_MAX_RD);
    if (n_robots * robot_dim > MRCC_MAX_D) return fail(MRCC_ERR_INVALID, "D > %d", MRCC_MAX_D);
    if (n_spheres < 0 || n_rects < 0 || n_spheres > (1 << 20) || n_rects > (1 << 20))
        return fail(MRCC_ERR_INVALID, "obstacle counts must be in 0..%d", 1 << 20);
    if (!robot_radius || (n_spheres > 0 && (!sphere_center || !sphere_radius)) ||
        (n_rects > 0 && (!rect_center || !rect_size)))
        return fail(MRCC_ERR_INVALID, "null argument");
    const int nr = n_robots, rd = robot_dim, ns = n_spheres, nq = n_rects;
    std::vector<float> r32(nr), rrT32(nr * nr), sc32(ns * rd + 1), srT32(nr * ns + 1),
        lo32(nq * rd + 1), hi32(nq * rd + 1), rp2_32(nr);
    std::vector<double> rrT64(nr * nr), sc64(ns * rd + 1), srT64(nr * ns + 1), lo64(nq * rd + 1),
        hi64(nq * rd + 1), rp2_64(nr);
    double B = 0.0;
    for (int i = 0; i < nr; ++i) {
        r32[i] = (float)robot_radius[i];
        rp2_64[i] = pow_fn(robot_radius[i], 2.0);
        rp2_32[i] = (float)rp2_64[i];
        for (int j = 0; j < nr; ++j) {
            rrT64[i * nr + j] = sqrt_threshold(robot_radius[i] + robot_radius[j]);
            rrT32[i * nr + j] = (float)rrT64[i * nr + j];
        }
        B = std::max(B, 2.0 * fabs(robot_radius[i]));
    }
    for (int o = 0; o < ns; ++o) {
        for (int k = 0; k < rd; ++k) {
            sc64[o * rd + k] = sphere_center[o * rd + k];
            sc32[o * rd + k] = (float)sphere_center[o * rd + k];
            B = std::max(B, fabs(sphere_center[o * rd + k]) + fabs(sphere_radius[o]));
        }
        for (int i = 0; i < nr; ++i) {
            srT64[i * ns + o] = sqrt_threshold(sphere_radius[o] + robot_radius[i]);
            srT32[i * ns + o] = (float)srT64[i * ns + o];
        }
    }
    for (int o = 0; o < nq; ++o)
        for (int k = 0; k < rd; ++k) {
            const double c = rect_center[o * rd + k], h = rect_size[o * rd + k] / 2;
            lo64[o * rd + k] = c - h;  // abstract_env.py:66
            hi64[o * rd + k] = c + h;  // abstract_env.py:67
            lo32[o * rd + k] = (float)lo64[o * rd + k];
            hi32[o * rd + k] = (float)hi64[o * rd + k];
            B = std::max(B, std::max(fabs(lo64[o * rd + k]), fabs(hi64[o * rd + k])));
        }
    SceneBox* box = new SceneBox();
    SceneRec* sc = &box->rec;
    memset(sc, 0, sizeof(SceneRec));
    sc->kind = KIND_DISKS;
    sc->D = nr * rd;
    sc->n_robots = nr;
    (void)hipGetDevice(&sc->device);
    Packer P;
    size_t o_r = P.add(r32), o_rrT32 = P.add(rrT32), o_rrT64 = P.add(rrT64), o_sc32 = P.add(sc32),
           o_sc64 = P.add(sc64), o_srT32 = P.add(srT32), o_srT64 = P.add(srT64), o_lo32 = P.add(lo32),
           o_hi32 = P.add(hi32), o_lo64 = P.add(lo64), o_hi64 = P.add(hi64), o_rp32 = P.add(rp2_32),
           o_rp64 = P.add(rp2_64);
    int rc = upload(sc, P);
    if (rc != MRCC_OK) {
        delete box;
        return rc;
    }
    void* d = sc->dev_mem;
    DiskView& V = sc->disk;
    V.nr = nr; V.rd = rd; V.D = nr * rd; V.ns = ns; V.nrect = nq;
    V.Bscene = (float)B;
    V.r32 = at<float>(d, o_r);
    V.rrT32 = at<float>(d, o_rrT32);
    V.rrT64 = at<double>(d, o_rrT64);
    V.sc32 = at<float>(d, o_sc32);
    V.sc64 = at<double>(d, o_sc64);
    V.srT32 = at<float>(d, o_srT32);
    V.srT64 = at<double>(d, o_srT64);
    V.lo32 = at<float>(d, o_lo32);
    V.hi32 = at<float>(d, o_hi32);
    V.lo64 = at<double>(d, o_lo64);
    V.hi64 = at<double>(d, o_hi64);
    V.rp2_32 = at<float>(d, o_rp32);
    V.rp2_64 = at<double>(d, o_rp64);
    *out = (mrcc_scene*)box;
    return MRCC_OK;
}

int mrcc_scene_create_planar_boxes(int n_robots, const double* robot_half_extents, int n_obstacles,
                                   const double* obstacle_pose, const double* obstacle_half_extents,
                                   mrcc_scene** out) {
    CHECK_INJECT();
    if (!out) return fail(MRCC_ERR_INVALID, "out is null");
    if (n_robots < 1 || n_robots > MRCC_PLANAR_MAX_ROBOTS)
        return fail(MRCC_ERR_INVALID, "planar: 1..%d robots", MRCC_PLANAR_MAX_ROBOTS);
    if (n_obstacles < 0 || n_obstacles > (1 << 20)) return fail(MRCC_ERR_INVALID, "obstacle count must be in 0..%d", 1 << 20);
    if (!robot_half_extents || (n_obstacles > 0 && (!obstacle_pose || !obstacle_half_extents)))
        return fail(MRCC_ERR_INVALID, "null argument");
    const int nr = n_robots, no = n_obstacles;
    std::vector<double> rh64(2 * nr), ob64(8 * no + 1);
    std::vector<float> rh32(2 * nr), ob32(8 * no + 1);
    double B = 0.0, H = 0.0;
    for (int i = 0; i < 2 * nr; ++i) {
        rh64[i] = robot_half_extents[i];
        rh32[i] = (float)rh64[i];
        H = std::max(H, fabs(rh64[i]));
    }
    for (int o = 0; o < no; ++o) {
        mrcc::SinCos t = mrcc::sincos_f64(obstacle_pose[3 * o + 2]);
        double* b = ob64.data() + 8 * o;
        b[0] = obstacle_pose[3 * o];
        b[1] = obstacle_pose[3 * o + 1];
        b[2] = t.c; b[3] = t.s;
        b[4] = -t.s; b[5] = t.c;
        b[6] = obstacle_half_extents[2 * o];
        b[7] = obstacle_half_extents[2 * o + 1];
        for (int k = 0; k < 8; ++k) ob32[8 * o + k] = (float)b[k];
        B = std::max(B, std::max(fabs(b[0]), fabs(b[1])));
        H = std::max(H, std::max(fabs(b[6]), fabs(b[7])));
    }
    SceneBox* box = new SceneBox();
    SceneRec* sc = &box->rec;
    memset(sc, 0, sizeof(SceneRec));
    sc->kind = KIND_PLANAR;
    sc->D = 3 * nr;
    sc->n_robots = nr;
    (void)hipGetDevice(&sc->device);
    Packer P;
    size_t o_rh32 = P.add(rh32), o_rh64 = P.add(rh64), o_ob32 = P.add(ob32), o_ob64 = P.add(ob64);
    int rc = upload(sc, P);
    if (rc != MRCC_OK) {
        delete box;
        return rc;
    }
    void* d = sc->dev_mem;
    PlanarView& V = sc->planar;
    V.nr = nr; V.no = no; V.D = 3 * nr;
    V.Bscene = (float)(B + 4.0 * H);
    V.rh32 = at<float>(d, o_rh32);
    V.rh64 = at<double>(d, o_rh64);
    V.ob32 = at<float>(d, o_ob32);
    V.ob64 = at<double>(d, o_ob64);
    *out = (mrcc_scene*)box;
    return MRCC_OK;
}

static int create_arm_scene(int n_arms, int dof, const double* base, const double* dh, const double* chain,
                            const int32_t* link_sphere_count, const double* sphere_local, int n_boxes,
                            const double* boxes, mrcc_scene** out) {
    if (!out) return fail(MRCC_ERR_INVALID, "out is null");
    if (n_arms < 1 || n_arms > MRCC_MAX_ARMS || dof < 1 || dof > MRCC_MAX_DOF)
        return fail(MRCC_ERR_INVALID, "arms: 1..%d arms of 1..%d joints", MRCC_MAX_ARMS, MRCC_MAX_DOF);
    if (n_boxes < 0 || n_boxes > 65535) return fail(MRCC_ERR_INVALID, "arms: 0..65535 boxes");
    if (!base || (!dh && !chain) || !link_sphere_count || (n_boxes > 0 && !boxes))
        return fail(MRCC_ERR_INVALID, "null argument");
    const int L = n_arms * dof;
    if (chain)
        for (int i = 0; i < 12 * L; ++i)
            if (!std::isfinite(chain[i])) return fail(MRCC_ERR_INVALID, "joint transform %d is not finite", i / 12);
    SceneBox* box = new SceneBox();
    SceneRec* sc = &box->rec;
    memset(sc, 0, sizeof(SceneRec));
    sc->kind = KIND_ARMS;
    (void)hipGetDevice(&sc->device);
    ArmHost& H = box->arm_host;
    H.na = n_arms;
    H.dof = dof;
    H.is_chain = chain ? 1 : 0;
    H.base.assign(base, base + 12 * n_arms);
    if (chain) H.chain.assign(chain, chain + 12 * L);
    else H.dh.assign(dh, dh + 3 * L);
    H.counts.assign(link_sphere_count, link_sphere_count + L);
    int S = 0;
    for (int l = 0; l < L; ++l) {
        if (H.counts[l] < 0 || H.counts[l] > 8) {
            delete box;
            return fail(MRCC_ERR_INVALID, "link %d: %d spheres (0..8 per link)", l, H.counts[l]);
        }
        S += H.counts[l];
    }
    H.sph.assign(sphere_local, sphere_local + 4 * S);
    H.boxes.assign(boxes, boxes + 6 * n_boxes);
    H.nm = 1;
    H.nslot = n_arms;
    H.m_active.assign(1, n_arms);
    H.m_grasp.assign(n_arms, 0);
    H.m_rel.assign(12 * n_arms, 0.0);
    H.m_half.assign(3 * n_arms, 0.0);
    int rc = build_arm_device(sc, H);
    if (rc != MRCC_OK) {
        if (sc->dev_mem) (void)hipFree(sc->dev_mem);
        delete box;
        return rc;
    }
    *out = (mrcc_scene*)box;
    return MRCC_OK;
}

int mrcc_scene_create_arms(int n_arms, int dof, const double* base, const double* dh,
                           const int32_t* link_sphere_count, const double* sphere_local,
                           int n_boxes, const double* boxes, mrcc_scene** out) {
    CHECK_INJECT();
    return create_arm_scene(n_arms, dof, base, dh, nullptr, link_sphere_count, sphere_local, n_boxes, boxes, out);
}

int mrcc_scene_create_arm_chains(int n_arms, int dof, const double* base, const double* joint_fixed,
                                 const int32_t* link_sphere_count, const double* sphere_local,
                                 int n_boxes, const double* boxes, mrcc_scene** out) {
    CHECK_INJECT();
    if (!joint_fixed) return fail(MRCC_ERR_INVALID, "null argument");
    return create_arm_scene(n_arms, dof, base, nullptr, joint_fixed, link_sphere_count, sphere_local, n_boxes,
                            boxes, out);
}

int mrcc_scene_set_modes(mrcc_scene* scene, int n_modes, const int32_t* n_active, int n_rest,
                         const uint8_t* has_grasp, const double* grasp_rel, const double* grasp_half) {
    CHECK_INJECT();
    if (!scene) return fail(MRCC_ERR_INVALID, "scene is null");
    SceneBox* box = (SceneBox*)scene;
    if (box->rec.kind != KIND_ARMS) return fail(MRCC_ERR_UNSUPPORTED, "modes need an arm scene");
    if (n_modes < 1) return fail(MRCC_ERR_INVALID, "n_modes < 1");
    ArmHost& H = box->arm_host;
    if (n_rest < 0 || H.na + n_rest > MRCC_MAX_SLOTS)
        return fail(MRCC_ERR_INVALID, "n_rest %d: at most %d boxes per mode (held and resting)", n_rest, MRCC_MAX_SLOTS);
    if (!n_active || !has_grasp || !grasp_rel || !grasp_half) return fail(MRCC_ERR_INVALID, "null argument");
    const int ns = H.na + n_rest;
    const size_t nms = (size_t)n_modes * ns;
    for (int m = 0; m < n_modes; ++m)
        if (n_active[m] < 1 || n_active[m] > H.na)
            return fail(MRCC_ERR_INVALID, "mode %d: n_active %d not in 1..%d", m, n_active[m], H.na);
    for (size_t i = 0; i < 12 * nms; ++i)
        if (!std::isfinite(grasp_rel[i])) return fail(MRCC_ERR_INVALID, "box transform %zu is not finite", i / 12);
    for (size_t i = 0; i < 3 * nms; ++i)
        if (!(grasp_half[i] >= 0.0) || !std::isfinite(grasp_half[i]))
            return fail(MRCC_ERR_INVALID, "box half extent %zu is negative or not finite", i);
    H.nm = n_modes;
    H.nslot = ns;
    H.m_active.assign(n_active, n_active + n_modes);
    H.m_grasp.assign(has_grasp, has_grasp + nms);
    H.m_rel.assign(grasp_rel, grasp_rel + 12 * nms);
    H.m_half.assign(grasp_half, grasp_half + 3 * nms);
    HIP_OK(hipDeviceSynchronize());  // kernels may still read the old arrays
    return build_arm_device(&box->rec, H);
}

int mrcc_scene_set_joint_types(mrcc_scene* scene, const uint8_t* joint_type, const double* travel) {
    CHECK_INJECT();
    if (!scene || !joint_type || !travel) return fail(MRCC_ERR_INVALID, "null argument");
    SceneBox* box = (SceneBox*)scene;
    if (box->rec.kind != KIND_ARMS) return fail(MRCC_ERR_UNSUPPORTED, "joint types apply to arm scenes");
    ArmHost& H = box->arm_host;
    const int L = H.na * H.dof;
    unsigned long long prism = 0;
    for (int l = 0; l < L; ++l) {
        if (joint_type[l] > 1) return fail(MRCC_ERR_INVALID, "joint %d: type %d is neither 0 (revolute) nor 1 (prismatic)", l, joint_type[l]);
        if (joint_type[l] == 0) continue;
        if (!H.is_chain) return fail(MRCC_ERR_UNSUPPORTED, "prismatic joints need a chain scene (mrcc_scene_create_arm_chains)");
        if (!(travel[l] >= 0.0) || !std::isfinite(travel[l]))
            return fail(MRCC_ERR_INVALID, "joint %d: travel must be finite and >= 0", l);
        prism |= 1ull << l;
    }
    H.prism = prism;
    H.travel.assign(L, 0.0);
    for (int l = 0; l < L; ++l)
        if ((prism >> l) & 1ull) H.travel[l] = travel[l];
    HIP_OK(hipDeviceSynchronize());  // kernels may still read the old arrays
    return build_arm_device(&box->rec, H);
}

int mrcc_scene_set_broadphase(mrcc_scene* scene, int enabled) {
    if (!scene) return fail(MRCC_ERR_INVALID, "scene is null");
    SceneRec* sc = rec_of(scene);
    if (sc->kind != KIND_ARMS) return fail(MRCC_ERR_UNSUPPORTED, "broad phase applies to arm scenes");
    sc->arm.broadphase = enabled ? 1 : 0;
    sc->armk.broadphase = sc->arm.broadphase;
    return MRCC_OK;
}

int mrcc_scene_destroy(mrcc_scene* scene) {
    if (!scene) return MRCC_OK;
    SceneBox* box = (SceneBox*)scene;
    SceneRec* sc = &box->rec;
    (void)hipDeviceSynchronize();
    if (sc->dev_mem) (void)hipFree(sc->dev_mem);
    if (sc->stage_host) (void)hipHostFree(sc->stage_host);  // stage_dev is its device address
    if (sc->batch_dev) (void)hipFree(sc->batch_dev);
    if (sc->batch_host) (void)hipHostFree(sc->batch_host);
    if (sc->list_host) (void)hipHostFree(sc->list_host);  // list_dev is its device address
    delete box;
    return MRCC_OK;
}

int mrcc_scene_info(const mrcc_scene* scene, int32_t* out4) {
    if (!scene || !out4) return fail(MRCC_ERR_INVALID, "null argument");
    const SceneRec* sc = rec_of(scene);
    out4[0] = sc->kind;
    out4[1] = sc->D;
    out4[2] = sc->n_robots;
    out4[3] = sc->kind == KIND_ARMS ? sc->arm.nm : 1;
    return MRCC_OK;
}

int mrcc_check_configs(const mrcc_scene* scene, const double* q, int64_t n, int64_t row_stride,
                       int64_t dim_stride, const int32_t* mode_ids, uint32_t robot_mask,
                       uint8_t* out, void* stream) {
    CHECK_INJECT();
    if (!scene || (!q && n) || (!out && n)) return fail(MRCC_ERR_INVALID, "null argument");
    if (n < 0) return fail(MRCC_ERR_INVALID, "n < 0");
    if (n == 0) return MRCC_OK;
    const SceneRec* sc = rec_of(scene);
    launch_configs(*sc, q, n, row_stride, dim_stride, mode_ids, robot_mask, out, 64 * device_cus(),
                   (hipStream_t)stream);
    HIP_OK(hipGetLastError());
    return MRCC_OK;
}

int mrcc_edge_plan_create(const mrcc_scene* scene, const mrcc_edge_batch* batch,
                          const mrcc_tune* tune, void* stream, mrcc_edge_plan** out) {
    CHECK_INJECT();
    if (!scene || !out) return fail(MRCC_ERR_INVALID, "null argument");
    SceneRec* sc = rec_of(scene);
    int rc = check_batch(sc, batch);
    if (rc != MRCC_OK) return rc;
    mrcc_tune t;
    rc = mrcc_internal::tune_defaults(sc, tune, &t);
    if (rc != MRCC_OK) return rc;
    mrcc_edge_plan* p = new mrcc_edge_plan();
    memset(p, 0, sizeof(*p));
    p->sc = sc;
    p->B = view_of(batch);
    p->tune = t;
    p->lgG = __builtin_ctz((unsigned)t.slots_per_pass);
    p->grid = mrcc_internal::plan_grid(sc, t);
    // workspace
    const size_t work_bytes = mrcc_internal::plan_layout(batch->n_edges, sc->D, nullptr, nullptr);
    if (hipMalloc(&p->work_mem, work_bytes) != hipSuccess) {
        delete p;
        return fail(MRCC_ERR_NOMEM, "workspace of %zu bytes", work_bytes);
    }
    unsigned char* w = (unsigned char*)p->work_mem;
    mrcc_internal::plan_layout(batch->n_edges, sc->D, w, &p->W);
    unsigned char* stats_mem = (unsigned char*)p->W.stats;
    if (!t.collect_stats) p->W.stats = nullptr;
    if (t.collect_stats) HIP_OK(hipMemset(stats_mem, 0, 64));
    HIP_OK(hipMemset(p->W.meta, 0, 64));
    if (t.time_kernels && !t.use_graph)
        for (int i = 0; i < 4; ++i) HIP_OK(hipEventCreate(&p->ev[i]));
    if (t.use_graph && batch->n_edges > 0) {
        HIP_OK(hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking));
        HIP_OK(hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal));
        plan_launch_sequence(p, p->cap_stream);
        hipError_t e1 = hipStreamEndCapture(p->cap_stream, &p->graph);
        if (e1 != hipSuccess) {
            mrcc_edge_plan_destroy(p);
            return fail(MRCC_ERR_HIP, "graph capture: %s", hipGetErrorString(e1));
        }
        hipError_t e2 = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
        if (e2 != hipSuccess) {
            mrcc_edge_plan_destroy(p);
            return fail(MRCC_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e2));
        }
    }
    HIP_OK(hipStreamSynchronize((hipStream_t)stream));
    *out = p;
    return MRCC_OK;
}

int mrcc_edge_plan_run(mrcc_edge_plan* p, void* stream) {
    CHECK_INJECT();
    if (!p) return fail(MRCC_ERR_INVALID, "plan is null");
    if (p->B.n == 0) return MRCC_OK;
    hipStream_t st = (hipStream_t)stream;
    if (p->exec) {
        HIP_OK(hipGraphLaunch(p->exec, st));
    } else {
        plan_launch_sequence(p, st);
        HIP_OK(hipGetLastError());
    }
    return MRCC_OK;
}

int mrcc_edge_plan_stats(mrcc_edge_plan* p, int64_t* out8) {
    if (!p || !out8) return fail(MRCC_ERR_INVALID, "null argument");
    HIP_OK(hipDeviceSynchronize());
    unsigned long long s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (p->W.stats) HIP_OK(hipMemcpy(s, p->W.stats, sizeof(s), hipMemcpyDeviceToHost));
    for (int i = 0; i < 8; ++i) out8[i] = (int64_t)s[i];
    return MRCC_OK;
}

int mrcc_edge_plan_timing(mrcc_edge_plan* p, double* out3) {
    if (!p || !out3) return fail(MRCC_ERR_INVALID, "null argument");
    if (!p->ev[0]) return fail(MRCC_ERR_UNSUPPORTED, "plan created without time_kernels (or with a graph)");
    if (p->timed_runs == 0) return fail(MRCC_ERR_INVALID, "plan has not run");
    HIP_OK(hipEventSynchronize(p->ev[3]));
    float a = 0.f, b = 0.f, c = 0.f;
    HIP_OK(hipEventElapsedTime(&a, p->ev[0], p->ev[1]));
    HIP_OK(hipEventElapsedTime(&b, p->ev[1], p->ev[2]));
    HIP_OK(hipEventElapsedTime(&c, p->ev[0], p->ev[3]));
    out3[0] = a;
    out3[1] = b;
    out3[2] = c;
    return MRCC_OK;
}

int mrcc_edge_plan_destroy(mrcc_edge_plan* p) {
    if (!p) return MRCC_OK;
    (void)hipDeviceSynchronize();
    for (int i = 0; i < 4; ++i)
        if (p->ev[i]) (void)hipEventDestroy(p->ev[i]);
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->cap_stream) (void)hipStreamDestroy(p->cap_stream);
    if (p->work_mem) (void)hipFree(p->work_mem);
    delete p;
    return MRCC_OK;
}

int mrcc_check_edges(const mrcc_scene* scene, const mrcc_edge_batch* batch, const mrcc_tune* tune,
                     void* stream) {
    CHECK_INJECT();
    if (!scene) return fail(MRCC_ERR_INVALID, "scene is null");
    int rc = check_batch(rec_of(scene), batch);
    if (rc != MRCC_OK) return rc;
    if (batch->n_edges == 0) return MRCC_OK;
    mrcc_tune t;
    memset(&t, 0, sizeof(t));
    if (tune) t = *tune;
    else {
        t.sorted = 1;
        t.persistent = 1;
    }
    t.use_graph = 0;
    mrcc_edge_plan* p = nullptr;
    rc = mrcc_edge_plan_create(scene, batch, &t, stream, &p);
    if (rc != MRCC_OK) return rc;
    rc = mrcc_edge_plan_run(p, stream);
    long long meta[8] = {0};
    if (rc == MRCC_OK) {
        hipError_t e = hipStreamSynchronize((hipStream_t)stream);
        if (e == hipSuccess) e = hipMemcpy(meta, p->W.meta, sizeof(meta), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fail(MRCC_ERR_HIP, "edge batch: %s", hipGetErrorString(e));
    }
    mrcc_edge_plan_destroy(p);
    if (rc == MRCC_OK && meta[2] > 0)
        return fail(MRCC_ERR_INVALID, "%lld edges have N_start > N (the reference asserts) or more than %d states",
                    meta[2], MRCC_MAX_EDGE_STATES);
    return rc;
}

// The staging block of the host-pointer per-call entry points: pinned, coherent host memory that
// the kernel reads and writes in place (no copy commands).  Layout (doubles): q1[MAX_D] q2[MAX_D],
// then the mode id (int32) and the verdict byte.  stage_dev is the device's address of the block.
static int ensure_stage(SceneRec* sc) {
    if (sc->stage_host) return MRCC_OK;
    const size_t bytes = 8 * (2 * MRCC_MAX_D + 4);
    void* h = nullptr;
    HIP_OK(hipHostMalloc(&h, bytes, hipHostMallocCoherent | hipHostMallocMapped));
    void* d = nullptr;
    const hipError_t e = hipHostGetDevicePointer(&d, h, 0);
    if (e != hipSuccess) {
        (void)hipHostFree(h);
        return fail(MRCC_ERR_HIP, "hipHostGetDevicePointer failed: %s", hipGetErrorString(e));
    }
    memset(h, 0, bytes);
    sc->stage_host = (double*)h;
    sc->stage_dev = (double*)d;
    sc->stage_out_dev = (unsigned char*)(sc->stage_dev + 2 * MRCC_MAX_D + 2);
    sc->stage_out_host = (unsigned char*)(sc->stage_host + 2 * MRCC_MAX_D + 2);
    return MRCC_OK;
}

#define MRCC_NO_VERDICT 0xFF  // the kernels write 0 or 1

// Spins until the byte *p, which a kernel writes into pinned host memory, differs from `unset`, for
// up to 200 us (a per-call check takes a few), then waits for the stream (which also reports an
// error of the work queued on it).
static int wait_host_byte(volatile const unsigned char* p, unsigned char unset, hipStream_t st) {
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned k = 1;; ++k) {
        if (*p != unset) return MRCC_OK;
        if ((k & 63) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(200)) break;
    }
    HIP_OK(hipStreamSynchronize(st));
    if (*p == unset) return fail(MRCC_ERR_HIP, "the kernel wrote no result");
    return MRCC_OK;
}

// Waits for the verdict byte that the launch on `st` writes into the staging block.  The byte is
// written once, after every wave has read its rows, so the next call may refill the block.
static int wait_verdict(SceneRec* sc, hipStream_t st, int32_t* result) {
    const int rc = wait_host_byte(sc->stage_out_host, MRCC_NO_VERDICT, st);
    if (rc != MRCC_OK) return rc;
    *result = *(volatile const unsigned char*)sc->stage_out_host ? 1 : 0;
    return MRCC_OK;
}

#define MRCC_HOST_LIST_ZC 64  // host edge lists up to this length use the zero-copy block

// The short-list block: rows, int arrays and verdicts of up to MRCC_HOST_LIST_ZC edges, then the
// completion flag.
static size_t list_block_bytes() {
    return 2 * 8 * (size_t)MRCC_HOST_LIST_ZC * MRCC_MAX_D + 4 * 4 * (size_t)MRCC_HOST_LIST_ZC +
           MRCC_HOST_LIST_ZC + 16 + 16;
}

static int ensure_list_stage(SceneRec* sc) {
    if (sc->list_host) return MRCC_OK;
    void* h = nullptr;
    HIP_OK(hipHostMalloc(&h, list_block_bytes(), hipHostMallocCoherent | hipHostMallocMapped));
    void* d = nullptr;
    const hipError_t e = hipHostGetDevicePointer(&d, h, 0);
    if (e != hipSuccess) {
        (void)hipHostFree(h);
        return fail(MRCC_ERR_HIP, "hipHostGetDevicePointer failed: %s", hipGetErrorString(e));
    }
    memset(h, 0, list_block_bytes());
    sc->list_host = (unsigned char*)h;
    sc->list_dev = (unsigned char*)d;
    return MRCC_OK;
}

// Mode ids of host-side arguments: arm scenes hold modes [0, n_modes); the other kinds ignore the
// mode, as the reference's abstract environment does (abstract_env.py:255-276).
static int check_mode(const SceneRec* sc, int32_t mode_id) {
    if (sc->kind == KIND_ARMS && (mode_id < 0 || mode_id >= sc->arm.nm))
        return fail(MRCC_ERR_INVALID, "mode id %d outside [0, %d)", mode_id, sc->arm.nm);
    return MRCC_OK;
}

// One-edge view with the scalar arguments by value (no staging of them)
static EdgeBatchView one_edge_view(const SceneRec* sc, const double* q1, const double* q2, double resolution,
                                   int32_t N, int32_t N_start, int32_t N_max, int32_t include_endpoints,
                                   int32_t mode_id, uint32_t robot_mask, uint8_t* out) {
    EdgeBatchView v;
    memset(&v, 0, sizeof(v));
    v.q1 = q1;
    v.q2 = q2;
    v.rs = sc->D;
    v.ds = 1;
    v.n = 1;
    v.res = resolution;
    v.incl = include_endpoints;
    v.mask = robot_mask;
    v.out = out;
    v.N0 = N > 0 ? N : 0;
    v.Ns0 = N_start;
    v.Nm0 = N_max < 0 ? -1 : N_max;
    v.md0 = mode_id;
    return v;
}

// Copies one or two host rows into the scene's staging block and clears its verdict byte
// (synchronous callers only: the previous call's kernel has finished reading the block).
static int stage_rows(SceneRec* sc, const double* q1h, const double* q2h) {
    int rc = ensure_stage(sc);
    if (rc != MRCC_OK) return rc;
    double* h = sc->stage_host;
    if (q1h) memcpy(h, q1h, 8 * (size_t)sc->D);
    if (q2h) memcpy(h + MRCC_MAX_D, q2h, 8 * (size_t)sc->D);
    *(volatile unsigned char*)sc->stage_out_host = MRCC_NO_VERDICT;
    return MRCC_OK;
}

int mrcc_check_edge(const mrcc_scene* scene, const double* q1, const double* q2, double resolution,
                    int32_t N, int32_t N_start, int32_t N_max, int32_t include_endpoints,
                    int32_t mode_id, uint32_t robot_mask, uint8_t* out, void* stream) {
    CHECK_INJECT();
    if (!scene || !q1 || !q2 || !out) return fail(MRCC_ERR_INVALID, "null argument");
    if (N <= 0 && !(resolution > 0.0)) return fail(MRCC_ERR_INVALID, "resolution must be > 0");
    if (N_start < 0) return fail(MRCC_ERR_INVALID, "N_start < 0");
    const SceneRec* sc = rec_of(scene);
    int rc = check_mode(sc, mode_id);
    if (rc != MRCC_OK) return rc;
    // every argument travels by value: nothing host-side is shared with an earlier asynchronous call
    const EdgeBatchView v = one_edge_view(sc, q1, q2, resolution, N, N_start, N_max, include_endpoints,
                                          mode_id, robot_mask, out);
    launch_edge1(*sc, v, (hipStream_t)stream);
    HIP_OK(hipGetLastError());
    return MRCC_OK;
}

int mrcc_check_edge_host(const mrcc_scene* scene, const double* q1, const double* q2,
                         double resolution, int32_t N, int32_t N_start, int32_t N_max,
                         int32_t include_endpoints, int32_t mode_id, uint32_t robot_mask,
                         int32_t* result) {
    CHECK_INJECT();
    if (!scene || !q1 || !q2 || !result) return fail(MRCC_ERR_INVALID, "null argument");
    if (N <= 0 && !(resolution > 0.0)) return fail(MRCC_ERR_INVALID, "resolution must be > 0");
    if (N_start < 0) return fail(MRCC_ERR_INVALID, "N_start < 0");
    SceneRec* sc = (SceneRec*)rec_of(scene);
    int rc = check_mode(sc, mode_id);
    if (rc != MRCC_OK) return rc;
    int Nh = N;
    if (Nh <= 0) {  // the same binary64 rule as the device (abstract_env.py:321-323)
        double m = fabs(q1[0] - q2[0]);
        for (int k = 1; k < sc->D; ++k) m = std::max(m, fabs(q1[k] - q2[k]));
        Nh = mrcc::edge_N_from_max(m, resolution);
    }
    if (N_start > Nh) return fail(MRCC_ERR_INVALID, "N_start %d > N %d (the reference asserts)", N_start, Nh);
    if (Nh > MRCC_MAX_EDGE_STATES) return fail(MRCC_ERR_INVALID, "edge of %d states (at most %d)", Nh, MRCC_MAX_EDGE_STATES);
    hipStream_t st = nullptr;
    rc = stage_rows(sc, q1, q2);
    if (rc != MRCC_OK) return rc;
    const EdgeBatchView v = one_edge_view(sc, sc->stage_dev, sc->stage_dev + MRCC_MAX_D, resolution, N, N_start,
                                          N_max, include_endpoints, mode_id, robot_mask, sc->stage_out_dev);
    // Arm edges of more than two rounds of the one-workgroup kernel (a round checks one state per
    // wave): the same kernel on several workgroups, one state per wave for up to 64 states, the
    // verdict byte preset to free (a workgroup writes 0 on a collision), then the flag kernel,
    // whose write reaches host memory after the workgroups' (stream order).  A free edge then
    // costs about one state's evaluation instead of one per round.
    if (sc->kind == KIND_ARMS) {
        const EdgeWin wh = edge_window(Nh, N_start, N_max < 0 ? -1 : N_max, include_endpoints);
        if (!wh.err && wh.M > 2 * sc->arm_nw_edge) {
            volatile unsigned char* flag = sc->stage_out_host + 8;  // the block's last 8 bytes
            *(volatile unsigned char*)sc->stage_out_host = 1;
            *flag = 0;
            const int nw = sc->arm_nw_edge;
            launch_arm_edge_spread(*sc, v, (std::min(64, wh.M) + nw - 1) / nw, st);
            HIP_OK(hipGetLastError());
            launch_flag(sc->stage_out_dev + 8, st);
            HIP_OK(hipGetLastError());
            rc = wait_host_byte(flag, 0, st);
            if (rc != MRCC_OK) return rc;
            *result = *(volatile const unsigned char*)sc->stage_out_host ? 1 : 0;
            return MRCC_OK;
        }
    }
    launch_edge1(*sc, v, st);
    HIP_OK(hipGetLastError());
    return wait_verdict(sc, st, result);
}

// Pinned and device staging of at least `bytes` for mrcc_check_edges_host (calls are synchronous,
// so a new call never overwrites a block a previous one still reads).
static int ensure_batch_stage(SceneRec* sc, size_t bytes) {
    if (bytes <= sc->batch_cap) return MRCC_OK;
    size_t cap = std::max(bytes, 2 * sc->batch_cap);
    cap = (cap + 4095) & ~(size_t)4095;
    if (sc->batch_dev) (void)hipFree(sc->batch_dev);
    if (sc->batch_host) (void)hipHostFree(sc->batch_host);
    sc->batch_dev = nullptr;
    sc->batch_host = nullptr;
    sc->batch_cap = 0;
    HIP_OK(hipMalloc(&sc->batch_dev, cap));
    HIP_OK(hipHostMalloc(&sc->batch_host, cap));
    sc->batch_cap = cap;
    return MRCC_OK;
}

int mrcc_check_edges_host(const mrcc_scene* scene, const double* q1, const double* q2, int64_t n,
                          double resolution, const int32_t* N, const int32_t* N_start,
                          const int32_t* N_max, int32_t include_endpoints, const int32_t* mode_ids,
                          uint32_t robot_mask, uint8_t* result) {
    CHECK_INJECT();
    if (!scene) return fail(MRCC_ERR_INVALID, "scene is null");
    if (n < 0) return fail(MRCC_ERR_INVALID, "n < 0");
    if (n == 0) return MRCC_OK;
    if (!q1 || !q2 || !result) return fail(MRCC_ERR_INVALID, "null argument");
    if (!N && !(resolution > 0.0)) return fail(MRCC_ERR_INVALID, "resolution must be > 0");
    SceneRec* sc = (SceneRec*)rec_of(scene);
    const int D = sc->D;
    // the reference's argument checks, per edge (abstract_env.py:321-331); entries N <= 0 of a
    // given N array are replaced by the resolution's N
    std::vector<int32_t> Nfill;
    if (N) Nfill.assign(N, N + n);
    int maxN = 2;  // the longest edge's state count (an upper bound of every window's)
    for (int64_t i = 0; i < n; ++i) {
        int Nh = N ? N[i] : 0;
        if (Nh <= 0) {
            const double* a = q1 + i * D;
            const double* b = q2 + i * D;
            double m = fabs(a[0] - b[0]);
            for (int k = 1; k < D; ++k) m = std::max(m, fabs(a[k] - b[k]));
            Nh = mrcc::edge_N_from_max(m, resolution);
        }
        const int ns = N_start ? N_start[i] : 0;
        if (ns < 0) return fail(MRCC_ERR_INVALID, "edge %lld: N_start < 0", (long long)i);
        if (mode_ids) {
            const int rm = check_mode(sc, mode_ids[i]);
            if (rm != MRCC_OK) return rm;
        }
        if (ns > Nh) return fail(MRCC_ERR_INVALID, "edge %lld: N_start %d > N %d (the reference asserts)", (long long)i, ns, Nh);
        if (Nh > MRCC_MAX_EDGE_STATES)
            return fail(MRCC_ERR_INVALID, "edge %lld of %d states (at most %d)", (long long)i, Nh, MRCC_MAX_EDGE_STATES);
        if (N) Nfill[i] = Nh;
        maxN = std::max(maxN, Nh);
    }
    // block: q1 rows, q2 rows, then the int32 arrays given, then one result byte per edge
    const size_t rows = 8 * (size_t)n * D;
    const int32_t* ints[4] = {N ? Nfill.data() : nullptr, N_start, N_max, mode_ids};
    size_t off[4], o = 2 * rows;
    for (int j = 0; j < 4; ++j) {
        off[j] = o;
        if (ints[j]) o += 4 * (size_t)n;
    }
    const size_t o_out = o, bytes = o + (size_t)n;
    // short lists (a planner's candidates): the kernel reads the rows from pinned host memory and
    // writes the verdicts there, no copy commands; longer lists: one copy each way
    const bool zc = n <= MRCC_HOST_LIST_ZC;
    int rc = zc ? ensure_list_stage(sc) : ensure_batch_stage(sc, bytes);
    if (rc != MRCC_OK) return rc;
    unsigned char* h = zc ? sc->list_host : sc->batch_host;
    unsigned char* d = zc ? sc->list_dev : sc->batch_dev;
    const size_t o_flag = bytes;
    memcpy(h, q1, rows);
    memcpy(h + rows, q2, rows);
    for (int j = 0; j < 4; ++j)
        if (ints[j]) memcpy(h + off[j], ints[j], 4 * (size_t)n);
    memset(h + o_out, 1, (size_t)n);  // verdicts start free; waves write 0 on a collision
    hipStream_t st = nullptr;
    if (zc) *(volatile unsigned char*)(h + o_flag) = 0;
    else HIP_OK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st));
    EdgeBatchView v;
    memset(&v, 0, sizeof(v));
    v.Nm0 = -1;  // N_max None unless given
    v.q1 = (const double*)d;
    v.q2 = (const double*)(d + rows);
    v.rs = D;
    v.ds = 1;
    v.n = n;
    v.res = resolution;
    v.N = N ? (const int*)(d + off[0]) : nullptr;
    v.N_start = N_start ? (const int*)(d + off[1]) : nullptr;
    v.N_max = N_max ? (const int*)(d + off[2]) : nullptr;
    v.modes = mode_ids ? (const int*)(d + off[3]) : nullptr;
    v.incl = include_endpoints;
    v.mask = robot_mask;
    v.out = d + o_out;
    // several waves per edge for short lists (states u, u + wpe, ... of the edge), one for long
    // ones; no more than the longest edge has states (a wave without one would only stage the scene)
    int wpe = n >= 2048 ? 1 : (int)std::min<int64_t>(64, std::max<int64_t>(1, 2048 / n));
    wpe = std::max(1, std::min(wpe, maxN));
    launch_direct(*sc, v, nullptr, n, nullptr, wpe, 0, 32 * device_cus(), nullptr, st);
    HIP_OK(hipGetLastError());
    if (zc) {
        // the flag kernel runs after every wave of the check (stream order) and its write reaches
        // host memory after theirs
        launch_flag(d + o_flag, st);
        HIP_OK(hipGetLastError());
        rc = wait_host_byte(h + o_flag, 0, st);
        if (rc != MRCC_OK) return rc;
    } else {
        HIP_OK(hipMemcpyAsync(h + o_out, d + o_out, (size_t)n, hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
    }
    volatile const unsigned char* out = h + o_out;
    for (int64_t i = 0; i < n; ++i) result[i] = out[i] ? 1 : 0;
    return MRCC_OK;
}

int mrcc_check_config_host(const mrcc_scene* scene, const double* q, int32_t mode_id,
                           uint32_t robot_mask, int32_t* result) {
    CHECK_INJECT();
    if (!scene || !q || !result) return fail(MRCC_ERR_INVALID, "null argument");
    SceneRec* sc = (SceneRec*)rec_of(scene);
    int rc = check_mode(sc, mode_id);
    if (rc != MRCC_OK) return rc;
    hipStream_t st = nullptr;
    rc = stage_rows(sc, q, nullptr);
    if (rc != MRCC_OK) return rc;
    if (sc->kind == KIND_DISKS || sc->kind == KIND_PLANAR) {
        // one workgroup whose lanes split the configuration's obstacle tests (one lane walking
        // every obstacle of a free configuration set the call's latency)
        const EdgeBatchView v = one_edge_view(sc, sc->stage_dev, sc->stage_dev, 1.0, 0, 0, -1, 0, mode_id,
                                              robot_mask, sc->stage_out_dev);
        launch_config1(*sc, v, st);
        HIP_OK(hipGetLastError());
        return wait_verdict(sc, st, result);
    }
    // the mode id travels in the staging block's int area (after the two rows)
    ((int32_t*)(sc->stage_host + 2 * MRCC_MAX_D))[0] = mode_id;
    const int32_t* md = (const int32_t*)(sc->stage_dev + 2 * MRCC_MAX_D);
    launch_configs(*sc, sc->stage_dev, 1, sc->D, 1, md, robot_mask, sc->stage_out_dev, 1, st);
    HIP_OK(hipGetLastError());
    return wait_verdict(sc, st, result);
}

int mrcc_check_edges_per_call(const mrcc_scene* scene, const mrcc_edge_batch* batch, void* stream) {
    CHECK_INJECT();
    if (!scene) return fail(MRCC_ERR_INVALID, "scene is null");
    const SceneRec* sc = rec_of(scene);
    int rc = check_batch(sc, batch);
    if (rc != MRCC_OK) return rc;
    EdgeBatchView v = view_of(batch);
    hipStream_t st = (hipStream_t)stream;
    for (long long e = 0; e < batch->n_edges; ++e) {
        // one launch per edge (the plain per-call kernel): the batch view offset to edge e
        EdgeBatchView u = v;
        if (!v.idx1) u.q1 = v.q1 + e * v.rs;
        else u.idx1 = v.idx1 + e;
        if (!v.idx2) u.q2 = v.q2 + e * v.rs;
        else u.idx2 = v.idx2 + e;
        if (v.N) u.N = v.N + e;
        if (v.N_start) u.N_start = v.N_start + e;
        if (v.N_max) u.N_max = v.N_max + e;
        if (v.modes) u.modes = v.modes + e;
        u.out = v.out + e;
        u.n = 1;
        launch_edge1(*sc, u, st);
    }
    HIP_OK(hipGetLastError());
    return MRCC_OK;
}

int mrcc_bfs_indices(int32_t N, int32_t* out) {
    if (N < 1 || !out) return fail(MRCC_ERR_INVALID, "N < 1");
    std::deque<std::pair<int, int>> q;
    q.emplace_back(0, N - 1);
    int idx = 0;
    while (!q.empty()) {
        auto [s, e] = q.front();
        q.pop_front();
        const int mid = (s + e) / 2;
        out[idx++] = mid;
        if (s <= mid - 1) q.emplace_back(s, mid - 1);
        if (mid + 1 <= e) q.emplace_back(mid + 1, e);
    }
    return MRCC_OK;
}

int64_t mrcc_edge_num_states(int32_t N, int32_t N_start, int32_t N_max, int32_t include_endpoints) {
    if (N < 1 || N_start < 0 || N_start > N) return -1;
    mrcc::EdgeWin w = mrcc::edge_window(N, N_start, N_max, include_endpoints);
    if (w.full) return w.M;
    int64_t c = 0;
    for (int k = 0; k < w.M; ++k)
        if (mrcc::edge_state_index(w, k, include_endpoints) >= 0) ++c;
    return c;
}

int mrcc_arm_fk(const mrcc_scene* scene, const double* q, int64_t n, int64_t row_stride,
                int64_t dim_stride, double* out, void* stream) {
    CHECK_INJECT();
    if (!scene || !q || !out) return fail(MRCC_ERR_INVALID, "null argument");
    const SceneRec* sc = rec_of(scene);
    if (sc->kind != KIND_ARMS) return fail(MRCC_ERR_UNSUPPORTED, "FK needs an arm scene");
    if (n <= 0) return MRCC_OK;
    launch_arm_fk(*sc, q, n, row_stride, dim_stride, out, (hipStream_t)stream);
    HIP_OK(hipGetLastError());
    return MRCC_OK;
}

int mrcc_sincos(const double* x, int64_t n, double* s, double* c, void* stream) {
    CHECK_INJECT();
    if (n <= 0) return MRCC_OK;
    if (!x || !s || !c) return fail(MRCC_ERR_INVALID, "null argument");
    launch_sincos(x, n, s, c, (hipStream_t)stream);
    HIP_OK(hipGetLastError());
    return MRCC_OK;
}

int mrcc_clearance_forward(const mrcc_scene* scene, const float* q, int64_t n, int32_t dim, float margin,
                           float* cost, void* stream) {
    CHECK_INJECT();
    if (!scene || (!q && n) || (!cost && n)) return fail(MRCC_ERR_INVALID, "null argument");
    const SceneRec* sc = rec_of(scene);
    if (sc->kind != KIND_ARMS) return fail(MRCC_ERR_UNSUPPORTED, "clearance needs an arm scene");
    if (dim != sc->D) return fail(MRCC_ERR_INVALID, "rows of %d coordinates, the scene has D = %d", dim, sc->D);
    if (n <= 0) return MRCC_OK;
    launch_clearance(*sc, q, n, margin, nullptr, cost, nullptr, (hipStream_t)stream);
    HIP_OK(hipGetLastError());
    return MRCC_OK;
}

int mrcc_clearance_backward(const mrcc_scene* scene, const float* q, int64_t n, int32_t dim, float margin,
                            const float* grad_cost, float* grad_q, void* stream) {
    CHECK_INJECT();
    if (!scene || (!q && n) || (!grad_cost && n) || (!grad_q && n))
        return fail(MRCC_ERR_INVALID, "null argument");
    const SceneRec* sc = rec_of(scene);
    if (sc->kind != KIND_ARMS) return fail(MRCC_ERR_UNSUPPORTED, "clearance needs an arm scene");
    if (dim != sc->D) return fail(MRCC_ERR_INVALID, "rows of %d coordinates, the scene has D = %d", dim, sc->D);
    if (n <= 0) return MRCC_OK;
    launch_clearance(*sc, q, n, margin, grad_cost, nullptr, grad_q, (hipStream_t)stream);
    HIP_OK(hipGetLastError());
    return MRCC_OK;
}

int mrcc_config_cost(const double* p, int64_t p_row_stride, const double* q, int64_t q_row_stride,
                     int64_t n, int32_t dim, const int32_t* slices, int32_t n_slices, int32_t metric,
                     int32_t reduction, double w, double* out, void* stream) {
    CHECK_INJECT();
    if (n < 0) return fail(MRCC_ERR_INVALID, "n < 0");
    if (dim < 1 || dim > MRCC_COST_MAX_D) return fail(MRCC_ERR_INVALID, "dim must be in 1..%d", MRCC_COST_MAX_D);
    if (metric < MRCC_COST_MAX || metric > MRCC_COST_ABSMAX) return fail(MRCC_ERR_INVALID, "unknown metric %d", metric);
    if (reduction < MRCC_REDUCE_MAX_SUM || reduction > MRCC_REDUCE_MAX)
        return fail(MRCC_ERR_INVALID, "unknown reduction %d", reduction);
    if (!slices || n_slices < 1 || n_slices > MRCC_COST_MAX_SLICES)
        return fail(MRCC_ERR_INVALID, "1..%d slices required", MRCC_COST_MAX_SLICES);
    if (p_row_stride < 0 || q_row_stride < 0) return fail(MRCC_ERR_INVALID, "negative row stride");
    CostArgs A;
    memset(&A, 0, sizeof(A));
    for (int i = 0; i < n_slices; ++i) {
        const int s = slices[2 * i], e = slices[2 * i + 1];
        if (s < 0 || e > dim || s > e || (s == e && metric != MRCC_COST_EUCLIDEAN))
            return fail(MRCC_ERR_INVALID, "slice %d [%d, %d) is empty or outside [0, %d)", i, s, e, dim);
        A.sl[2 * i] = s;
        A.sl[2 * i + 1] = e;
    }
    if (n == 0) return MRCC_OK;
    if (!p || !q || !out) return fail(MRCC_ERR_INVALID, "null argument");
    A.P = p;
    A.ps = p_row_stride;
    A.Q = q;
    A.qs = q_row_stride;
    A.n = n;
    A.D = dim;
    A.ns = n_slices;
    A.metric = metric;
    A.red = reduction;
    A.w = w;
    A.out = out;
    launch_config_cost(A, 32 * device_cus(), (hipStream_t)stream);
    HIP_OK(hipGetLastError());
    return MRCC_OK;
}

}  // extern "C"
