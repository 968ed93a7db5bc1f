// percall_bench.cpp — times the planners' per-call pattern through the C ABI, without Python.
//
// Counterpart of the reference's cpp/main.cpp:83-104 (10,000 is_collision_free calls, ms/call):
// builds the abstract.test scene (abstract_env.py:363-421), then times 10,000
// mrcc_check_config_host calls and 10,000 mrcc_check_edge_host calls on random configurations in
// [-2, 2]^4 (resolution 0.01) and prints ms per call.
//
// Build:  hipcc -O2 -std=c++17 -I../../include percall_bench.cpp -L.. -lmrcc -Wl,-rpath,'$ORIGIN/..' -o percall_bench
#include <chrono>
#include <cstdio>
#include <random>

#include "mrcc.h"

int main() {
    if (mrcc_init(0) != MRCC_OK) {
        std::fprintf(stderr, "mrcc_init: %s\n", mrcc_last_error());
        return 1;
    }
    const double radius[2] = {0.1, 0.1};
    const double sphere_c[2] = {0.0, 0.0}, sphere_r[1] = {0.2};
    const double rect_c[2] = {0.0, 0.4}, rect_b[2] = {0.5, 0.5};
    mrcc_scene* sc = nullptr;
    if (mrcc_scene_create_disks(2, 2, radius, 1, sphere_c, sphere_r, 1, rect_c, rect_b, &sc) != MRCC_OK) {
        std::fprintf(stderr, "scene: %s\n", mrcc_last_error());
        return 1;
    }
    std::mt19937_64 rng(0);
    std::uniform_real_distribution<double> U(-2.0, 2.0);
    const int n = 10000;
    int32_t res = 0;
    long n_free = 0;
    double q[4], q2[4];
    for (int w = 0; w < 100; ++w) {  // warm-up
        for (double& v : q) v = U(rng);
        mrcc_check_config_host(sc, q, 0, 0xffffffffu, &res);
    }
    auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < n; ++i) {
        for (double& v : q) v = U(rng);
        if (mrcc_check_config_host(sc, q, 0, 0xffffffffu, &res) != MRCC_OK) return 1;
        n_free += res;
    }
    auto t1 = std::chrono::steady_clock::now();
    long e_free = 0;
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < 4; ++k) {
            q[k] = U(rng);
            q2[k] = q[k] + 0.1 * U(rng);
        }
        if (mrcc_check_edge_host(sc, q, q2, 0.01, 0, 0, -1, 0, 0, 0xffffffffu, &res) != MRCC_OK) return 1;
        e_free += res;
    }
    auto t2 = std::chrono::steady_clock::now();
    const double ms_c = std::chrono::duration<double, std::milli>(t1 - t0).count() / n;
    const double ms_e = std::chrono::duration<double, std::milli>(t2 - t1).count() / n;
    std::printf("{\"config_ms_per_call\": %.5f, \"edge_ms_per_call\": %.5f, \"configs_free\": %ld, \"edges_free\": %ld, \"calls\": %d}\n",
                ms_c, ms_e, n_free, e_free, n);
    mrcc_scene_destroy(sc);
    return 0;
}
