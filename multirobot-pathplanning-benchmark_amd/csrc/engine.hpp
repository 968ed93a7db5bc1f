// engine.hpp — host-side records of the collision engine and the kernel launchers.
#pragma once

#include "../../include/mrcc.h"
#include "evaluators.hpp"

// A scene as held by the library: the kind, its device view and the device allocation behind it.
struct SceneRec {
    int kind;
    int D;
    int n_robots;
    int device;
    DiskView disk;
    PlanarView planar;
    ArmView arm;          // every array of an arm scene (clearance and kinematics kernels)
    ArmK armk;            // the edge and configuration kernels' view of the same scene
    size_t arm_lds;       // dynamic LDS bytes of the arm kernels (arm_nw waves per workgroup)
    size_t arm_lds1;      // dynamic LDS bytes of an arm kernel with one wave per workgroup
    int arm_nw;           // waves per workgroup of the arm kernels
    int arm_nw_pers;      // waves per workgroup of the persistent kernel
    int arm_wg_pers;      // its resident workgroups per CU (registers and LDS)
    int arm_blocks_pers;  // its largest useful grid: resident workgroups on every CU
    size_t arm_lds_pers;  // its dynamic LDS bytes
    int arm_nw_edge;      // waves of the per-call kernel's one workgroup
    size_t arm_lds_edge;  // its dynamic LDS bytes
    void* dev_mem;        // one allocation holding every array the views point to
    size_t dev_bytes;
    // host copies needed by host code
    int arm_na, arm_dof, arm_S, arm_nm;
    int arm_has_grasp;    // some mode holds a box: the evaluator with the held-box phase
    // staging block of the host-pointer per-call entry points (lazily allocated): pinned coherent
    // host memory, read and written by the kernels in place; stage_dev is its device address
    double* stage_dev;
    double* stage_host;
    unsigned char* stage_out_dev;
    unsigned char* stage_out_host;
    // staging of the host-memory batch entry point (grown on demand)
    unsigned char* batch_dev;
    unsigned char* batch_host;
    size_t batch_cap;
    // block of the short host lists (at most MRCC_HOST_LIST_ZC edges): pinned coherent host memory
    // that the kernels read and write in place; list_dev is its device address; a flag byte after
    // the verdicts is set by a one-thread kernel after the check
    unsigned char* list_host;
    unsigned char* list_dev;
};

void launch_plan_build(const SceneRec& sc, const EdgeBatchView& B, const PlanWork& W, int lgG,
                       hipStream_t st);
void launch_plan_run(const SceneRec& sc, const EdgeBatchView& B, const PlanWork& W, int lgG,
                     int spt, int grid, hipStream_t st);
void launch_direct(const SceneRec& sc, const EdgeBatchView& B, const int* list, long long n_list,
                   const long long* n_dev, int wpe, int init_out, int max_grid,
                   unsigned long long* stats, hipStream_t st);
void launch_fill_u8(unsigned char* p, long long n, unsigned char v, hipStream_t st);
// plain per-call form: edge 0 of B in one workgroup (k_arm_edge / k_lane_edge)
void launch_edge1(const SceneRec& sc, const EdgeBatchView& B, hipStream_t st);
// one thread writes 1 to *flag (after the work queued before it on the stream)
void launch_flag(unsigned char* flag, hipStream_t st);
// one configuration (row q1 of B) of a disk or planar scene: the per-call lane kernel, obstacles
// split over lanes
void launch_config1(const SceneRec& sc, const EdgeBatchView& B, hipStream_t st);
// k_arm_edge on `blocks` workgroups (blocks > 1: the verdict must be preset to 1)
void launch_arm_edge_spread(const SceneRec& sc, const EdgeBatchView& B, int blocks, hipStream_t st);
void launch_configs(const SceneRec& sc, const double* q, long long n, long long rs, long long ds,
                    const int* modes, unsigned mask, unsigned char* out, int max_grid,
                    hipStream_t st);
void launch_arm_fk(const SceneRec& sc, const double* q, long long n, long long rs, long long ds,
                   double* out, hipStream_t st);
void launch_sincos(const double* x, long long n, double* s, double* c, hipStream_t st);
void launch_clearance(const SceneRec& sc, const float* q, long long n, float margin,
                      const float* grad_cost, float* cost, float* grad_q, hipStream_t st);

// configuration-pair costs (cost.hip): row r of the result reduces d = P[r*ps + k] - Q[r*qs + k]
// (k < D) over the slices [sl[2i], sl[2i+1])
#define MRCC_COST_MAX_SLICES 64
#define MRCC_COST_MAX_D 1024
#define MRCC_COST_TILE_BYTES 8192  // LDS of one workgroup's tile of rows (narrow rows)
struct CostArgs {
    const double* P;
    long long ps;
    const double* Q;
    long long qs;
    long long n;
    int D, ns, metric, red;
    double w;
    double* out;
    int sl[2 * MRCC_COST_MAX_SLICES];
};
void launch_config_cost(const CostArgs& A, int max_grid, hipStream_t st);

// api.hip internals shared with the no-allocation entry point (noalloc.hip)
namespace mrcc_internal {
int fail(int code, const char* fmt, ...);  // sets the thread's last-error message, returns code
int injected();                            // an injected failure (mrcc_debug_inject_failure), or 0
int device_cus();
const SceneRec* scene_rec(const mrcc_scene* s);
int check_batch(const SceneRec* sc, const mrcc_edge_batch* b);
EdgeBatchView batch_view(const mrcc_edge_batch* b);
// execution shape: defaults filled in and checked (mrcc_edge_plan_create's rules)
int tune_defaults(const SceneRec* sc, const mrcc_tune* tune, mrcc_tune* t);
int plan_grid(const SceneRec* sc, const mrcc_tune& t);
// Bytes of a plan's workspace for n_edges edges of D coordinates; with W, lays it out from base.
size_t plan_layout(long long n_edges, int D, unsigned char* base, PlanWork* W);
// queue build, persistent kernel and long-edge fallback on stream st
void launch_plan_sequence(const SceneRec& sc, const EdgeBatchView& B, const PlanWork& W, int lgG, int spt,
                          int grid, hipStream_t st);
}  // namespace mrcc_internal

// noalloc.hip: looks up the current device's scratch array (called by mrcc_init)
int mrcc_noalloc_setup();
