/*
 * mrcc.h — C ABI of libmrcc.so, the MI355X (gfx950) batched collision-checking engine.
 *
 * The engine replaces the geometric part of the planners' environment collision API:
 *   BaseProblem.is_collision_free            src/multi_robot_multi_goal_planning/problems/planning_env.py:1724-1734
 *   BaseProblem.is_collision_free_for_robot  planning_env.py:1736-1744 (semantics: rai_base_env.py:515-615)
 *   BaseProblem.is_edge_collision_free       planning_env.py:1746-1763 (algorithm: abstract_env.py:301-354)
 *   BaseProblem.is_path_collision_free       planning_env.py:1765-1881
 * The reference has no FFI for this path: backends are Python BaseProblem subclasses
 * (registry.py:8-25).  The Python package binds these symbols with ctypes; every symbol below is
 * what such a binding needs.
 *
 * Conventions
 *   - Every function returns int status: MRCC_OK (0) or a negative MRCC_ERR_* code; the message of
 *     the last failure on the calling thread is returned by mrcc_last_error().  No exceptions cross
 *     the ABI.
 *   - Buffers named "device" are caller-owned device memory (e.g. torch tensors, passed as
 *     data_ptr()).  Scenes, plans and tables are library-owned.
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream).  All device work is
 *     stream-ordered; nothing synchronises unless stated.
 *   - Configuration rows: element (row r, coordinate k) lives at q[r*row_stride + k*dim_stride]
 *     (strides in elements).  row_stride=D, dim_stride=1 is the reference's flat per-config layout
 *     (NpConfiguration.q, configuration.py:222-243); row_stride=1, dim_stride=n is SoA.
 *   - Result bytes: 1 = collision free, 0 = in collision (the reference's bool).
 *   - A scene/plan handle is not thread-safe; use one per thread or lock externally.
 */
#ifndef MRCC_H
#define MRCC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MRCC_ABI_VERSION 3

enum {
    MRCC_OK = 0,
    MRCC_ERR_INVALID = -1,     /* bad argument (the reference raises ValueError / asserts) */
    MRCC_ERR_HIP = -2,         /* HIP runtime error */
    MRCC_ERR_NOMEM = -3,       /* device allocation failed */
    MRCC_ERR_UNSUPPORTED = -4, /* scene kind does not support the call (NotImplementedError) */
    MRCC_ERR_INJECTED = -5     /* failure injected by mrcc_debug_inject_failure */
};

enum { MRCC_SCENE_DISKS = 1, MRCC_SCENE_PLANAR_BOXES = 2, MRCC_SCENE_ARMS = 3 };

typedef struct mrcc_scene mrcc_scene;
typedef struct mrcc_edge_plan mrcc_edge_plan;

/* ---- library ---------------------------------------------------------------------------- */
int mrcc_abi_version(void);
const char* mrcc_last_error(void);
/* Selects the HIP device used by subsequent calls of this thread and builds per-device tables. */
int mrcc_init(int device);
/* Makes the next `count` API calls fail with `code` before touching the device (host-side tests of
 * the conservative error contract, cf. mujoco_env.py:797-807 "exception => collision"). */
int mrcc_debug_inject_failure(int code, int count);

/* ---- scenes ------------------------------------------------------------------------------ */
/* Disk robots and sphere / axis-aligned-rectangle obstacles in `robot_dim` dimensions
 * (AbstractEnvironment, abstract_env.py:41-109, :255-276).  Robot i owns coordinates
 * [i*robot_dim, (i+1)*robot_dim).  rect_size is the FULL edge length b of Rectangle(center, b)
 * (abstract_env.py:61-67).  robot_dim <= 12, n_robots <= 16. */
int mrcc_scene_create_disks(int n_robots, int robot_dim, const double* robot_radius,
                            int n_spheres, const double* sphere_center, const double* sphere_radius,
                            int n_rects, const double* rect_center, const double* rect_size,
                            mrcc_scene** out);

/* Planar oriented-box robots with (x, y, theta) each and oriented-box obstacles (x, y, yaw) —
 * the 2D rectangle scenes of the rai catalog (e.g. rai_config.py make_random_two_dim) with a
 * closed separating-axis predicate.  n_robots <= 16. */
int mrcc_scene_create_planar_boxes(int n_robots, const double* robot_half_extents /*[n,2]*/,
                                   int n_obstacles, const double* obstacle_pose /*[m,3] x,y,yaw*/,
                                   const double* obstacle_half_extents /*[m,2]*/,
                                   mrcc_scene** out);

/* Serial revolute arms (standard DH, T_k = T_{k-1} Rz(q_k) Tz(d_k) Tx(a_k) Rx(alpha_k)) with
 * link spheres and axis-aligned box obstacles — the sphere-based FK + narrow phase that the
 * native backends compute (rai_base_env.py:454-477, pinocchio_env.py:426-474,
 * mr_vamp_env.py:682-689).  base: [n_arms,12] = row-major 3x3 rotation then translation.
 * dh: [n_arms,dof,3] = (d, a, alpha).  link_sphere_count: [n_arms*dof]; spheres are listed
 * arm-major then link-major in sphere_local [S,4] = (x, y, z, r) in the link frame.
 * boxes: [n_boxes,6] = (center xyz, half-extents xyz).  n_arms <= 8, dof <= 8, 0..8 spheres/link (at most 512),
 * n_boxes <= 65535. */
int mrcc_scene_create_arms(int n_arms, int dof, const double* base, const double* dh,
                           const int32_t* link_sphere_count, const double* sphere_local,
                           int n_boxes, const double* boxes, mrcc_scene** out);

/* Serial revolute arms of any joint geometry (the robot descriptions the mujoco and pinocchio
 * backends load, mujoco_env.py / pinocchio_env.py, e.g. assets/models/mujoco/mujoco_ur10/ur10e.xml):
 * T_k = T_{k-1} Rz(q_k) C_k, with C_k a fixed transform per joint that carries link k's frame
 * (its joint axis along z) to the next.  joint_fixed: [n_arms,dof,12] = C_k row-major 3x3
 * rotation and translation, interleaved by rows: (R00 R01 R02 t0, R10 R11 R12 t1, R20 R21 R22 t2).
 * Spheres are given in the frame after C_k, as for mrcc_scene_create_arms; every other argument
 * and limit as there.  A DH joint is C = [[1,0,0,a],[0,cos alpha,-sin alpha,0],[0,sin alpha,cos alpha,d]],
 * and then both constructors give the same transforms. */
int mrcc_scene_create_arm_chains(int n_arms, int dof, const double* base, const double* joint_fixed,
                                 const int32_t* link_sphere_count, const double* sphere_local,
                                 int n_boxes, const double* boxes, mrcc_scene** out);

/* Modes of an arm scene (mode -> scene graph, planning_env.py:1711-1722; attachment semantics
 * rai_base_env.py:704-839, pinocchio_env.py:266-372).  Mode m: arms [0, n_active[m]) exist.  Each
 * mode has n_arms + n_rest box slots (at most 32); slot s of mode m is present iff
 * has_grasp[m*slots+s], with transform grasp_rel[(m*slots+s)*12 ..] (row-major R then t) and half
 * extents grasp_half[(m*slots+s)*3 ..]:
 *   slot a < n_arms  a box held by arm a, pose T_ee(q) * T_rel (T_rel frozen at mode entry);
 *   slot n_arms + j  a box resting at the world pose T_rel (a movable object put down).
 * A held box does not collide with the spheres of its holding link (the last link of its arm); it
 * is tested against every other sphere, its own arm's included, the obstacle boxes and the other
 * boxes of the mode.  A resting box is tested against every sphere and the held boxes (static
 * against static is no collision of a configuration).  Mode 0 exists by default (all arms, no
 * boxes). */
int mrcc_scene_set_modes(mrcc_scene* scene, int n_modes, const int32_t* n_active, int n_rest,
                         const uint8_t* has_grasp, const double* grasp_rel,
                         const double* grasp_half);
/* Joint types of an arm scene: joint_type[l] (l = arm * dof + k) 0 = revolute, 1 = prismatic.
 * A prismatic joint is a chain joint T <- T Tz(q) C (the reference's slide joints,
 * mujoco_env.py worlds with MJCF slide joints; URDF prismatic joints of pinocchio_env.py), with
 * C as in mrcc_scene_create_arm_chains; DH scenes reject it (MRCC_ERR_UNSUPPORTED).  travel[l]
 * bounds |q| of a prismatic joint (normally max(|lower|, |upper|) of its limits): the binary32
 * margins hold for coordinates within it, and a configuration with |q_l| > travel[l] counts as a
 * collision, as one with a non-finite coordinate does.  Default: every joint revolute. */
int mrcc_scene_set_joint_types(mrcc_scene* scene, const uint8_t* joint_type, const double* travel);
/* Arm scenes: 1 (default) = link bounding spheres cull sphere tests whose link pair (or link and
 * box) is certainly separated; 0 = every sphere pair of the pair set is tested (brute force, used
 * to measure the narrow phase against the FP32 roofline).  Results are identical. */
int mrcc_scene_set_broadphase(mrcc_scene* scene, int enabled);
int mrcc_scene_destroy(mrcc_scene* scene);
/* kind, config dimension D, robot count, mode count */
int mrcc_scene_info(const mrcc_scene* scene, int32_t* out4);

/* ---- configuration validity -------------------------------------------------------------- */
/* out[r] = is_collision_free(q_r, mode_r) (planning_env.py:1724-1734) for r < n, or, when
 * robot_mask is not all-ones, is_collision_free_for_robot(robots in mask, q_r, mode_r)
 * (planning_env.py:1736-1744; a collision counts iff it involves a robot in the mask and no
 * robot outside it, rai_base_env.py:560-586).  mode_ids: device int32[n] or NULL (mode 0). */
int mrcc_check_configs(const mrcc_scene* scene, const double* q, int64_t n, int64_t row_stride,
                       int64_t dim_stride, const int32_t* mode_ids, uint32_t robot_mask,
                       uint8_t* out, void* stream);

/* ---- edge validity ----------------------------------------------------------------------- */
/* One batch of edges: out[e] = is_edge_collision_free(q1_e, q2_e, mode_e, resolution,
 * include_endpoints=include_endpoints, N_start=N_start[e], N_max=N_max[e], N=N[e])
 * (abstract_env.py:301-354).  Row of q1 for edge e is idx1[e] (or e when idx1 is NULL), same for
 * q2 / idx2, so a path or a roadmap can pass its vertex array once. */
typedef struct mrcc_edge_batch {
    const double* q1;          /* device */
    const double* q2;          /* device */
    int64_t row_stride;
    int64_t dim_stride;
    const int32_t* idx1;       /* device int32[n_edges] or NULL */
    const int32_t* idx2;       /* device int32[n_edges] or NULL */
    int64_t n_edges;
    double resolution;         /* > 0 */
    const int32_t* N;          /* device int32[n_edges] or NULL: N = max(2, int(maxabs/res)+1) */
    const int32_t* N_start;    /* device int32[n_edges] or NULL: 0 */
    const int32_t* N_max;      /* device int32[n_edges] or NULL: N ; negative entries mean None */
    int32_t include_endpoints;
    const int32_t* mode_ids;   /* device int32[n_edges] or NULL */
    uint32_t robot_mask;       /* 0xffffffff = full check */
    uint8_t* out;              /* device uint8[n_edges] */
} mrcc_edge_batch;

/* Execution shape of the batched edge check (chosen by the autotuner when NULL / zero). */
typedef struct mrcc_tune {
    int32_t slots_per_pass;    /* states of one edge issued per pass (coarse-to-fine order) */
    int32_t slots_per_ticket;  /* queue slots a wave claims per atomic ticket (arm scenes:
                                  <= 64, one lane resolves one slot) */
    int32_t waves_per_cu;      /* persistent grid = waves_per_cu * CUs */
    int32_t sorted;            /* 1: edges ordered longest-first */
    int32_t persistent;        /* 1: persistent work-queue kernel; 0: one workgroup per edge */
    int32_t use_graph;         /* 1: capture the kernel sequence in a hipGraph at plan creation */
    int32_t collect_stats;     /* 1: count evaluated states and FP64 rechecks */
    int32_t time_kernels;      /* 1: HIP events around the queue build and the main kernel of
                                  each run (plans without a graph); see mrcc_edge_plan_timing */
} mrcc_tune;

/* Plan = workspace + (optionally) an instantiated hipGraph bound to the batch's buffers.
 * Creating a plan synchronises `stream`; running it is asynchronous. */
int mrcc_edge_plan_create(const mrcc_scene* scene, const mrcc_edge_batch* batch,
                          const mrcc_tune* tune, void* stream, mrcc_edge_plan** out);
int mrcc_edge_plan_run(mrcc_edge_plan* plan, void* stream);
/* Synchronises; out8 = {states evaluated, pair rechecks in FP64, slots skipped (edge already in
 * collision), queue slots issued, and for arm scenes the tests executed: link-box, sphere-box,
 * link-pair, sphere-pair} since plan creation (zeros unless collect_stats). */
int mrcc_edge_plan_stats(mrcc_edge_plan* plan, int64_t* out8);
/* Synchronises; out3 = {queue build, main kernel, whole run} in milliseconds of the last run,
 * from HIP events recorded on the run's stream.  Needs time_kernels = 1 and use_graph = 0
 * (MRCC_ERR_UNSUPPORTED otherwise).  No counterpart in the reference: instrumentation. */
int mrcc_edge_plan_timing(mrcc_edge_plan* plan, double* out3);
int mrcc_edge_plan_destroy(mrcc_edge_plan* plan);
/* One-shot batched check (plan + run + destroy; synchronises). */
int mrcc_check_edges(const mrcc_scene* scene, const mrcc_edge_batch* batch,
                     const mrcc_tune* tune, void* stream);
/* Plain per-call form: one kernel launch checks one edge (q1, q2: device double[D]), the way
 * the unchanged planners call is_edge_collision_free.  N <= 0: computed; N_max < 0: None. */
int mrcc_check_edge(const mrcc_scene* scene, const double* q1, const double* q2,
                    double resolution, int32_t N, int32_t N_start, int32_t N_max,
                    int32_t include_endpoints, int32_t mode_id, uint32_t robot_mask,
                    uint8_t* out, void* stream);
/* The planners' call pattern on host memory: q1, q2 (or q) are host double[D]; they are copied
 * into a pinned, coherent staging block of the scene that the kernel reads in place and into which
 * it writes the verdict (no copy commands); one kernel launch on the null stream; the call returns
 * once the verdict has arrived (spinning on it for up to 200 us, then waiting for the stream).
 * *result = 1 when collision free, 0 otherwise (is_edge_collision_free / is_collision_free of one
 * call).  Arm edges of more than two rounds of states (a round: one state per wave of the
 * workgroup) run on several workgroups, one state per wave up to 64 states, followed by a
 * one-thread flag kernel that the call spins on.  A configuration of a disk or planar scene is
 * checked by one workgroup whose lanes split its obstacle tests.  Calls on one scene must not overlap (they
 * share the staging block). */
int mrcc_check_edge_host(const mrcc_scene* scene, const double* q1, const double* q2,
                         double resolution, int32_t N, int32_t N_start, int32_t N_max,
                         int32_t include_endpoints, int32_t mode_id, uint32_t robot_mask,
                         int32_t* result);
int mrcc_check_config_host(const mrcc_scene* scene, const double* q, int32_t mode_id,
                           uint32_t robot_mask, int32_t* result);
/* A short list of edges on host memory (a planner's candidate list: find_parent / rewire
 * rrtstar_base.py:1339-1446, PRM edges prm_graph.py:663-684, EIT windows planner_eitstar.py:630-665):
 * rows q1[i*D ..], q2[i*D ..] (host double), optional per-edge host int32 arrays N (<= 0: from the
 * resolution), N_start, N_max (< 0: None), mode_ids (NULL: defaults).  Lists of up to 64 edges
 * go through a pinned coherent block of the scene that the kernel reads and writes in place (the
 * host spins on a flag written after the check); longer lists through one pinned staging copy each
 * way.  One launch (several wavefronts per edge for short lists), synchronous.
 * result[i] = 1 when edge i is collision free (is_edge_collision_free, planning_env.py:1746-1763),
 * 0 otherwise.  N_start > N (the reference asserts) or N_start < 0 -> MRCC_ERR_INVALID. */
int mrcc_check_edges_host(const mrcc_scene* scene, const double* q1, const double* q2, int64_t n,
                          double resolution, const int32_t* N, const int32_t* N_start,
                          const int32_t* N_max, int32_t include_endpoints, const int32_t* mode_ids,
                          uint32_t robot_mask, uint8_t* result);
/* The plain per-call form applied to every edge of a batch: n_edges launches of the
 * mrcc_check_edge kernel (host loop, no Python in between). */
int mrcc_check_edges_per_call(const mrcc_scene* scene, const mrcc_edge_batch* batch, void* stream);
/* The batched check of mrcc_check_edges without any device allocation.  No counterpart in the
 * reference: it exists for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The plan workspace is a file-scope device
 * array of 16 MiB per GPU (kept until the process ends): (24 + 8 D) bytes per edge plus about 264
 * KiB of tables, so about 65,000 edges of 28 coordinates; a larger batch is rejected with
 * MRCC_ERR_INVALID.  Stream-ordered and asynchronous (nothing synchronises, nothing is allocated):
 * edges the reference would reject (N_start > N, more than 2^24 states, an unknown mode id) get
 * verdict 0, and when n_invalid (device int64) is given their count is written there.  tune: NULL
 * or a persistent shape without statistics or timing (use_graph is ignored).  Every call on one GPU
 * shares that GPU's workspace: calls that overlap on one GPU (other streams or threads) race. */
int mrcc_check_edges_noalloc(const mrcc_scene* scene, const mrcc_edge_batch* batch,
                             const mrcc_tune* tune, int64_t* n_invalid, void* stream);

/* ---- host helpers (no device work) ------------------------------------------------------- */
/* generate_binary_search_indices(N) (planning_env.py:34-51) into out[N]. */
int mrcc_bfs_indices(int32_t N, int32_t* out);
/* Number of states the reference edge loop passes to is_collision_free (abstract_env.py:337-352). */
int64_t mrcc_edge_num_states(int32_t N, int32_t N_start, int32_t N_max, int32_t include_endpoints);

/* ---- kinematics (device) ----------------------------------------------------------------- */
/* FP64 forward kinematics of an arm scene: out[r][link][12] (row-major R then t) for all
 * n_arms*dof links of each config row; the exact transforms the FP64 near-contact recheck uses. */
int mrcc_arm_fk(const mrcc_scene* scene, const double* q, int64_t n, int64_t row_stride,
                int64_t dim_stride, double* out, void* stream);
/* The engine's FP64 sincos on device (x, s, c: device double[n]). */
int mrcc_sincos(const double* x, int64_t n, double* s, double* c, void* stream);

/* ---- differentiable clearance cost (arm scenes, FP32) ------------------------------------ */
/* cost[r] = sum over the checked pair set of max(0, margin + r_i + r_j - d_ij)^2 (sphere-sphere)
 * and max(0, margin + r_i - d(sphere, box))^2 (sphere-box) for config row r (q: device float,
 * rows contiguous of dim coordinates; dim must equal the scene's D = n_arms*dof, else
 * MRCC_ERR_INVALID). */
int mrcc_clearance_forward(const mrcc_scene* scene, const float* q, int64_t n, int32_t dim, float margin,
                           float* cost, void* stream);
/* grad_q[r][k] = grad_cost[r] * d cost[r] / d q[r][k]; FK is recomputed, nothing is stored by
 * the forward. */
int mrcc_clearance_backward(const mrcc_scene* scene, const float* q, int64_t n, int32_t dim, float margin,
                            const float* grad_cost, float* grad_q, void* stream);

/* ---- configuration-pair costs and distances (device) ------------------------------------- */
/* The planners' metric functions (core/configuration.py):
 *   batch_config_cost / _batch_config_cost_impl  :437-510 (config_cost :352-363)
 *   NpConfiguration._batch_dist                   :303-329 (batch_config_dist :342-349)
 * For row r < n the difference d_k = p[r*p_row_stride + k] - q[r*q_row_stride + k] (k < dim;
 * p_row_stride = 0 broadcasts one configuration) is reduced over each slice i, coordinates
 * [slices[2i], slices[2i+1]), to a value v_i by `metric`:
 *   MRCC_COST_MAX        np.max(np.abs(d[s:e])) (:503; a NaN propagates)
 *   MRCC_COST_EUCLIDEAN  sqrt of sum_squared += d_k * d_k in coordinate order, contracted
 *                        (compute_sliced_euclidean_dists :101-126)
 *   MRCC_COST_ABSMAX     m = |d_s|, then m = |d_k| > m ? |d_k| : m (compute_abs_max_reduction
 *                        :174-197)
 * and the v_i are reduced by `reduction`:
 *   MRCC_REDUCE_MAX_SUM  max_i v_i + w * (v_0 + v_1 + ...), contracted (compute_max_sum_reduction
 *                        :150-171; max by v_i > m)
 *   MRCC_REDUCE_SUM      0.0 + v_0 + v_1 + ... (compute_sum_reduction :129-147)
 *   MRCC_REDUCE_MAX      m = v_0, then v_i > m ? v_i : m (compute_max_reduction :200-219)
 * p, q, out: device double; slices: host int32[2*n_slices], 1 <= n_slices <= 64, 0 <= s <= e <= dim
 * (s < e unless metric is MRCC_COST_EUCLIDEAN); 1 <= dim <= 1024.  Stream-ordered. */
enum { MRCC_COST_MAX = 0, MRCC_COST_EUCLIDEAN = 1, MRCC_COST_ABSMAX = 2 };
enum { MRCC_REDUCE_MAX_SUM = 0, MRCC_REDUCE_SUM = 1, MRCC_REDUCE_MAX = 2 };
int mrcc_config_cost(const double* p, int64_t p_row_stride, const double* q, int64_t q_row_stride,
                     int64_t n, int32_t dim, const int32_t* slices, int32_t n_slices, int32_t metric,
                     int32_t reduction, double w, double* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* MRCC_H */
