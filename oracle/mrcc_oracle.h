/*
 * mrcc_oracle.h — CPU oracle of the collision-checking hot path (FP64, plain C).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and only as the checker / the timed CPU baseline.  The product (libmrcc.so and
 * the Python package) never calls it.
 *
 * What it restates (reference = vhartman/multirobot-pathplanning-benchmark, paths relative to
 * src/multi_robot_multi_goal_planning/problems/):
 *   or_bfs_indices            planning_env.py:34-51   generate_binary_search_indices
 *   or_edge_N                 abstract_env.py:321-323 + core/configuration.py:174-197, :328-339
 *   or_disk_config_free       abstract_env.py:255-276 (+ Sphere :41-49, Rectangle :61-84)
 *   or_edge_free              abstract_env.py:301-354 is_edge_collision_free
 *   or_*_config_free(mask)    planning_env.py:1736-1744 / rai_base_env.py:515-615 per-robot rule
 * and the engine's own geometric predicates for scenes the reference computes inside external
 * libraries (rai / pinocchio / VAMP; no in-tree implementation):
 *   or_planar_config_free     2D oriented boxes, closed separating-axis test
 *   or_arm_fk / or_arm_config_free   DH arms, link spheres, AABB obstacles, grasped boxes
 *   or_scene_arm_chains       arms of any joint geometry (T <- T Rz(q) C), as the mujoco /
 *                             pinocchio robot descriptions give them
 *   or_clearance_cost / _grad        summed squared penetration and its joint gradient
 * and the planners' cost / distance functions beside the path (core/configuration.py):
 *   or_config_cost            :101-219 numba reductions, :303-329 _batch_dist, :437-510 batch_config_cost
 * Arithmetic is IEEE binary64 with contraction disabled (-ffp-contract=off); where the reference
 * goes through numpy the numpy rounding sequence is reproduced (fused np.linalg.norm on small
 * vectors, pairwise np.sum, libm pow for r**2).
 */
#ifndef MRCC_ORACLE_H
#define MRCC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_scene or_scene;

/* FP64 sincos of the engine (Cody-Waite reduction by pi/2 in three parts, Taylor polynomials on
 * [-pi/4, pi/4], Horner without fused multiply-add). */
void or_sincos(double x, double* s, double* c);

int or_bfs_indices(int N, int32_t* out);
int or_edge_N(const double* q1, const double* q2, int D, double res);
int64_t or_edge_num_states(int N, int N_start, int N_max, int include_endpoints);

or_scene* or_scene_disks(int n_robots, int robot_dim, const double* robot_radius, int n_spheres,
                         const double* sphere_center, const double* sphere_radius, int n_rects,
                         const double* rect_center, const double* rect_size);
or_scene* or_scene_planar(int n_robots, const double* robot_half, int n_obs,
                          const double* obs_pose, const double* obs_half);
or_scene* or_scene_arms(int n_arms, int dof, const double* base, const double* dh,
                        const int32_t* link_sphere_count, const double* sphere_local,
                        int n_boxes, const double* boxes);
/* arms whose joints are T <- T Rz(q) C, C [dof*n_arms,12] rows (R00 R01 R02 t0, ..., t2) */
or_scene* or_scene_arm_chains(int n_arms, int dof, const double* base, const double* joint_fixed,
                              const int32_t* link_sphere_count, const double* sphere_local,
                              int n_boxes, const double* boxes);
/* joint types of an arm scene, types [n_arms*dof]: 0 revolute, 1 prismatic (chain scenes only:
 * T <- T Tz(q) C); a prismatic joint with |q| > travel[l] makes the configuration a collision */
int or_scene_set_joint_types(or_scene* s, const uint8_t* types, const double* travel);
/* modes: slots per mode = n_arms + n_rest; slot a < n_arms is the box held by arm a (pose
 * T_ee(q) rel), slot n_arms + j a box resting at world pose rel; has_grasp [n_modes*slots] */
int or_scene_set_modes(or_scene* s, int n_modes, const int32_t* n_active, int n_rest,
                       const uint8_t* has_grasp, const double* grasp_rel,
                       const double* grasp_half);
void or_scene_free(or_scene* s);
int or_scene_dim(const or_scene* s);

/* 1 = collision free, 0 = collision, <0 = error */
int or_config_free(const or_scene* s, const double* q, int mode, uint32_t mask);
/* N <= 0: computed; N_max < 0: None.  *n_checked (may be NULL) += states evaluated. */
int or_edge_free(const or_scene* s, const double* q1, const double* q2, double res, int N,
                 int N_start, int N_max, int include_endpoints, int mode, uint32_t mask,
                 int64_t* n_checked);

/* batch drivers (OpenMP over rows / edges; threads <= 0: all) — the CPU baseline */
int or_check_configs(const or_scene* s, const double* q, int64_t n, int64_t row_stride,
                     const int32_t* modes, uint32_t mask, uint8_t* out, int threads);
int or_check_edges(const or_scene* s, const double* q1, const double* q2, int64_t n,
                   int64_t row_stride, double res, const int32_t* N, const int32_t* N_start,
                   const int32_t* N_max, int include_endpoints, const int32_t* modes,
                   uint32_t mask, uint8_t* out, int threads, int64_t* n_checked);

/* arms: FP64 forward kinematics, out[(arm*dof+link)*12 ..] */
int or_arm_fk(const or_scene* s, const double* q, double* out);
/* clearance cost (mode 0, all arms) and its gradient w.r.t. q (FP64) */
double or_clearance_cost(const or_scene* s, const double* q, double margin);
int or_clearance_grad(const or_scene* s, const double* q, double margin, double* grad);
int or_clearance_cost_batch(const or_scene* s, const double* q, int64_t n, double margin,
                            double* cost, int threads);

/* Costs / distances of configuration pairs (core/configuration.py): per row r, d_k =
 * p[r*ps + k] - q[r*qs + k]; per slice [sl[2i], sl[2i+1]) a value by `metric`
 * (0 np.max(np.abs) :503, 1 compute_sliced_euclidean_dists :101-126, 2
 * compute_abs_max_reduction :174-197), then a reduction over the slices (0
 * compute_max_sum_reduction :150-171, 1 compute_sum_reduction :129-147, 2 compute_max_reduction
 * :200-219).  Contracted steps are explicit fma().  Returns 0, or -1 for a bad argument. */
int or_config_cost(const double* p, int64_t ps, const double* q, int64_t qs, int64_t n, int dim,
                   const int32_t* sl, int n_slices, int metric, int reduction, double w, double* out,
                   int threads);

/* single predicates (for pinning tests): 1 overlap / 0 separated */
int or_test_obb2(const double* A8, const double* B8);
int or_test_obb3(const double* A12, const double* a3, const double* B12, const double* b3);
int or_test_sph_obb(const double* c3, double r, const double* G12, const double* h3);

#ifdef __cplusplus
}
#endif
#endif
