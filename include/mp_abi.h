/*
 * mp_abi.h -- C ABI of the MI355X (gfx950) motion-planning hot path.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes.  Device pointers are
 * HIP device addresses; streams are hipStream_t passed as void* (NULL = the library's
 * stream).  Every function returns MP_OK (0) or a negative MP_E* code; mp_last_error()
 * returns a message for the calling thread's last failure.  Nothing here exits the process.
 *
 * Paths below are relative to the reference checkout (jih189/isaac_sim_ros2_humble).  Each
 * entry point names the reference interface it replaces.
 */
#ifndef MP_ABI_H
#define MP_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MP_OK 0
#define MP_E_INVALID (-1)  /* bad argument / size mismatch (reference: std::runtime_error) */
#define MP_E_HIP (-2)      /* HIP runtime error (reference: CUDA_CHECK -> exit) */
#define MP_E_RTC (-3)      /* hipRTC compile error (reference: NVRTC_SAFE_CALL -> exit) */
#define MP_E_NOMEM (-4)    /* device allocation failed (reference: isValid()==false / nullptr) */
#define MP_E_CAPACITY (-5) /* obstacle cache / tree / path capacity exceeded */
#define MP_E_STATE (-6)    /* library not initialised / object in wrong state */

typedef struct mp_robot mp_robot;
typedef struct mp_scene mp_scene;
typedef struct mp_server mp_server;

/* ---------------------------------------------------------------------------------------- */
/* Lifecycle and device plumbing                                                            */
/* ---------------------------------------------------------------------------------------- */

/* Select the HIP device and create the library stream.  Replaces the implicit context
 * creation of KernelFunction::create (cudampl/CUDAMPLib/include/NvrtcUtil.h:127-139). */
int mp_init(int device);
int mp_shutdown(void);
const char* mp_last_error(void);
/* Directory of the on-disk code-object cache (default: $MP_KERNEL_CACHE or
 * <library dir>/_kcache).  Replaces the CWD "<kernel>.ptx" cache (NvrtcUtil.h:62-125), keyed
 * by a hash of the generated source, compile options and hipRTC version. */
int mp_set_cache_dir(const char* dir);

int mp_device_alloc(void** dptr, size_t bytes);
int mp_device_free(void* dptr);
int mp_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream);
int mp_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream);
int mp_memset_d8(void* dst, int value, size_t bytes, void* stream);
int mp_stream_sync(void* stream);
void* mp_default_stream(void);
/* Elapsed ms of the most recent launch of the named kernel family on the library stream
 * ("check", "fk", "edges", "nn", "sample", "plan"), measured with HIP events around it. */
double mp_last_kernel_ms(const char* family);
/* Algorithmic flops (FK, sphere transforms, bounding-sphere and pair tests actually evaluated,
 * at the per-operation costs of the planning cost model) accumulated by the check / cost /
 * edge kernels since the last reset. */
int mp_work_flops(uint64_t* out, int reset);
/* Planner phase profile (diagnostic; no counterpart in the reference).  While enabled, planning
 * launches add the wall-clock ticks (100 MHz) each wave spends per phase, and event counts, to
 * library counters: [0] scheduling, [1] sample + nearest neighbour, [2] extension check,
 * [3] node publication, [4] connection nearest neighbour, [5] connection checks, [6] path
 * extraction, [7] whole iterations, [8] wave lifetime, [9] iterations, [10] inserts,
 * [11] reserved (0).  enable: 1 on, 0 off, -1 unchanged.  counters (may be NULL)
 * receives the first n (<= MP_PLAN_PROFILE_N) counters, which are then reset. */
#define MP_PLAN_PROFILE_N 12
int mp_plan_profile(int enable, uint64_t* counters, int n);

/* ---------------------------------------------------------------------------------------- */
/* Robot                                                                                    */
/* ---------------------------------------------------------------------------------------- */
/* RobotInfo layout (src/foliation_planner/include/foliation_planner/robot_info.hpp:13-220),
 * as taken by SingleArmSpace (cudampl/CUDAMPLib/include/spaces/SingleArmSpace.h:24-40) and
 * CPRRTC::RobotSolver (cudampl/CPRRTC/include/RobotSolver.h:84-101). */
typedef struct {
    const char* name;
    int num_links;                 /* L = number of joints (one joint per link, root first) */
    const int* joint_types;        /* [L] 0 root, 1 revolute, 2 prismatic, 5 fixed */
    const double* joint_origins;   /* [L][16] row-major Isometry3d::matrix() */
    const double* joint_axes;      /* [L][3] */
    const int* parents;            /* [L] parent link index, -1 for the root */
    const int* active;             /* [L] 1 if the joint is a planning DOF */
    const float* lower;            /* [L] (active joints) */
    const float* upper;            /* [L] */
    const float* defaults;         /* [L] values of inactive joints */
    int num_spheres;               /* S */
    const int* sphere_link;        /* [S] */
    const float* sphere_pos;       /* [S][3] centre in link frame */
    const float* sphere_radius;    /* [S] */
    const unsigned char* collision_enabled; /* [L][L] link-pair self-collision enable map */
} mp_robot_desc;

/* Generate robot-specialised HIP source (unrolled FK, self-collision pair set) and compile it
 * with hipRTC for gfx950 (cached).  Replaces SingleArmSpace's ctor JIT build
 * (cudampl/CUDAMPLib/src/spaces/SingleArmSpace.cu:10-161) and RobotSolver's
 * (cudampl/CPRRTC/src/RobotSolver.cu:37-141). */
int mp_robot_create(const mp_robot_desc* desc, mp_robot** out);
int mp_robot_destroy(mp_robot* robot);
/* dof, spheres, self pairs, links */
int mp_robot_info(const mp_robot* robot, int* dof, int* num_spheres, int* num_pairs, int* num_links);
/* Generated source text (for inspection); returns its length, copies up to cap bytes. */
size_t mp_robot_source(const mp_robot* robot, char* buf, size_t cap);
/* Compile-only: generate + hipRTC-compile + store in the cache, no device needed.  Used by
 * the build step to pre-populate the cache for the bundled robots. */
int mp_robot_precompile(const mp_robot_desc* desc);
/* Compile-only for a robot set (the module a mixed batch or a server uses). */
int mp_precompile(const mp_robot_desc* const* descs, int n);
/* Generated module source for a robot set without compiling it (no device needed); returns
 * the length, copies up to cap-1 bytes plus a terminating NUL. */
size_t mp_generate_source(const mp_robot_desc* const* descs, int n, char* buf, size_t cap);

/* ---------------------------------------------------------------------------------------- */
/* Scene                                                                                    */
/* ---------------------------------------------------------------------------------------- */
/* Obstacle records (float32, base frame):
 *   spheres   [ns][4]   x, y, z, radius                 EnvConstraintSphere (EnvConstraintSphere.cu:10-33),
 *                                                       CPRRTC Sphere POD (cudampl/CPRRTC/include/Obstacle.h:4-10)
 *   boxes     [nb][18]  pos(3), R(9 row-major), min(3), max(3)
 *                                                       EnvConstraintCuboid (EnvConstraintCuboid.cu:36-79)
 *   cylinders [nc][14]  pos(3), R(9), radius, height    EnvConstraintCylinder (EnvConstraintCylinder.cu:50-100)
 *   capsules  [ncap][7] p1(3), p2(3), radius            pRRTC Capsule (pRRTC/src/planning/utils.cuh:113-127)
 * The scene is resident in HBM until destroyed. */
int mp_scene_create(const float* spheres, int ns, const float* boxes, int nb, const float* cylinders, int nc,
                    const float* capsules, int ncap, mp_scene** out);
/* Replace the obstacles (RobotSolver::updateEnvObstacle, RobotSolver.cu:179-204); the counts
 * must fit the capacities given at creation (MP_E_CAPACITY otherwise, where the reference
 * printed and returned). */
int mp_scene_update(mp_scene* scene, const float* spheres, int ns, const float* boxes, int nb,
                    const float* cylinders, int nc, const float* capsules, int ncap);
/* Joint bounds of every check, cost and plan on this scene.  In CUDAMPLib, joint bounds are a
 * constraint of the space like the obstacles: SingleArmSpace::checkStates evaluates a
 * BoundaryConstraint only when the space holds one (SingleArmSpace.cu:1572-1602; its cost,
 * BoundaryConstraint.cu:56-92: the squared violations of the active joints summed, valid below
 * 1e-4).  A scene is that constraint set, so it carries the bounds:
 *   MP_BOUNDS_ROBOT   the robot's joint limits (the default);
 *   MP_BOUNDS_NONE    no joint-bound check (a space without a BoundaryConstraint);
 *   MP_BOUNDS_CUSTOM  lower[n], upper[n] over the active joints (a BoundaryConstraint's own
 *                     bounds); joints from n on are unbounded, and +-INFINITY leaves a side open.
 * The cost row MP_COST_BOUNDS follows the same bounds.  MP_E_STATE while server queries use the
 * scene. */
#define MP_BOUNDS_ROBOT 0
#define MP_BOUNDS_NONE 1
#define MP_BOUNDS_CUSTOM 2
int mp_scene_set_joint_bounds(mp_scene* scene, int mode, const float* lower, const float* upper, int n);
int mp_scene_destroy(mp_scene* scene);

/* ---------------------------------------------------------------------------------------- */
/* Data-parallel primitives (device pointers, asynchronous on `stream`)                     */
/* ---------------------------------------------------------------------------------------- */
/* Sphere centres of n configurations.  q: [n][dof] active joints; centres: [n][S][3].
 * Replaces SingleArmStates::update (SingleArmStates.cu:696-723): FK kernel + sphere kernel
 * fused, no link poses written. */
int mp_fk_spheres(const mp_robot* robot, const float* d_q, int64_t n, float* d_centres, void* stream);

/* Link poses of n configurations: poses [n][L][16], the pose of every link in the base frame as
 * a row-major 4x4 with last row (0, 0, 0, 1); L is mp_robot_info's num_links, in the robot
 * description's link order.  Replaces kin_forward_nvrtc_kernel and
 * SingleArmStates::calculateForwardKinematicsNvrtv (KernelGenerator.h:222-406,
 * SingleArmStates.cu:752-768). */
int mp_fk_link_poses(const mp_robot* robot, const float* d_q, int64_t n, float* d_poses, void* stream);

/* Collision verdicts of n configurations: verdict[i] = 1 iff every constraint cost is 0
 * (self pairs, all obstacle types, joint bounds).  Early-out per wave.  d_centres optional.
 * Replaces states->update(); space->checkStates(states, feasible)
 * (SingleArmSpace.cu:1572-1602, one cost kernel per constraint + calculateTotalCosts). */
int mp_check_states(const mp_robot* robot, const mp_scene* scene, const float* d_q, int64_t n,
                    uint8_t* d_verdict, float* d_centres, void* stream);

/* Full-evaluation total cost per configuration (no early-out), the value
 * BaseStates::getTotalCostsHost returns after checkStates (States.cu:5-63). */
int mp_state_costs(const mp_robot* robot, const mp_scene* scene, const float* d_q, int64_t n,
                   float* d_cost, uint8_t* d_verdict, void* stream);

/* Per-constraint costs, full evaluation: parts [MP_COST_KINDS][n] (row k, state i at k*n + i),
 * rows in the order below; cost_opt [n] is their sum in that order (the value mp_state_costs
 * writes) and verdict_opt [n] is (sum == 0).  The [constraint][state] cost layout of BaseStates
 * (cudampl/CUDAMPLib/include/base/States.h:205), filled by each constraint's
 * computeCost (EnvConstraintSphere.cu:35-117, EnvConstraintCuboid.cu:84-171,
 * EnvConstraintCylinder.cu:107-196, SelfCollisionConstraint.cu:52-133,
 * BoundaryConstraint.cu:46-93); capsules have no CUDAMPLib constraint (pRRTC's
 * sphere_capsule, pRRTC/src/planning/utils.cuh:113-127, as a penetration depth). */
#define MP_COST_KINDS 6
#define MP_COST_ENV_SPHERE 0
#define MP_COST_ENV_CUBOID 1
#define MP_COST_ENV_CYLINDER 2
#define MP_COST_ENV_CAPSULE 3
#define MP_COST_SELF 4
#define MP_COST_BOUNDS 5
int mp_constraint_costs(const mp_robot* robot, const mp_scene* scene, const float* d_q, int64_t n, float* d_parts,
                        float* d_cost_opt, uint8_t* d_verdict_opt, void* stream);

/* Edge checks with checkMotions semantics (SingleArmSpace.cu:646-848): steps =
 * floor(d/res)+1, states q0 + dir*i (i < steps-1) then q1; ok[e] = all valid, len[e] = d.
 * No host round trip, no temporary allocation. */
int mp_check_edges(const mp_robot* robot, const mp_scene* scene, const float* d_q0, const float* d_q1,
                   int m, float resolution, uint8_t* d_ok, float* d_len, void* stream);

/* Nearest neighbour (squared L2 over dof, ties -> lowest index) of nq queries [nq][dof] in a
 * structure-of-arrays tree [dof][capacity] holding n_nodes nodes.  Replaces
 * find_the_nearest_neighbors_kernel (SingleArmStates.cu:1265-1431). */
int mp_nn_argmin(const float* d_tree_soa, int n_nodes, int capacity, int dof, const float* d_q, int nq,
                 int* d_idx, void* stream);

/* Uniform samples within the robot's active-joint bounds from a Philox4x32-10 stream:
 * q[i] = sample(seed, offset + i).  Replaces initCurand + sample_kernel
 * (SingleArmSpace.cu:184-217, 251-322). */
int mp_sample(const mp_robot* robot, uint64_t seed, uint64_t offset, int64_t n, float* d_q, void* stream);

/* Low-discrepancy samples: q[i] = Halton point offset + i + 1 in the prime bases 3, 5, 7, ...
 * (one per active joint), each coordinate shifted modulo 1 by a per-seed Philox rotation and
 * scaled into the bounds.  The sampler of pRRTC (pRRTC.cu:43-105: halton_initialize /
 * halton_next), made stateless so that any worker can draw any index of the stream. */
int mp_sample_halton(const mp_robot* robot, uint64_t seed, uint64_t offset, int64_t n, float* d_q, void* stream);

/* Steering: to[i] <- from[i] + (max/d)(to[i]-from[i]) when d > max (interpolate_kernel,
 * SingleArmSpace.cu:219-249). */
int mp_steer(const float* d_from, float* d_to, int n, int dof, float max_distance, void* stream);

/* ---------------------------------------------------------------------------------------- */
/* Planning                                                                                 */
/* ---------------------------------------------------------------------------------------- */
typedef struct {
    float resolution;      /* edge check resolution, rad (RobotSolver default 0.02) */
    int max_step;          /* configurations per extend/connect chunk (CPRRTC 32) */
    int max_iterations;    /* samples per query (CPRRTC max_iterations_ = 10000) */
    int tree_capacity;     /* nodes per tree (power of two recommended) */
    int waves;             /* persistent grid size in waves (0 = tuned default) */
    int max_path;          /* waypoint capacity per query */
    uint64_t seed;         /* sample stream seed; query k uses seed + k */
    int exclusive;         /* 1: one team grows each query (deterministic per seed, identical
                              alone and in a batch); 0: all teams share all queries */
    int team_waves;        /* waves per team (1, 2 or 4; 0 = tuned default): the waves of a
                              team share one iteration's nearest-neighbour scans and collision
                              checks (threads per tree iteration = 64 * team_waves) */
    int teams_per_query;   /* most teams growing one query's trees at once (0 = tuned, or
                              unbounded): more teams on one tree mostly add redundant
                              extensions; surplus teams leave the launch */
    int lanes_per_iteration; /* 64: one iteration per team pass, chunks of up to 64 states;
                              32: two iterations per pass, lanes 0-31 and 32-63 (the
                              reference's 32-thread CPRRTC block, needs max_step <= 32);
                              0 = tuned default */
    int sampler;           /* MP_SAMPLER_UNIFORM (Philox, as mp_sample; the CPRRTC sampler) or
                              MP_SAMPLER_HALTON (as mp_sample_halton; pRRTC's sampler) */
    int dynamic_domain;    /* 1: dynamic-domain sampling, pRRTC's (pRRTC.cu:420, 456-468, 584-598):
                              a sample farther from its nearest node than that node's radius is
                              dropped; a node's radius is unbounded until an extension from it
                              fails, then dd_radius, shrunk by (1 - dd_alpha) on each later
                              failure (not below dd_min_radius) and grown by (1 + dd_alpha) on
                              each success */
    float dd_radius;       /* rad (pRRTC 4.0) */
    float dd_alpha;        /* (pRRTC 1e-4) */
    float dd_min_radius;   /* rad (pRRTC 1.0) */
    int balance;           /* which tree an iteration grows (MP_BALANCE_*) */
    float tree_ratio;      /* the balance modes' size ratio (0: the mode's pRRTC default, 0.5 for
                              MP_BALANCE_DYNAMIC, 1.0 for MP_BALANCE_ALTERNATE) */
} mp_rrtc_params;

#define MP_SAMPLER_UNIFORM 0
#define MP_SAMPLER_HALTON 1

/* Tree balance (pRRTC's balance settings, pRRTC/src/planning/pRRTC.cu:348-371; n_s, n_g the
 * trees' sizes when an iteration starts, k its sample index, W the iterations running at once):
 *   MP_BALANCE_SMALLER    grow the smaller tree (CPRRTC, RobotSolver.cu:570-588; the default)
 *   MP_BALANCE_SPLIT      half the iterations grow each tree (pRRTC balance 0): k even -> start
 *   MP_BALANCE_DYNAMIC    pRRTC balance 1: while |n_s - n_g| < 1.5 W, a fraction 1 - n_s/(n_s+n_g)
 *                         of the iterations grows the start tree (chosen by a golden-ratio
 *                         sequence over k); otherwise the start tree when n_s/(n_s+n_g) <
 *                         tree_ratio, else the goal tree
 *   MP_BALANCE_ALTERNATE  pRRTC balance 2 (a worker switches trees unless they are unbalanced),
 *                         stateless: the smaller tree while the larger has at least (1 +
 *                         tree_ratio) times its nodes, else alternate with k
 * In paired iterations (lanes_per_iteration 32) both halves grow the tree chosen for the first. */
#define MP_BALANCE_SMALLER 0
#define MP_BALANCE_SPLIT 1
#define MP_BALANCE_DYNAMIC 2
#define MP_BALANCE_ALTERNATE 3

/* Defaults: resolution 0.02, max_step 32, max_iterations 10000, tree_capacity 65536,
 * waves 0, max_path 4096, seed 0, exclusive 0, team_waves 0, teams_per_query 0,
 * lanes_per_iteration 0, sampler MP_SAMPLER_UNIFORM, dynamic_domain 0 (dd_radius 4, dd_alpha 1e-4,
 * dd_min_radius 1). */
void mp_rrtc_default_params(mp_rrtc_params* p);

#define MP_PLAN_RUNNING 0
#define MP_PLAN_SOLVED 1
#define MP_PLAN_EXHAUSTED 2   /* max_iterations samples used (reference: returns {}) */
#define MP_PLAN_TREE_FULL 3
#define MP_PLAN_INVALID 4     /* start or goal in collision / out of bounds */
#define MP_PLAN_PATH_LOST 5   /* solved, but a tree node on the path never became readable
                                 within the path walker's wait bound: no path is returned */
#define MP_PLAN_PATH_TOO_LONG 6 /* solved, but the path has more waypoints than max_path or the
                                 caller's buffer holds: num_waypoints is the length needed and
                                 nothing is copied */

typedef struct {
    int status;            /* MP_PLAN_* */
    int num_waypoints;     /* waypoints written to `waypoints` (MP_PLAN_PATH_TOO_LONG: the
                              number the path needs; nothing is written) */
    float* waypoints;      /* caller buffer [capacity][dof], host memory */
    int capacity;
    int iterations;        /* samples consumed */
    int start_tree_size;
    int goal_tree_size;
    int64_t configs_checked;
    double latency_ms;     /* submit -> result on host */
    double work_flops;     /* algorithmic flops of the collision checks run for this query */
    double device_ms;      /* device time from admission (start/goal check) to the final status */
} mp_path;

/* One query, one launch: the CPRRTC::RobotSolver::solve equivalent (RobotSolver.cu:348-473):
 * all waves of the grid grow the two trees of this query; the path is start ... goal
 * (start-tree half root->connection, goal-tree half connection->root, RobotSolver.cu:267-346). */
int mp_rrtc_solve(mp_robot* robot, const mp_scene* scene, const float* start, const float* goal,
                  const mp_rrtc_params* params, mp_path* out);

/* mp_rrtc_solve without allocating: no counterpart in the reference.  For callers that cannot
 * allocate device memory (code under stream capture, tight serving loops, harnesses that forbid
 * allocation).  The query's trees, parents and domain radii, descriptor, state, control block and
 * path live in a static device array of MP_NOALLOC_SCRATCH_BYTES per device (held until the
 * process ends; its address is looked up at mp_init); the call allocates nothing and copies the
 * path into out->waypoints.  Same semantics and results as mp_rrtc_solve (exclusive mode: the
 * same path bit for bit).  MP_E_CAPACITY when 2 x tree_capacity x dof floats of trees, 4 x
 * tree_capacity ints of parents and radii, and max_path x dof floats of path (each rounded up to
 * 256 B, plus 1.25 KiB of descriptors) exceed the array.  Calls on one device share the array: the
 * library serialises them, and each clears what must start empty. */
#define MP_NOALLOC_SCRATCH_BYTES ((size_t)4 << 20)
int mp_rrtc_solve_noalloc(mp_robot* robot, const mp_scene* scene, const float* start, const float* goal,
                          const mp_rrtc_params* params, mp_path* out);

/* nq independent queries (robots may differ) in ONE persistent launch with per-query work
 * queues; robots[k], scenes[k], starts[k] / goals[k] ([dof_k] floats at offset k*16). */
int mp_rrtc_solve_batch(int nq, mp_robot* const* robots, const mp_scene* const* scenes, const float* starts16,
                        const float* goals16, const mp_rrtc_params* params, mp_path* out);

/* Autotuner of the planning launch shape (no counterpart in the reference; CPRRTC fixes one
 * block of threads per worker).  Solves the sample queries once per candidate shape and seed
 * (team_waves 1, 2, 4 x total waves x teams per query 0, 16, 32, 64, 128 x lanes per iteration
 * 64, 32 (32 when max_step <= 32); per_query = 1 tunes the one-query-per-launch shape by solving
 * the sample one query at a time, 0 tunes the batch shape with the sample as one batch), scores
 * each by its mean planning-kernel time, and returns the fastest in *best (params with
 * team_waves, waves, teams_per_query and lanes_per_iteration set).  The choice is kept for this
 * robot set and class, in memory and in the code-object cache directory, and used by every later
 * solve or server whose params leave those four fields at 0.  results (may be NULL) receives up
 * to cap candidates. */
typedef struct {
    int team_waves;
    int waves;             /* 0 = every team the device keeps resident */
    int teams_per_query;   /* 0 = unbounded */
    int lanes_per_iteration;
    double kernel_ms;      /* mean planning-kernel time over the seeds (summed over queries
                              when per_query) */
} mp_tune_result;
int mp_rrtc_autotune(int nq, mp_robot* const* robots, const mp_scene* const* scenes, const float* starts16,
                     const float* goals16, const mp_rrtc_params* params, int per_query, int reps,
                     mp_tune_result* results, int cap, int* n_results, mp_rrtc_params* best);

/* ---------------------------------------------------------------------------------------- */
/* Roadmap planner (RRG), device resident                                                   */
/* ---------------------------------------------------------------------------------------- */
/* The extend-connect loop of CUDAMPLib::RRG::solve (cudampl/CUDAMPLib/src/planners/RRG.cpp:
 * 241-424) with the roadmap in device memory.  Vertices 0..ns-1 are the start states (group 1),
 * ns..ns+ng-1 the goal states (group 2).  Each pass: sample k states from the space's Philox
 * stream (pass p takes indices sample_offset + p k .. + k - 1, as mp_sample), pull each to within
 * the side's travel distance of its nearest side-group member, check it, check the motions from
 * each free sample to its nearest member of each group, then add the samples that reach a group
 * (edges weighted by the motion length; a sample reaching both joins the groups and solves) or,
 * when none reaches the side's group, halve its travel distance (floor 0.1) and stay on the
 * side.  The side is group 1 on even loop indices.  No host round trip per pass: the caller
 * checks its termination condition and calls mp_roadmap_iterate once per loop pass; the graph is
 * read once at the end (Dijkstra stays with the caller, RRG.cpp:433-491).  The robot and scene
 * must outlive the roadmap. */
typedef struct mp_roadmap mp_roadmap;
typedef struct {
    int solved;            /* a vertex joins both groups */
    int passes;            /* loop passes run */
    int loop_index;        /* t: the next pass's side is group 1 when even */
    int vertices, edges, start_group, goal_group;
    float travel_start, travel_goal;  /* current travel distances of the two groups */
} mp_roadmap_info;
/* samples_per_iteration in [1, 256] (RRG::setSampleAttemptsInEachIteration); starts [ns][dof],
 * goals [ng][dof] on the host. */
int mp_roadmap_create(mp_robot* robot, const mp_scene* scene, const float* starts, int ns, const float* goals, int ng,
                      int samples_per_iteration, float max_travel_distance, float resolution, uint64_t seed,
                      uint64_t sample_offset, mp_roadmap** out);
/* Enqueues one loop pass (asynchronous; at most a few passes are in flight).  *solved = 1 once a
 * finished pass has joined the groups; no pass is enqueued then. */
int mp_roadmap_iterate(mp_roadmap* rm, int* solved);
/* Waits for the enqueued passes and reports the loop state. */
int mp_roadmap_info_get(mp_roadmap* rm, mp_roadmap_info* info);
/* Waits, then copies the graph: vertices [nv][dof], groups [nv] (1, 2, or 0 for the vertex that
 * joins both), edges [ne][2] (added vertex, the member it reaches), weights [ne], and the members
 * of each group in order of addition, start_members [start_group], goal_members [goal_group]
 * (vertex ids).  Any pointer may be NULL. */
int mp_roadmap_read(mp_roadmap* rm, float* vertices, int* groups, int* edges, float* weights, int* start_members,
                    int* goal_members);
int mp_roadmap_destroy(mp_roadmap* rm);

/* ---------------------------------------------------------------------------------------- */
/* Serving: continuous batching                                                             */
/* ---------------------------------------------------------------------------------------- */
typedef struct {
    int num_robots;
    mp_robot* const* robots;   /* robots the server can plan for (one module for the set) */
    size_t tree_memory_bytes;  /* device memory for tree buffers; admission is by free memory */
    int tree_capacity;         /* nodes per tree */
    int max_active;            /* upper bound on concurrently active queries */
    int step_iterations;       /* team passes per step launch at least (0 = 2 per team); a step
                                * gives every active query 32 passes when that is more */
    int waves;                 /* grid size in waves (0 = auto) */
    int use_graph;             /* capture the step in a hipGraph and replay it */
    mp_rrtc_params params;
} mp_server_params;

typedef struct {
    int64_t submitted, completed, solved;
    int active, queued;
    double p50_ms, p99_ms, mean_ms, throughput_qps;
    int64_t steps;
    double step_ms;  /* worker time inside steps: enqueue, device work, synchronisation */
    double loop_ms;  /* worker time in whole steps: admission, the step, retiring finished queries */
    int64_t backlog_steps;  /* steps that ran while queries waited for a slot */
} mp_server_stats;

/* A server thread owns the device; submit/poll are thread-safe. */
int mp_server_create(const mp_server_params* params, mp_server** out);
/* The scene is pinned from submit until the query finishes (or the server is destroyed):
 * mp_scene_update and mp_scene_destroy return MP_E_STATE for it meanwhile. */
int mp_server_submit(mp_server* s, int robot_index, const mp_scene* scene, const float* start, const float* goal,
                     uint64_t seed, int64_t* ticket);
/* Returns 1 and fills `out` when the query has finished, 0 if still pending.  out->waypoints
 * holds out->capacity waypoints of the ticket's robot (capacity x its dof floats); a solved path
 * longer than that is reported as MP_PLAN_PATH_TOO_LONG with its length and not copied.  Each
 * result is handed out once. */
int mp_server_poll(mp_server* s, int64_t ticket, mp_path* out);
int mp_server_wait_idle(mp_server* s, double timeout_s);
int mp_server_stats_get(mp_server* s, mp_server_stats* out);
int mp_server_destroy(mp_server* s);

#ifdef __cplusplus
}
#endif
#endif
