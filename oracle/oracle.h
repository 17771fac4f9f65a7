/*
 * CPU oracle for the motion-planning hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library,
 * and only as the checker / the timed CPU baseline.  The product path (the HIP library) never
 * links or calls it.
 *
 * Contents (each function cites the reference behaviour it restates; paths are relative to the
 * reference checkout):
 *   - orc_fk64        FP64 forward kinematics in the reference's 4x4 product order
 *                     (cudampl/CUDAMPLib/include/spaces/KernelGenerator.h:254-406,
 *                      cudampl/CPRRTC/src/RobotSolver.cu:845-1110,
 *                      cudampl/CUDAMPLib/src/states/SingleArmStates.cu:340-393).
 *   - orc_fk32        FP32 forward kinematics under the project's FP32 evaluation rules
 *                     (explicit fmaf chains, own sin/cos), the rules the HIP kernels follow.
 *   - orc_check32     FP32 collision verdicts/costs: self pairs, env spheres, oriented boxes,
 *                     finite cylinders, capsules, joint bounds; valid <=> total cost == 0
 *                     (SelfCollisionConstraint.cu:7-133, EnvConstraintSphere.cu:35-117,
 *                      EnvConstraintCuboid.cu:13-171, EnvConstraintCylinder.cu:17-196,
 *                      BoundaryConstraint.cu:46-93, SingleArmSpace.cu:1572-1602;
 *                      capsule: pRRTC/src/planning/utils.cuh:113-127).
 *   - orc_edges32     edge interpolation check (SingleArmSpace.cu:646-848, KernelGenerator.h:408-522).
 *   - orc_nn          brute-force nearest neighbour (SingleArmStates.cu:1265-1431), ties -> lowest index.
 *   - orc_steer       steering toward a target by at most a distance (SingleArmSpace.cu:219-249).
 *   - orc_validate_path  independent path re-check at a fine resolution.
 *   - orc_rrtconnect  CPU RRT-Connect with OMPL RRTConnect semantics (range, discrete motion
 *                     validation) over the FP32 checker: the CPU baseline.
 *   - orc_philox_uniform  Philox4x32-10 counter RNG (Salmon et al., SC'11) used for sampling.
 */
#ifndef MP_ORACLE_H
#define MP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_robot orc_robot;
typedef struct orc_scene orc_scene;

/* Robot model in RobotInfo layout (robot_info.hpp:13-220).  L links (one joint per link). */
orc_robot* orc_robot_create(int L, const int* joint_types, const double* joint_origins /*[L][16]*/,
                            const double* joint_axes /*[L][3]*/, const int* parents, const int* active,
                            const float* lower, const float* upper, const float* defaults,
                            int S, const int* sphere_link, const float* sphere_pos /*[S][3]*/,
                            const float* sphere_radius, const unsigned char* enabled /*[L][L]*/);
void orc_robot_destroy(orc_robot* r);
int orc_robot_dof(const orc_robot* r);
int orc_robot_num_pairs(const orc_robot* r);
/* pair list: out_i/out_j [P] (may be NULL), thresholds [P] = fl((ri+rj)^2) */
void orc_robot_pairs(const orc_robot* r, int* out_i, int* out_j, float* out_thr);
/* per-config full-evaluation flop count for a scene (cost model: FK + 18*S + pairs) */
double orc_full_flops(const orc_robot* r, const orc_scene* s);

/* Scene.  spheres [ns][4] = x,y,z,r; boxes [nb][18] = pos3, R9 row-major, min3, max3;
 * cylinders [nc][14] = pos3, R9, radius, height (axis = local z);
 * capsules [ncap][7] = p1(3), p2(3), radius. */
orc_scene* orc_scene_create(const float* spheres, int ns, const float* boxes, int nb,
                            const float* cylinders, int nc, const float* capsules, int ncap);
/* Joint bounds of the scene's checks (mp_scene_set_joint_bounds): mode 0 the robot's limits,
 * 1 none, 2 lower[n] / upper[n] over the active joints (joints from n on unbounded). */
void orc_scene_set_joint_bounds(orc_scene* s, int mode, const float* lower, const float* upper, int n);
void orc_scene_destroy(orc_scene* s);

void orc_sincosf(float x, float* s, float* c);

/* FK.  q: active joints [dof].  centres: [S][3].  link_poses (optional): [L][16] row-major. */
void orc_fk64(const orc_robot* r, const double* q, double* centres, double* link_poses);
void orc_fk32(const orc_robot* r, const float* q, float* centres);
/* FP32 link poses [L][16] row-major (last row 0 0 0 1), the frames orc_fk32 places spheres by. */
void orc_link_poses32(const orc_robot* r, const float* q, float* poses);

/* Check n configurations q[n][dof].  verdict[n] 1 = valid.  cost (optional) [n] total cost.
 * centres (optional) [n][S][3].  canon_flops (optional) [n]: flops of the canonical-order
 * early-out evaluation (FK + transforms + self pairs, then sphere-major obstacle pairs,
 * stopping at the first hit).  nthreads >= 1 partitions configs over host threads. */
void orc_check32(const orc_robot* r, const orc_scene* s, const float* q, int64_t n,
                 uint8_t* verdict, float* cost, float* centres, double* canon_flops, int nthreads);
/* Per-constraint costs parts[6][n] (row k, configuration i at k*n + i): 0 env spheres, 1 boxes,
 * 2 cylinders, 3 capsules, 4 self pairs, 5 joint bounds -- the [constraint][state] cost rows of
 * BaseStates (States.h:205); orc_check32's total cost is their sum in this order. */
void orc_constraint_costs(const orc_robot* r, const orc_scene* s, const float* q, int64_t n, float* parts);
/* Verdicts from given sphere centres [n][S][3] plus q for the bound check (collision stage only). */
void orc_check_centres32(const orc_robot* r, const orc_scene* s, const float* q, const float* centres,
                         int64_t n, uint8_t* verdict);

/* Edge check with CUDAMPLib checkMotions semantics.  q0/q1 [m][dof]. */
void orc_edges32(const orc_robot* r, const orc_scene* s, const float* q0, const float* q1, int m,
                 float resolution, uint8_t* ok, float* length);

/* Steering (SingleArmSpace.cu:219-249): in place on to[n][dof]. */
void orc_steer(const float* from, float* to, int n, int dof, float max_distance);

/* nearest neighbour: tree [n][dof] row-major, query [dof]; returns index (ties -> lowest). */
int orc_nn(const float* tree, int n, int dof, const float* q);

/* Path re-check: every segment densified at `resolution` (n = ceil(d/res) pieces, all
 * interior points and endpoints checked).  Returns 1 if valid; *bad_segment = first bad
 * segment index or -1 (waypoint 0 invalid gives -2). */
int orc_validate_path(const orc_robot* r, const orc_scene* s, const float* path, int k,
                      float resolution, int* bad_segment);

/* Philox4x32-10 uniform sample: config index idx of stream (seed): lo + u*(hi-lo) per active
 * joint, u = (x >> 8) * 2^-24 with x the Philox output word for dimension j. */
void orc_sample(const orc_robot* r, uint64_t seed, uint64_t idx, float* q);
/* Halton sample (pRRTC.cu:43-105, stateless): radical inverse of idx + 1 in the prime base of
 * each active joint (3, 5, 7, ...), shifted modulo 1 by a per-seed Philox rotation, scaled into
 * the bounds.  orc_radical_inverse(i, b) = (digits of i in base b, reversed) / b^digits. */
void orc_sample_halton(const orc_robot* r, uint64_t seed, uint64_t idx, float* q);
float orc_radical_inverse(uint32_t i, uint32_t b);
void orc_philox4x32(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                    uint32_t out[4]);

/* Roadmap planner: CUDAMPLib::RRG::solve's extend-connect loop (RRG.cpp:241-424) for up to
 * max_passes loop passes or until a vertex joins the start and goal groups.  Pass p samples
 * orc_sample(seed, offset + p k + i), i < k; each sample is pulled to within the side's travel
 * distance of its nearest side-group member (orc_steer) and checked; motions (orc_edges32) go from
 * each free sample to its nearest member of each group; the samples reaching a group are added in
 * sample order with their edges (weight = motion length), or, when none reaches the side's group,
 * its travel distance halves (floor 0.1) and the side repeats.  The side is group 1 on even loop
 * indices.  Vertices 0..ns-1 are the starts (group 1), then the goals (group 2).  Outputs:
 * vertices [vcap][dof], groups [vcap] (1, 2, or 0 joins both), edges [ecap][2] (added vertex,
 * reached member), weights [ecap], start_members / goal_members [vcap] (vertex ids in order of
 * addition).  Returns 0, or -2 when vcap / ecap is too small. */
typedef struct {
    int solved, passes, loop_index, vertices, edges, start_group, goal_group;
    float travel_start, travel_goal;
} orc_rrg_info;
int orc_rrg(const orc_robot* r, const orc_scene* s, const float* starts, int ns, const float* goals, int ng, int k,
            float max_travel, float resolution, uint64_t seed, uint64_t offset, int max_passes, float* vertices,
            int* groups, int vcap, int* edges, float* weights, int ecap, int* start_members, int* goal_members,
            orc_rrg_info* info);

typedef struct {
    int solved;
    int iterations;
    int start_tree;
    int goal_tree;
    int64_t states_checked;
    double seconds;
} orc_plan_stats;

/* OMPL RRTConnect semantics on the CPU.  range = max extension (rad), resolution = motion
 * validation spacing (rad).  Stops at max_iterations or time_limit seconds.  out_path has room
 * for out_cap waypoints [out_cap][dof]; returns number written (0 on failure). */
int orc_rrtconnect(const orc_robot* r, const orc_scene* s, const float* start, const float* goal,
                   float range, float resolution, uint64_t seed, int max_iterations, double time_limit,
                   float* out_path, int out_cap, orc_plan_stats* stats);

/* CPRRTC's planner as the GPU planner runs it with one team per query (exclusive mode), lanes 64
 * (one iteration per pass) or 32 (two per pass): sequential, deterministic per seed (orc_sample's
 * stream).  Returns the number of waypoints written (0 unless solved and the path fits out_cap). */
typedef struct {
    int status;      /* 1 solved, 2 exhausted, 3 tree full, 4 invalid start or goal */
    int iterations;  /* samples drawn */
    int start_tree, goal_tree, path_len;
} orc_cprrtc_info;
int orc_cprrtc(const orc_robot* r, const orc_scene* s, const float* start, const float* goal, float res,
               int max_step, uint64_t seed, int max_iters, int cap, int lanes, float* out_path, int out_cap,
               orc_cprrtc_info* info);

#ifdef __cplusplus
}
#endif
#endif
