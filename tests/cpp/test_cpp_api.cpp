// Exercises the C++ planner API (CPRRTC::RobotSolver and the CUDAMPLib classes) the way the
// reference's harnesses do (src/foliation_planner/src/cuda_test.cpp TEST_CPRRTC / TEST_Planner).
// Input: a robot dump written by tests/test_gpu_cpp_api.py ("key v0 v1 ..." lines).  Output:
// "key value..." lines that the Python test checks against the CPU oracle.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

// the include block of src/foliation_planner/src/evaluate_cudampl.cpp:14-20, verbatim
#include <CUDAMPLib/spaces/SingleArmSpace.h>
#include <CUDAMPLib/constraints/EnvConstraintSphere.h>
#include <CUDAMPLib/constraints/SelfCollisionConstraint.h>
#include <CUDAMPLib/tasks/SingleArmTask.h>
#include <CUDAMPLib/planners/RRG.h>
#include <CUDAMPLib/termination/StepTermination.h>
#include <CUDAMPLib/termination/TimeoutTermination.h>
// and the rest of cuda_test.cpp:14-28 (its TaskSpaceConstraint.h is outside this library)
#include <CUDAMPLib/constraints/EnvConstraintCuboid.h>
#include <CUDAMPLib/constraints/EnvConstraintCylinder.h>
#include <CUDAMPLib/constraints/BoundaryConstraint.h>
#include <CUDAMPLib/planners/cRRTC.h>
#include <CPRRTC/RobotSolver.h>
#include "mp_cpp_util.h"

using V = std::vector<double>;

static std::map<std::string, V> load(const char* path) {
    std::map<std::string, V> m;
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream is(line);
        std::string k;
        is >> k;
        double x;
        while (is >> x) m[k].push_back(x);
    }
    return m;
}

static void print_path(const char* key, const std::vector<std::vector<float>>& p) {
    std::printf("%s %zu", key, p.size());
    for (auto& w : p)
        for (float x : w) std::printf(" %.9g", x);
    std::printf("\n");
}

// The RRG evaluation of src/foliation_planner/src/evaluate_cudampl.cpp:160-250 for each task
// "task<k>_obstacles" (x y z per sphere), "task<k>_radius", "task<k>_start", "task<k>_goal":
// self-collision plus a sphere-obstacle constraint named "obstacle_constraint" (replaced per
// task), a 0.02-resolution space, RRG with max travel distance 6.0 and 60 samples per iteration,
// a 10 s timeout, the path without densification.
static int evaluate(std::map<std::string, V>& m, const std::vector<int>& jt, const std::vector<mpx::Isometry3d>& poses,
                    const std::vector<mpx::Vector3d>& axes, const std::vector<int>& par, const std::vector<int>& smap,
                    const std::vector<std::vector<float>>& spos, const std::vector<float>& rad,
                    const std::vector<bool>& act, const std::vector<float>& lo, const std::vector<float>& hi,
                    const std::vector<float>& def, const std::vector<std::string>& names,
                    const std::vector<std::vector<bool>>& en, size_t dof) {
    std::vector<CUDAMPLib::BaseConstraintPtr> constraints;
    CUDAMPLib::SelfCollisionConstraintPtr self_collision_constraint =
        std::make_shared<CUDAMPLib::SelfCollisionConstraint>("self_collision_constraint", smap, rad, en);
    constraints.push_back(self_collision_constraint);
    for (int i = 0; m.count("task" + std::to_string(i) + "_start"); i++) {
        const std::string k = "task" + std::to_string(i) + "_";
        for (size_t j = 0; j < constraints.size(); j++)
            if (constraints[j]->getName() == "obstacle_constraint") {
                constraints.erase(constraints.begin() + j);
                break;
            }
        std::vector<std::vector<float>> obstacle_pos;
        const V& ob = m[k + "obstacles"];
        for (size_t o = 0; o + 2 < ob.size(); o += 3) obstacle_pos.push_back({(float)ob[o], (float)ob[o + 1], (float)ob[o + 2]});
        std::vector<float> radius(m[k + "radius"].begin(), m[k + "radius"].end());
        CUDAMPLib::EnvConstraintSpherePtr env_constraint =
            std::make_shared<CUDAMPLib::EnvConstraintSphere>("obstacle_constraint", obstacle_pos, radius);
        constraints.push_back(env_constraint);
        CUDAMPLib::SingleArmSpacePtr single_arm_space = std::make_shared<CUDAMPLib::SingleArmSpace>(
            dof, constraints, jt, poses, axes, par, smap, spos, rad, act, lo, hi, def, names, 0.02);
        std::vector<std::vector<float>> start_joint_values_set{std::vector<float>(m[k + "start"].begin(), m[k + "start"].end())};
        std::vector<std::vector<float>> goal_joint_values_set{std::vector<float>(m[k + "goal"].begin(), m[k + "goal"].end())};
        CUDAMPLib::SingleArmTaskPtr problem_task =
            std::make_shared<CUDAMPLib::SingleArmTask>(start_joint_values_set, goal_joint_values_set);
        CUDAMPLib::RRGPtr planner = std::make_shared<CUDAMPLib::RRG>(single_arm_space);
        planner->setMaxTravelDistance(6.0);
        planner->setSampleAttemptsInEachIteration(60);
        planner->setMotionTask(problem_task, false);
        auto start_time = std::chrono::high_resolution_clock::now();
        CUDAMPLib::TimeoutTerminationPtr termination_condition = std::make_shared<CUDAMPLib::TimeoutTermination>(10.0);
        planner->solve(termination_condition);
        auto end_time = std::chrono::high_resolution_clock::now();
        auto time_taken = std::chrono::duration_cast<std::chrono::milliseconds>(end_time - start_time).count();
        std::printf("%ssolved %d %ld\n", k.c_str(), (int)problem_task->hasSolution(), (long)time_taken);
        print_path((k + "path").c_str(), problem_task->getSolution());
    }
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    auto m = load(argv[1]);
    const bool eval_mode = argc > 2 && std::string(argv[2]) == "evaluate";
    const int L = (int)m["joint_types"].size(), S = (int)m["sphere_radius"].size();
    std::vector<int> jt, par, smap;
    std::vector<mpx::Isometry3d> poses(L);
    std::vector<mpx::Vector3d> axes(L), spos3(S);
    std::vector<std::vector<float>> spos(S);
    std::vector<float> rad, lo, hi, def;
    std::vector<bool> act;
    std::vector<std::string> names;
    for (int i = 0; i < L; i++) names.push_back("l" + std::to_string(i));
    std::vector<std::vector<bool>> en(L, std::vector<bool>(L));
    for (int i = 0; i < L; i++) {
        jt.push_back((int)m["joint_types"][i]);
        par.push_back((int)m["parents"][i]);
        act.push_back(m["active"][i] != 0);
        lo.push_back((float)m["lower"][i]);
        hi.push_back((float)m["upper"][i]);
        def.push_back((float)m["defaults"][i]);
        for (int r = 0; r < 4; r++)
            for (int c = 0; c < 4; c++) poses[i].matrix()(r, c) = m["joint_origins"][16 * i + 4 * r + c];
        for (int a = 0; a < 3; a++) axes[i](a) = m["joint_axes"][3 * i + a];
        for (int j = 0; j < L; j++) en[i][j] = m["enabled"][i * L + j] != 0;
    }
    for (int s = 0; s < S; s++) {
        smap.push_back((int)m["sphere_link"][s]);
        rad.push_back((float)m["sphere_radius"][s]);
        for (int a = 0; a < 3; a++) {
            spos3[s](a) = m["sphere_pos"][3 * s + a];
            spos[s].push_back((float)m["sphere_pos"][3 * s + a]);
        }
    }
    size_t dof = 0;
    for (bool b : act) dof += b;
    if (eval_mode) return evaluate(m, jt, poses, axes, par, smap, spos, rad, act, lo, hi, def, names, en, dof);
    std::vector<float> start(m["start"].begin(), m["start"].end()), goal(m["goal"].begin(), m["goal"].end());
    const V& bx = m["box"];  // x_min x_max y_min y_max z_min z_max x y z roll pitch yaw

    // --- CPRRTC::RobotSolver (cuda_test.cpp:1902-1931)
    CPRRTC::RobotSolver solver("robot", dof, jt, poses, axes, par, smap, spos3, rad, act, lo, hi, def, names, en, 0.02f);
    solver.setEnvObstacleCache(4, 2, 1);
    std::vector<CPRRTC::Sphere> sph;
    std::vector<CPRRTC::Cuboid> cub{{(float)bx[0], (float)bx[1], (float)bx[2], (float)bx[3], (float)bx[4], (float)bx[5],
                                     (float)bx[6], (float)bx[7], (float)bx[8], (float)bx[9], (float)bx[10], (float)bx[11]}};
    std::vector<CPRRTC::Cylinder> cyl;
    solver.updateEnvObstacle(sph, cub, cyl);
    auto path = solver.solve(start, goal);
    print_path("cprrtc_path", path);
    std::vector<float> shortv(dof - 1, 0.0f);
    std::printf("cprrtc_bad_size %zu\n", solver.solve(shortv, goal).size());
    std::vector<CPRRTC::Cuboid> many(3, cub[0]);
    solver.updateEnvObstacle(sph, many, cyl);  // over capacity: message, obstacles unchanged
    std::printf("cprrtc_after_overflow %d\n", solver.solve(start, goal).empty() ? 0 : 1);

    // --- CUDAMPLib (cuda_test.cpp TEST_Planner / evaluate_cudampl.cpp)
    float Rm[9];
    mpcpp::rpy_to_matrix((float)bx[9], (float)bx[10], (float)bx[11], Rm);
    auto env = std::make_shared<CUDAMPLib::EnvConstraintCuboid>(
        "obstacle", std::vector<std::vector<float>>{{(float)bx[6], (float)bx[7], (float)bx[8]}},
        std::vector<std::vector<float>>{std::vector<float>(Rm, Rm + 9)},
        std::vector<std::vector<float>>{{(float)bx[1], (float)bx[3], (float)bx[5]}},
        std::vector<std::vector<float>>{{(float)bx[0], (float)bx[2], (float)bx[4]}});
    auto self = std::make_shared<CUDAMPLib::SelfCollisionConstraint>("self", smap, rad, en);
    auto bound = std::make_shared<CUDAMPLib::BoundaryConstraint>("bound", lo, hi, act);
    std::vector<CUDAMPLib::BaseConstraintPtr> cons{env, self, bound};
    auto space = std::make_shared<CUDAMPLib::SingleArmSpace>(dof, cons, jt, poses, axes, par, smap, spos, rad, act, lo, hi,
                                                             def, names, 0.02f);
    auto states = space->createStatesFromVector({start, goal});
    std::vector<bool> feas;
    space->checkStates(states, feas);
    std::printf("space_feasible %d %d\n", (int)feas[0], (int)feas[1]);
    // per-constraint costs (env, self, bound) of the start, a state inside the wall and one past
    // the last joint's limit by 0.5
    {
        std::vector<float> wall = start, beyond = start;
        for (auto& v : wall) v = 0.0f;
        beyond[dof - 1] = 1.5f;
        auto cst = space->createStatesFromVector({start, wall, beyond});
        space->checkStates(cst);
        auto costs = cst->getCostsHost();
        auto total = cst->getTotalCostsHost();
        std::printf("space_costs");
        for (size_t i = 0; i < costs.size(); i++) {
            for (float c : costs[i]) std::printf(" %.9g", c);
            std::printf(" %.9g", total[i]);
        }
        std::printf("\n");
    }
    // joint bounds are checked only through a BoundaryConstraint, with its own bounds: a state
    // beyond the last active joint's limit by 0.5 is feasible in a space without one; a narrow
    // BoundaryConstraint (that joint within +-0.1) rejects a state at 0.3 and keeps the start
    {
        std::vector<float> outside = start;
        outside[dof - 1] = 1.5f;
        std::vector<CUDAMPLib::BaseConstraintPtr> cons_nb{env, self};
        auto space_nb = std::make_shared<CUDAMPLib::SingleArmSpace>(dof, cons_nb, jt, poses, axes, par, smap, spos, rad,
                                                                    act, lo, hi, def, names, 0.02f);
        std::vector<bool> f_nb, f_b, f_n;
        space_nb->checkStates(space_nb->createStatesFromVector({outside}), f_nb);
        space->checkStates(space->createStatesFromVector({outside}), f_b);
        std::vector<float> nlo = lo, nhi = hi;
        for (int i = 0, a = 0; i < L; i++)
            if (act[i] && a++ == (int)dof - 1) {
                nlo[i] = -0.1f;
                nhi[i] = 0.1f;
            }
        auto narrow = std::make_shared<CUDAMPLib::BoundaryConstraint>("bound", nlo, nhi, act);
        std::vector<CUDAMPLib::BaseConstraintPtr> cons_n{env, self, narrow};
        auto space_n = std::make_shared<CUDAMPLib::SingleArmSpace>(dof, cons_n, jt, poses, axes, par, smap, spos, rad,
                                                                   act, lo, hi, def, names, 0.02f);
        std::vector<float> z03 = start;
        z03[dof - 1] = 0.3f;
        space_n->checkStates(space_n->createStatesFromVector({start, z03}), f_n);
        std::printf("space_bounds %d %d %d %d\n", (int)f_nb[0], (int)f_b[0], (int)f_n[0], (int)f_n[1]);
    }
    std::vector<bool> mf;
    std::vector<float> mc;
    space->checkMotions(space->createStatesFromVector({start}), space->createStatesFromVector({goal}), mf, mc);
    std::printf("space_motion %d %.9g\n", (int)mf[0], mc[0]);
    // forward kinematics through the states (update / link-pose and sphere-centre getters)
    auto fk_states = std::dynamic_pointer_cast<CUDAMPLib::SingleArmStates>(space->createStatesFromVector({start, goal}));
    auto last_link = fk_states->getLinkPoseInBaseLinkHost("l" + std::to_string(L - 1));
    auto centres = fk_states->getSelfCollisionSpheresPosInBaseLinkHost();
    std::printf("fk_pose");
    for (auto& T : last_link)
        for (int r = 0; r < 4; r++)
            for (int c = 0; c < 4; c++) std::printf(" %.9g", T.matrix()(r, c));
    std::printf("\nfk_centres %.9g %.9g %.9g %.9g %.9g %.9g\n", centres[0][0][0], centres[0][0][1], centres[0][0][2],
                centres[1][0][0], centres[1][0][1], centres[1][0][2]);
    const std::string gen = space->generateFKKernelSourceCode();
    std::printf("space_source %d %d\n", (int)(gen.find("mp_plan") != std::string::npos),
                (int)(gen == space->generateRobotCollisionModelSourceCode()));
    auto samples = space->sample(64);
    std::vector<bool> sf;
    space->checkStates(samples, sf);
    int nf = 0;
    for (bool b : sf) nf += b;
    std::printf("space_samples %d %d\n", samples->getNumOfStates(), nf);

    auto task = std::make_shared<CUDAMPLib::SingleArmTask>(std::vector<std::vector<float>>{start},
                                                           std::vector<std::vector<float>>{goal});
    auto planner = std::make_shared<CUDAMPLib::RRG>(space);
    planner->setMotionTask(task, false);
    planner->solve(std::make_shared<CUDAMPLib::TimeoutTermination>(10.0f));
    std::printf("rrg_solved %d %s\n", (int)task->hasSolution(), task->getFailureReason().c_str());
    print_path("rrg_path", task->getSolution());
    CUDAMPLib::BaseStatesPtr gs, gg;
    planner->getStartAndGoalGroupStates(gs, gg);
    std::printf("rrg_groups %d %d\n", gs->getNumOfStates(), gg->getNumOfStates());

    // a goal region of two states, the first inside the wall: the path ends at the free one
    std::vector<float> in_wall(dof, 0.0f);
    auto task_multi = std::make_shared<CUDAMPLib::SingleArmTask>(std::vector<std::vector<float>>{start},
                                                                 std::vector<std::vector<float>>{in_wall, goal});
    auto planner_multi = std::make_shared<CUDAMPLib::RRG>(space);
    planner_multi->setSampleAttemptsInEachIteration(60);
    planner_multi->setMotionTask(task_multi, false);
    planner_multi->solve(std::make_shared<CUDAMPLib::TimeoutTermination>(10.0f));
    std::printf("rrg_multi_solved %d\n", (int)task_multi->hasSolution());
    print_path("rrg_multi_path", task_multi->getSolution());

    // state manager (SingleArmStateManager): add, nearest per group, gather, steer toward a node
    auto mgr = space->createStateManager();
    std::vector<std::vector<float>> stored{{0.1f, 0.2f, 0.3f}, {-0.4f, 0.5f, 0.0f}, {0.9f, -0.9f, 0.1f},
                                           {0.0f, 0.0f, 0.8f}, {-0.7f, -0.2f, -0.6f}};
    auto i1 = mgr->add_states(space->createStatesFromVector({stored[0], stored[1], stored[2]}));
    auto i2 = mgr->add_states(space->createStatesFromVector({stored[3], stored[4]}));
    std::printf("mgr_add %d %d %d %d %d %d\n", i1[0], i1[1], i1[2], i2[0], i2[1], mgr->get_num_of_states());
    std::vector<std::vector<float>> queries{{0.0f, 0.3f, 0.2f}, {0.8f, -0.5f, 0.5f}, {-0.6f, 0.1f, -0.3f}};
    std::vector<std::vector<int>> nn;
    mgr->find_the_nearest_neighbors(space->createStatesFromVector(queries), {{0, 1, 2}, {3, 4}}, nn);
    std::printf("mgr_nn");
    for (auto& r : nn) std::printf(" %d %d", r[0], r[1]);
    std::printf("\n");
    std::vector<std::vector<int>> knn;
    mgr->find_k_nearest_neighbors(2, space->createStatesFromVector(queries), {{0, 1, 2, 3, 4}}, knn);
    std::printf("mgr_knn");
    for (auto& r : knn) std::printf(" %d %d", r[0], r[1]);
    std::printf("\n");
    auto got = mgr->get_states({4, 0});
    print_path("mgr_get", space->getJointVectorInActiveJointsFromStates(got));
    auto cat = mgr->concatinate_states({got, space->createStatesFromVector(queries)});
    std::printf("mgr_cat %d\n", cat->getNumOfStates());
    auto far = space->createStatesFromVector({{0.9f, 0.9f, 0.9f}});
    std::vector<int> to0{0};
    mgr->interpolateToStates(far, to0, 0.1f);
    print_path("mgr_steer", space->getJointVectorInActiveJointsFromStates(far));
    std::printf("rrg_manager %d\n", planner->getStateManager()->get_num_of_states());

    // a goal region: a space whose bounds box the goal (SingleArmTask's region constructor)
    std::vector<float> rlo = lo, rhi = hi;
    for (int i = 0, a = 0; i < L; i++)
        if (act[i]) {
            rlo[i] = goal[a] - 0.05f;
            rhi[i] = goal[a] + 0.05f;
            a++;
        }
    auto rbound = std::make_shared<CUDAMPLib::BoundaryConstraint>("bound", rlo, rhi, act);
    std::vector<CUDAMPLib::BaseConstraintPtr> rcons{env, self, rbound};
    auto region = std::make_shared<CUDAMPLib::SingleArmSpace>(dof, rcons, jt, poses, axes, par, smap, spos, rad, act,
                                                              rlo, rhi, def, names, 0.02f);
    auto task_region = std::make_shared<CUDAMPLib::SingleArmTask>(std::vector<std::vector<float>>{start}, region, 20);
    auto planner_region = std::make_shared<CUDAMPLib::RRG>(space);
    planner_region->setMotionTask(task_region, false);
    planner_region->solve(std::make_shared<CUDAMPLib::TimeoutTermination>(10.0f));
    std::printf("rrg_region %d %d\n", (int)task_region->getGoalStatesVector().size(), (int)task_region->hasSolution());
    print_path("rrg_region_path", task_region->getSolution());

    // a free direct motion is returned at once, densified to the space's resolution
    std::vector<float> near_goal = start;
    near_goal[1] += 0.3f;
    auto task_direct = std::make_shared<CUDAMPLib::SingleArmTask>(std::vector<std::vector<float>>{start},
                                                                  std::vector<std::vector<float>>{near_goal});
    auto planner_direct = std::make_shared<CUDAMPLib::RRG>(space);
    planner_direct->setMotionTask(task_direct, false);
    planner_direct->solve(std::make_shared<CUDAMPLib::StepTermination>(0));
    print_path("rrg_direct_path", task_direct->getSolution());

    // a start inside the wall is rejected; a blocked query with no steps left terminates
    auto task_bad = std::make_shared<CUDAMPLib::SingleArmTask>(std::vector<std::vector<float>>{in_wall},
                                                               std::vector<std::vector<float>>{goal});
    auto planner_bad = std::make_shared<CUDAMPLib::RRG>(space);
    planner_bad->setMotionTask(task_bad, false);
    planner_bad->solve(std::make_shared<CUDAMPLib::StepTermination>(100));
    std::printf("rrg_invalid %d %s\n", (int)task_bad->hasSolution(), task_bad->getFailureReason().c_str());
    auto task_s0 = std::make_shared<CUDAMPLib::SingleArmTask>(std::vector<std::vector<float>>{start},
                                                              std::vector<std::vector<float>>{goal});
    auto planner_s0 = std::make_shared<CUDAMPLib::RRG>(space);
    planner_s0->setMotionTask(task_s0, false);
    planner_s0->solve(std::make_shared<CUDAMPLib::StepTermination>(0));
    std::printf("rrg_step0 %d %s\n", (int)task_s0->hasSolution(), task_s0->getFailureReason().c_str());

    auto task0 = std::make_shared<CUDAMPLib::SingleArmTask>(std::vector<std::vector<float>>{start},
                                                            std::vector<std::vector<float>>{goal});
    auto planner0 = std::make_shared<CUDAMPLib::cRRTC>(space);
    planner0->setMotionTask(task0, true);
    planner0->solve(std::make_shared<CUDAMPLib::StepTermination>(0));
    std::printf("step0 %d %s\n", (int)task0->hasSolution(), task0->getFailureReason().c_str());

    // cRRTC to a solution (cuda_test.cpp TEST_Planner with the cRRTC planner): the full path is
    // densified to the space's resolution; without it, the planner's waypoints
    for (int full = 1; full >= 0; full--) {
        auto task_c = std::make_shared<CUDAMPLib::SingleArmTask>(std::vector<std::vector<float>>{start},
                                                                 std::vector<std::vector<float>>{goal});
        auto planner_c = std::make_shared<CUDAMPLib::cRRTC>(space);
        planner_c->setMotionTask(task_c, full == 1);
        planner_c->solve(std::make_shared<CUDAMPLib::TimeoutTermination>(10.0f));
        std::printf("crrtc_solved_%d %d\n", full, (int)task_c->hasSolution());
        print_path(full ? "crrtc_full_path" : "crrtc_path", task_c->getSolution());
    }
    return 0;
}
