"""MotionBenchMaker problem loader: scene / request YAML -> Scene, start, goal.

Restates the reference's MBM loaders (src/foliation_planner/include/foliation_planner/
obstacle_generator.hpp:914-975 for scenes, MPM_helper.hpp:10-92 for requests):
  * scene: `world.collision_objects[*]`, each with `primitives` (shape_msgs/SolidPrimitive:
    BOX dimensions = full sizes x y z, CYLINDER dimensions = [height, radius], SPHERE = [radius])
    and `primitive_poses` (position, quaternion x y z w), in the robot base frame; an object-level
    `pose`, when present, is composed in front of the primitive poses;
  * request: start values from `start_state.joint_state` (name, position), goal values from the
    first `goal_constraints[0].joint_constraints`; joints missing from a request keep 0.
Inactive joints of the robot take their values from the request's start state, as the
reference's MBM harness does (cuda_test.cpp:2806-2837 with robot_info.hpp:127-155).

Panda problems name joints panda_jointN / panda_finger_jointN.  They are planned for the FER
model (robots/fer.json: franka_description/robots/fer, the Panda arm), whose kinematics.yaml
equals FR3's but whose joint limits are the Panda's; the FER model keeps FR3's fr3_* joint names,
onto which the panda_* names map.  Under FR3's narrower limits many Panda goals lie out of
bounds.
"""
import copy

import numpy as np
import yaml

from .scene import Scene, box_record, cylinder_record


def quat_matrix(q):
    x, y, z, w = [float(v) for v in q]
    n = np.sqrt(x * x + y * y + z * z + w * w)
    x, y, z, w = x / n, y / n, z / n, w / n
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _pose(p):
    return np.asarray(p.get("position", [0, 0, 0]), np.float64), quat_matrix(p.get("orientation", [0, 0, 0, 1]))


def load_scene(path):
    with open(path) as f:
        doc = yaml.safe_load(f)
    boxes, cyls, sph = [], [], []
    for obj in (doc.get("world") or {}).get("collision_objects") or []:
        op, oR = (np.zeros(3), np.eye(3))
        if isinstance(obj.get("pose"), dict):
            op, oR = _pose(obj["pose"])
        for prim, pp in zip(obj.get("primitives") or [], obj.get("primitive_poses") or []):
            p, R = _pose(pp)
            pos, rot = op + oR @ p, oR @ R
            dims = [float(d) for d in prim["dimensions"]]
            t = prim["type"]
            if t == "box":
                boxes.append(box_record(pos, rot, np.asarray(dims[:3]) / 2))
            elif t == "cylinder":
                cyls.append(cylinder_record(pos, rot, dims[1], dims[0]))
            elif t == "sphere":
                sph.append(np.concatenate([pos, [dims[0]]]).astype(np.float32))
    return Scene(spheres=np.array(sph).reshape(-1, 4), boxes=np.array(boxes).reshape(-1, 18),
                 cylinders=np.array(cyls).reshape(-1, 14))


def _canon(name, model):
    if name in model.joint_names:
        return name
    if name.startswith("panda_"):
        alt = "fr3_" + name[len("panda_"):]
        if alt in model.joint_names:
            return alt
    return None


def load_request(path, model):
    """(start[dof], goal[dof], inactive values {joint name: value}) for `model`."""
    with open(path) as f:
        doc = yaml.safe_load(f)
    js = doc["start_state"]["joint_state"]
    start_map = {}
    for n, v in zip(js["name"], js["position"]):
        c = _canon(n, model)
        if c is not None:
            start_map[c] = float(v)
    goal_map = {}
    for jc in doc["goal_constraints"][0]["joint_constraints"]:
        c = _canon(jc["joint_name"], model)
        if c is not None:
            goal_map[c] = float(jc["position"])
    active = [n for n, a in zip(model.joint_names, model.active) if a]
    start = np.array([start_map.get(n, 0.0) for n in active], np.float32)
    goal = np.array([goal_map.get(n, 0.0) for n in active], np.float32)
    inactive = {n: v for n, v in start_map.items() if n not in active}
    return start, goal, inactive


def with_defaults(model, values):
    """Copy of the model whose inactive joints take `values` (joint name -> value)."""
    m = copy.deepcopy(model)
    for i, n in enumerate(m.joint_names):
        if n in values and not m.active[i]:
            m.defaults[i] = np.float32(values[n])
    return m


# Inactive-joint values of every MBM request start state (uniform across the 14 problem sets).
MBM_INACTIVE = {
    "fetch": {"bellows_joint": 0.05, "head_pan_joint": 0.0, "head_tilt_joint": 0.0, "l_gripper_finger_joint": 0.05,
              "l_wheel_joint": 0.0, "r_gripper_finger_joint": 0.05, "r_wheel_joint": 0.0},
    "fer": {"fr3_finger_joint1": 0.065, "fr3_finger_joint2": 0.065},
}

# Robot model of each problem family: sets named *_panda use FER, *_fetch use Fetch.
MBM_ROBOTS = ("fer", "fetch")


def mbm_model(model):
    """The robot as configured by the MBM requests (inactive joints at the start-state values)."""
    m = with_defaults(model, MBM_INACTIVE[model.name])
    m.name = model.name + "_mbm"
    return m


def problem_models():
    """[FER, Fetch] as configured by the MBM requests: the models `load_problems` indexes."""
    from .model import load_robot
    return [mbm_model(load_robot(n)) for n in MBM_ROBOTS]


def load_problems(root, models):
    """All (set name, model index, Scene, start, goal) under `root`/<set>/scene*.yaml;
    `models` = [Panda-family model, Fetch model], normally `problem_models()`."""
    import glob
    import os
    out = []
    for d in sorted(os.listdir(root)):
        if not os.path.isdir(os.path.join(root, d)):
            continue
        k = 1 if d.endswith("fetch") else 0
        for sf in sorted(glob.glob(os.path.join(root, d, "scene*.yaml"))):
            start, goal, _ = load_request(sf.replace("scene", "request"), models[k])
            out.append((d, k, load_scene(sf), start, goal))
    return out
