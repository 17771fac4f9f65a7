"""Robot-model converter: URDF / xacro kinematics / SRDF / collision-sphere YAML -> flat JSON.

This is the ROS-free stand-in for ``RobotInfo`` (src/foliation_planner/include/foliation_planner/
robot_info.hpp:13-220): it produces the same flat arrays that RobotInfo hands to
``SingleArmSpace`` / ``CPRRTC::RobotSolver`` (joint types, joint origin transforms, joint axes,
parent-link map, active-joint map, bounds, default values, collision spheres and the link-pair
self-collision enable map), without MoveIt.

Semantics reproduced from robot_info.hpp and the MoveIt/urdfdom behaviour it relies on:
  * link order = depth-first pre-order from the URDF root, children visited in the order
    urdfdom stores them (its joint table is a std::map keyed by joint name, so children are
    sorted by joint name);
  * one joint per link; the root link gets joint type 0, identity origin, zero axis
    (robot_info.hpp:167-175);
  * joint types use MoveIt's enum: 1 revolute (URDF revolute and continuous), 2 prismatic,
    5 fixed (robot_info.hpp:118-164);
  * continuous joints get MoveIt's bounds [-pi, pi];
  * default value of a revolute/prismatic joint = 0 when 0 lies inside its bounds, else the
    bounds' midpoint (MoveIt JointModel::getVariableDefaultPositions); fixed joints 0;
  * joint origin = translation + rotation built the way urdfdom/MoveIt do it
    (rpy -> quaternion -> Eigen rotation matrix), stored as a row-major 4x4 (Isometry3d::matrix());
  * collision spheres in YAML document order, per link (robot_info.hpp:233-291);
  * enable map: all true, SRDF ``disable_collisions`` pairs cleared symmetrically
    (robot_info.hpp:56-93).

The JSON files written next to this script are data derived from the reference's robot files;
this script is the only place that reads the reference checkout.

Usage: python convert_robots.py [--reference /root/reference]
"""
import argparse
import json
import math
import os
import re
import xml.etree.ElementTree as ET

import yaml

HERE = os.path.dirname(os.path.abspath(__file__))

JT_ROOT, JT_REVOLUTE, JT_PRISMATIC, JT_FIXED = 0, 1, 2, 5


def rpy_to_matrix(roll, pitch, yaw):
    """urdf::Rotation::setFromRPY -> quaternion (normalised) -> Eigen::Quaterniond::toRotationMatrix."""
    phi, the, psi = roll / 2.0, pitch / 2.0, yaw / 2.0
    x = math.sin(phi) * math.cos(the) * math.cos(psi) - math.cos(phi) * math.sin(the) * math.sin(psi)
    y = math.cos(phi) * math.sin(the) * math.cos(psi) + math.sin(phi) * math.cos(the) * math.sin(psi)
    z = math.cos(phi) * math.cos(the) * math.sin(psi) - math.sin(phi) * math.sin(the) * math.cos(psi)
    w = math.cos(phi) * math.cos(the) * math.cos(psi) + math.sin(phi) * math.sin(the) * math.sin(psi)
    n = math.sqrt(x * x + y * y + z * z + w * w)
    x, y, z, w = x / n, y / n, z / n, w / n
    tx, ty, tz = 2 * x, 2 * y, 2 * z
    twx, twy, twz = tx * w, ty * w, tz * w
    txx, txy, txz = tx * x, ty * x, tz * x
    tyy, tyz, tzz = ty * y, tz * y, tz * z
    return [[1 - (tyy + tzz), txy - twz, txz + twy],
            [txy + twz, 1 - (txx + tzz), tyz - twx],
            [txz - twy, tyz + twx, 1 - (txx + tyy)]]


def pose_matrix(xyz, rpy):
    R = rpy_to_matrix(*rpy)
    return [R[0][0], R[0][1], R[0][2], xyz[0],
            R[1][0], R[1][1], R[1][2], xyz[1],
            R[2][0], R[2][1], R[2][2], xyz[2],
            0.0, 0.0, 0.0, 1.0]


def moveit_default(lo, hi):
    return 0.0 if (lo <= 0.0 <= hi) else (lo + hi) / 2.0


class Joint:
    def __init__(self, name, jtype, parent, child, xyz, rpy, axis, lower, upper):
        self.name, self.jtype, self.parent, self.child = name, jtype, parent, child
        self.xyz, self.rpy, self.axis = xyz, rpy, axis
        self.lower, self.upper = lower, upper


def build_model(name, group, links, joints, active_joint_names, sphere_yaml, disabled_pairs, root=None):
    """Assemble the RobotInfo arrays from a link list + joint list (URDF semantics)."""
    by_child = {j.child: j for j in joints}
    children = {l: [] for l in links}
    for j in sorted(joints, key=lambda j: j.name):  # urdfdom: std::map<joint name, joint>
        children[j.parent].append(j.child)
    roots = [l for l in links if l not in by_child]
    assert len(roots) == 1, roots
    root = roots[0]
    order = []

    def dfs(l):
        order.append(l)
        for c in children[l]:
            dfs(c)
    dfs(root)
    assert len(order) == len(links)
    index = {l: i for i, l in enumerate(order)}

    jtypes, jposes, jaxes, parents, jnames, active, lower, upper, defaults = [], [], [], [], [], [], [], [], []
    for l in order:
        if l == root:
            jtypes.append(JT_ROOT)
            jposes.append(pose_matrix([0, 0, 0], [0, 0, 0]))
            jaxes.append([0.0, 0.0, 0.0])
            parents.append(-1)
            jnames.append("")
            active.append(False)
            lower.append(0.0)
            upper.append(0.0)
            defaults.append(0.0)
            continue
        j = by_child[l]
        jnames.append(j.name)
        parents.append(index[j.parent])
        jposes.append(pose_matrix(j.xyz, j.rpy))
        if j.jtype in ("revolute", "continuous"):
            jtypes.append(JT_REVOLUTE)
            jaxes.append([float(v) for v in j.axis])
            lo, hi = (-math.pi, math.pi) if j.jtype == "continuous" else (j.lower, j.upper)
            defaults.append(moveit_default(lo, hi))
        elif j.jtype == "prismatic":
            jtypes.append(JT_PRISMATIC)
            jaxes.append([float(v) for v in j.axis])
            lo, hi = j.lower, j.upper
            defaults.append(moveit_default(lo, hi))
        elif j.jtype == "fixed":
            jtypes.append(JT_FIXED)
            jaxes.append([0.0, 0.0, 0.0])
            lo, hi = 0.0, 0.0
            defaults.append(0.0)
        else:
            raise ValueError(j.jtype)
        is_active = j.name in active_joint_names
        active.append(is_active)
        lower.append(float(lo) if is_active else 0.0)
        upper.append(float(hi) if is_active else 0.0)

    # group order of active joints must match link order (RobotInfo maps bounds by link order)
    sph_link, sph_pos, sph_r = [], [], []
    for lname, spheres in sphere_yaml.items():
        if lname not in index:
            continue
        for s in spheres:
            sph_link.append(index[lname])
            sph_pos.append([float(v) for v in s["center"]])
            sph_r.append(float(s["radius"]))

    L = len(order)
    enabled = [[1] * L for _ in range(L)]
    for a, b in disabled_pairs:
        if a in index and b in index:
            enabled[index[a]][index[b]] = 0
            enabled[index[b]][index[a]] = 0

    npairs = 0
    for i in range(len(sph_r)):
        for j in range(i + 1, len(sph_r)):
            if sph_link[i] != sph_link[j] and enabled[sph_link[i]][sph_link[j]]:
                npairs += 1

    return {
        "name": name,
        "group": group,
        "links": order,
        "joint_names": jnames,
        "joint_types": jtypes,
        "joint_origins": jposes,
        "joint_axes": jaxes,
        "parents": parents,
        "active": active,
        "lower": lower,
        "upper": upper,
        "defaults": defaults,
        "sphere_link": sph_link,
        "sphere_pos": sph_pos,
        "sphere_radius": sph_r,
        "collision_enabled": enabled,
        "dof": sum(1 for a in active if a),
        "num_self_pairs": npairs,
    }


def parse_fetch(ref):
    base = os.path.join(ref, "src/fetch_ros")
    root = ET.parse(os.path.join(base, "fetch_description/robots/fetch.urdf")).getroot()
    links = [l.get("name") for l in root.findall("link")]
    joints = []
    for j in root.findall("joint"):
        o = j.find("origin")
        xyz = [float(v) for v in o.get("xyz", "0 0 0").split()] if o is not None else [0.0, 0.0, 0.0]
        rpy = [float(v) for v in o.get("rpy", "0 0 0").split()] if o is not None else [0.0, 0.0, 0.0]
        a = j.find("axis")
        axis = [float(v) for v in a.get("xyz").split()] if a is not None else [1.0, 0.0, 0.0]
        lim = j.find("limit")
        lo = float(lim.get("lower", 0.0)) if lim is not None and lim.get("lower") is not None else 0.0
        hi = float(lim.get("upper", 0.0)) if lim is not None and lim.get("upper") is not None else 0.0
        joints.append(Joint(j.get("name"), j.get("type"), j.find("parent").get("link"),
                            j.find("child").get("link"), xyz, rpy, axis, lo, hi))
    srdf = ET.parse(os.path.join(base, "fetch_moveit2_config/config/fetch.srdf")).getroot()
    group = "arm_with_torso"
    g = [x for x in srdf.findall("group") if x.get("name") == group][0]
    active = [x.get("name") for x in g.findall("joint")]
    disabled = [(d.get("link1"), d.get("link2")) for d in srdf.findall("disable_collisions")]
    with open(os.path.join(base, "fetch_description/robots/fetch.collision_spheres.yaml")) as f:
        sy = yaml.safe_load(f)["collision_spheres"]
    return build_model("fetch", group, links, joints, active, sy, disabled)


def parse_fr3(ref, arm="fr3"):
    """FR3, or with arm="fer" the FER (Panda) arm: franka_description/robots/fer has the same
    kinematics.yaml as robots/fr3 but its own joint_limits.yaml.  The FER model keeps the FR3
    link and joint names, sphere model and SRDF pairs (the Panda problems of MotionBenchMaker
    name joints panda_jointN, mapped onto these by the problem loader); only the bounds differ."""
    base = os.path.join(ref, "src/franka_ros2")
    desc = os.path.join(base, "franka_description")
    with open(os.path.join(desc, "robots/%s/kinematics.yaml" % arm)) as f:
        kin = yaml.safe_load(f)
    with open(os.path.join(desc, "robots/%s/joint_limits.yaml" % arm)) as f:
        lim = yaml.safe_load(f)
    p = "fr3_"
    links = ["base"] + [p + "link%d" % i for i in range(9)] + [p + "hand", p + "hand_tcp", p + "leftfinger", p + "rightfinger"]

    def kin_origin(jn):
        k = kin[jn]["kinematic"]
        return [float(k["x"]), float(k["y"]), float(k["z"])], [float(k["roll"]), float(k["pitch"]), float(k["yaw"])]

    joints = [Joint(p + "base_joint", "fixed", "base", p + "link0", [0.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0, 0, 0], 0, 0)]
    # franka_arm.xacro:44-151: joint1..7 revolute about z, origins from kinematics.yaml, limits from joint_limits.yaml
    for i in range(1, 8):
        xyz, rpy = kin_origin("joint%d" % i)
        l = lim["joint%d" % i]["limit"]
        joints.append(Joint(p + "joint%d" % i, "revolute", p + "link%d" % (i - 1), p + "link%d" % i, xyz, rpy,
                            [0.0, 0.0, 1.0], float(l["lower"]), float(l["upper"])))
    xyz, rpy = kin_origin("joint8")
    joints.append(Joint(p + "joint8", "fixed", p + "link7", p + "link8", xyz, rpy, [0, 0, 0], 0, 0))
    # end_effectors/common/franka_hand.xacro with franka_hand_arguments.xacro defaults
    joints.append(Joint(p + "hand_joint", "fixed", p + "link8", p + "hand", [0.0, 0.0, 0.0], [0.0, 0.0, -math.pi / 4], [0, 0, 0], 0, 0))
    joints.append(Joint(p + "hand_tcp_joint", "fixed", p + "hand", p + "hand_tcp", [0.0, 0.0, 0.1034], [0.0, 0.0, 0.0], [0, 0, 0], 0, 0))
    joints.append(Joint(p + "finger_joint1", "prismatic", p + "hand", p + "leftfinger", [0.0, 0.0, 0.0584], [0.0, 0.0, 0.0], [0.0, 1.0, 0.0], 0.0, 0.04))
    joints.append(Joint(p + "finger_joint2", "prismatic", p + "hand", p + "rightfinger", [0.0, 0.0, 0.0584], [0.0, 0.0, math.pi], [0.0, 1.0, 0.0], 0.0, 0.04))
    active = [p + "joint%d" % i for i in range(1, 8)]
    disabled = []
    srdf_dir = os.path.join(base, "franka_fr3_moveit_config/srdf")
    for fn in ("fr3_arm.xacro", "hand.xacro", "fr3_arm_hand.xacro"):
        with open(os.path.join(srdf_dir, fn)) as f:
            txt = f.read().replace("${arm_id}", "fr3")
        for m in re.finditer(r'<disable_collisions\s+link1="([^"]+)"\s+link2="([^"]+)"', txt):
            disabled.append((m.group(1), m.group(2)))
    with open(os.path.join(desc, "robots/franka.yml")) as f:
        sy = yaml.safe_load(f)["collision_spheres"]
    return build_model(arm, "fr3_arm", links, joints, active, sy, disabled)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    ap.add_argument("--out", default=HERE)
    a = ap.parse_args()
    for fn, model in (("fr3.json", parse_fr3(a.reference)), ("fer.json", parse_fr3(a.reference, "fer")),
                      ("fetch.json", parse_fetch(a.reference))):
        with open(os.path.join(a.out, fn), "w") as f:
            json.dump(model, f, indent=1)
        print(fn, "links", len(model["links"]), "dof", model["dof"], "spheres", len(model["sphere_radius"]),
              "self pairs", model["num_self_pairs"])


if __name__ == "__main__":
    main()
