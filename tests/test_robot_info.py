"""The ROS-free C++ RobotInfo (cpp/include/foliation_planner/robot_info.hpp) reads Fetch's URDF,
SRDF and collision-sphere YAML (the reference's robot files, committed under tests/golden/robots)
into the same arrays as the bundled Fetch model: link order, joint types, parents, active map,
bounds, MoveIt default values, spheres, self-collision enable map, joint origins and axes.  The
bundled model's FK is pinned by closed forms in test_oracle.py, so this pins the loader too."""
import os
import subprocess

import numpy as np
import pytest

from isaac_sim_ros2_humble_amd.model import load_robot

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
FIX = os.path.join(HERE, "golden", "robots", "fetch")


@pytest.fixture(scope="module")
def loaded(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("ri") / "test_robot_info")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I", os.path.join(REPO, "isaac_sim_ros2_humble_amd", "cpp", "include"),
                    os.path.join(HERE, "cpp", "test_robot_info.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe, os.path.join(FIX, "fetch.urdf"), os.path.join(FIX, "fetch.srdf"), "arm_with_torso",
                          os.path.join(FIX, "fetch.collision_spheres.yaml")], capture_output=True, text=True, check=True)
    res = {}
    for line in out.stdout.splitlines():
        k, *v = line.split()
        res[k] = v
    return res


def test_fetch_arrays_match_the_bundled_model(loaded, fetch):
    r = loaded
    assert r["links"] == fetch.links
    assert [("" if j == "-" else j) for j in r["joint_names"]] == list(fetch.joint_names)
    assert int(r["dimension"][0]) == fetch.dof == 8
    for k, want in (("joint_types", fetch.joint_types), ("parents", fetch.parents), ("active", fetch.active),
                    ("sphere_link", fetch.sphere_link), ("collision_enabled", fetch.enabled.reshape(-1))):
        assert [int(x) for x in r[k]] == [int(x) for x in np.asarray(want).reshape(-1)], k
    for k, want in (("lower", fetch.lower), ("upper", fetch.upper), ("defaults", fetch.defaults),
                    ("sphere_radius", fetch.sphere_radius), ("sphere_pos", fetch.sphere_pos.reshape(-1))):
        assert np.array_equal(np.array([float(x) for x in r[k]], np.float32), np.asarray(want, np.float32)), k
    T = np.array([float(x) for x in r["joint_origins"]]).reshape(-1, 16)
    assert np.allclose(T, np.asarray(fetch.joint_origins).reshape(-1, 16), atol=1e-12)
    A = np.array([float(x) for x in r["joint_axes"]]).reshape(-1, 3)
    assert np.allclose(A, np.asarray(fetch.joint_axes), atol=1e-12)


def test_fetch_model_semantics(loaded):
    """MoveIt semantics the loader reproduces: continuous joints get [-pi, pi]; the prismatic
    torso keeps its URDF limits; 0 is the default inside the bounds; the root has type 0."""
    r = loaded
    names = ["" if j == "-" else j for j in r["joint_names"]]
    lo = np.array([float(x) for x in r["lower"]])
    hi = np.array([float(x) for x in r["upper"]])
    types = [int(x) for x in r["joint_types"]]
    assert types[0] == 0 and int(r["parents"][0]) == -1
    i = names.index("torso_lift_joint")
    assert types[i] == 2 and lo[i] == 0.0 and abs(hi[i] - 0.38615) < 1e-6
    for j in ("upperarm_roll_joint", "forearm_roll_joint", "wrist_roll_joint"):  # continuous in fetch.urdf
        k = names.index(j)
        assert types[k] == 1 and abs(lo[k] + np.pi) < 1e-6 and abs(hi[k] - np.pi) < 1e-6
    assert len(r["sphere_radius"]) == 110
