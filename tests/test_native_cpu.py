"""CPU-side checks of the native library: it loads, exports every entry point the C ABI header
declares, generates robot source without a device, and reports errors without a GPU."""
import ctypes as C
import os
import re
import subprocess

import numpy as np

from isaac_sim_ros2_humble_amd import native
from isaac_sim_ros2_humble_amd.model import load_robot

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "mp_abi.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mp_[a-z0-9_]+)\s*\(", txt)))


def test_header_functions_exported():
    lib = native.lib()
    names = declared_functions()
    assert len(names) >= 35
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(native.EXPORTED) >= set(names)


def test_generated_source_structure(fr3, fetch):
    src = native.generate_source([fr3, fetch])
    assert "struct Robot0" in src and "struct Robot1" in src
    for k in ("mp_check_0", "mp_check_1", "mp_plan", "mp_plan_init", "mp_edges_1", "mp_nn_generic"):
        assert k in src
    # every self pair of the robot appears exactly once in the hit code of its struct
    r0 = src[src.index("struct Robot0"):src.index("struct Robot1")]
    hit = r0[r0.index("self_hit"):r0.index("self_cost")]
    assert hit.count("h |=") == 754
    r1 = src[src.index("struct Robot1"):src.index("mp_load_cfg") if "mp_load_cfg" in src else None]
    hit1 = r1[r1.index("self_hit"):r1.index("self_cost")]
    assert hit1.count("h |=") == 2534
    # FK evaluates one sin/cos per active joint only, in the sphere frames and the link poses
    fk = r0[r0.index("void fk("):r0.index("NLINK")]
    assert fk.count("mp_sincosf(") == 7
    poses = r0[r0.index("void poses("):r0.index("write_centres")]
    assert poses.count("mp_sincosf(") == 7 and "static constexpr int NLINK = 14;" in r0
    assert poses.count("] = 1.0f;") >= 14  # the (0, 0, 0, 1) row of every link
    assert "mp_poses_0" in src and "mp_poses_1" in src


def test_query_descriptor_loads_every_field_the_kernels_read(fr3):
    # the planning kernels read these QueryDev fields after mp_load_query (admission empties a
    # reused slot over dof_max coordinate rows)
    src = native.generate_source([fr3])
    body = src[src.index("mp_load_query("):]
    body = body[:body.index("\n}\n")]
    for f in ("tree[0]", "tree[1]", "parent[0]", "parent[1]", "path", "seed", "robot", "dof", "cap", "max_iters",
              "max_path", "dof_max", "sampler", "scene.box", "scene.sph", "scene.cyl", "scene.cap", "scene.bnd", "scene.bbox", "scene.bcyl",
              "scene.bcap"):
        assert "Q.%s = mp_sld(&p->%s)" % (f, f) in body, f


def test_generated_checks_split_across_a_team(fr3):
    """Every self-collision block is guarded by its team share and every obstacle loop strides by
    the team size, so a team of W waves covers each block and obstacle exactly once."""
    src = native.generate_source([fr3])
    r0 = src[src.index("struct Robot0"):]
    hit = r0[r0.index("self_hit"):r0.index("self_cost")]
    blocks = hit.count("T.mine(")
    assert blocks == len(set(int(x.split(")")[0]) for x in hit.split("T.mine(")[1:]))
    assert [int(x.split(")")[0]) for x in hit.split("T.mine(")[1:]] == list(range(blocks))
    env = r0[r0.index("env_hit"):r0.index("env_cost")]
    assert env.count("o += os, ++it)") == 4 and "const int o0 = mp_uniform(T.w), os = mp_uniform(T.W);" in env


def test_generated_source_is_deterministic_and_robot_specific(fr3):
    a = native.generate_source([fr3])
    assert a == native.generate_source([fr3])
    m2 = load_robot("fr3")
    m2.sphere_radius = m2.sphere_radius.copy()
    m2.sphere_radius[5] += np.float32(0.001)
    assert native.generate_source([m2]) != a  # a different robot never reuses a cached module


def test_invalid_robot_rejected():
    m = load_robot("fr3")
    m.parents = m.parents.copy()
    m.parents[3] = 7  # not topologically ordered
    try:
        native.generate_source([m])
    except native.MPError as e:
        assert "topological" in str(e)
    else:
        raise AssertionError("expected an error")


def test_ctypes_structs_match_the_header(tmp_path):
    """Every ctypes mirror of a C ABI struct has the header's size and field offsets (compiled
    with the host C compiler from include/mp_abi.h)."""
    pairs = [(native.RobotDesc, "mp_robot_desc"), (native.RRTCParams, "mp_rrtc_params"), (native.Path, "mp_path"),
             (native.TuneResult, "mp_tune_result"), (native.ServerParams, "mp_server_params"),
             (native.ServerStats, "mp_server_stats"), (native.RoadmapInfo, "mp_roadmap_info")]
    body = []
    for py, c in pairs:
        body.append('printf("%%zu", sizeof(%s));' % c)
        body += ['printf(" %%zu", offsetof(%s, %s));' % (c, f) for f, _ in py._fields_]
        body.append('printf("\\n");')
    src = tmp_path / "layout.c"
    src.write_text("#include <stdio.h>\n#include <stddef.h>\n#include \"mp_abi.h\"\nint main(void) {\n" +
                   "\n".join(body) + "\nreturn 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    lines = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    for (py, c), line in zip(pairs, lines):
        got = [int(x) for x in line.split()]
        assert got[0] == C.sizeof(py), c
        assert got[1:] == [getattr(py, f).offset for f, _ in py._fields_], c


def _gen_error(d):
    arr = (C.POINTER(native.RobotDesc) * 1)(C.pointer(d))
    n = native.lib().mp_generate_source(C.cast(arr, C.c_void_p), 1, None, 0)
    assert n == 0
    return native.lib().mp_last_error().decode()


def test_malformed_robot_descriptions_rejected():
    """Bad counts and null arrays are MP_E_INVALID errors, not crashes (no device needed)."""
    m = load_robot("fr3")
    for field, value, msg in (("num_links", 0, "link/sphere counts"), ("num_links", -3, "link/sphere counts"),
                              ("num_spheres", -1, "link/sphere counts"), ("joint_types", None, "null"),
                              ("parents", None, "null"), ("collision_enabled", None, "null"),
                              ("sphere_radius", None, "null")):
        d, keep = native.robot_desc(m)
        setattr(d, field, value)
        assert msg in _gen_error(d), field
    # no spheres at all is a valid (collision-free) robot
    d, keep = native.robot_desc(m)
    d.num_spheres = 0
    d.sphere_link = d.sphere_pos = d.sphere_radius = None
    arr = (C.POINTER(native.RobotDesc) * 1)(C.pointer(d))
    assert native.lib().mp_generate_source(C.cast(arr, C.c_void_p), 1, None, 0) > 0


def test_plan_statuses_named():
    # every MP_PLAN_* code of the header has a name in the Python layer
    txt = open(HEADER).read()
    codes = {int(v): k for k, v in re.findall(r"#define (MP_PLAN_[A-Z_]+) (\d+)\b", txt) if k != "MP_PLAN_PROFILE_N"}
    assert set(codes) == set(native.PLAN_STATUS)
    for v, k in codes.items():
        assert native.PLAN_STATUS[v] == k[len("MP_PLAN_"):].lower()
