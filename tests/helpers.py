"""Small synthetic robots, geometry references and the path contract check used by the tests."""
import copy

import numpy as np

from isaac_sim_ros2_humble_amd.model import RobotModel


def point_robot(radius=0.05, lim=2.0, base_sphere=None):
    """Three prismatic joints x, y, z carrying one sphere: configuration = sphere centre.
    base_sphere = (x, y, z, r) adds a sphere fixed on the base link (a self-collision partner)."""
    L = 4
    T = np.tile(np.eye(4).reshape(16), (L, 1))
    sl, sp, sr = [3], [[0.0, 0.0, 0.0]], [radius]
    if base_sphere is not None:
        sl.append(0)
        sp.append(list(base_sphere[:3]))
        sr.append(base_sphere[3])
    return RobotModel(
        name="point" if base_sphere is None else "point2", links=["base", "lx", "ly", "lz"],
        joint_types=np.array([0, 2, 2, 2], np.int32), joint_origins=T,
        joint_axes=np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float64),
        parents=np.array([-1, 0, 1, 2], np.int32), active=np.array([0, 1, 1, 1], np.int32),
        lower=np.array([0, -lim, -lim, -lim], np.float32), upper=np.array([0, lim, lim, lim], np.float32),
        defaults=np.zeros(4, np.float32), sphere_link=np.array(sl, np.int32),
        sphere_pos=np.array(sp, np.float32), sphere_radius=np.array(sr, np.float32),
        enabled=np.ones((L, L), np.uint8), joint_names=["", "jx", "jy", "jz"])


def point_box_distance(p, pos, R, lo, hi):
    """Distance (float64) from points p[n,3] to the oriented box {pos + R x : lo <= x <= hi}."""
    local = (p - pos) @ R  # R^T (p - pos)
    d = local - np.clip(local, lo, hi)
    return np.linalg.norm(d, axis=-1)


def point_cylinder_sdf(p, pos, R, radius, height):
    local = (p - pos) @ R
    dxy = np.hypot(local[:, 0], local[:, 1]) - radius
    dz = np.abs(local[:, 2]) - height / 2
    outside = np.hypot(np.maximum(dxy, 0), np.maximum(dz, 0))
    inside = np.minimum(np.maximum(dxy, dz), 0)
    return outside + inside


def point_segment_distance(p, a, b):
    v = b - a
    t = np.clip(((p - a) @ v) / (v @ v), 0, 1)
    return np.linalg.norm(p - (a + t[:, None] * v), axis=-1)


def planner_states(path, resolution):
    """The states the planner checked along a returned path: each segment a->b is a tree edge
    (states a + k * res * (b - a)/|b - a|, k = 0 .. |b - a|/res, endpoints included) or the join
    between the trees (shorter than 2 res; only its endpoints are checked, CPRRTC semantics)."""
    out = [path[:1].astype(np.float64)]
    for a, b in zip(path[:-1].astype(np.float64), path[1:].astype(np.float64)):
        # tree edges are whole multiples of res (up to float32 rounding, < 1e-5 relative); the
        # join is shorter than 2 res and gets its endpoint only
        n = max(int(np.floor(np.linalg.norm(b - a) / resolution + 1e-4)), 1)
        out.append(a + (b - a) * (np.arange(1, n + 1)[:, None] / n))
    return np.concatenate(out).astype(np.float32)


def assert_valid(model, scene, res, start, goal, resolution=0.02):
    """Status, endpoints and the planner's contract (every state it checked along the path is
    free); returns the re-validation at 0.005 rad, which a path may fail where it grazes an
    obstacle between two checked states (a rate, not a per-path guarantee)."""
    assert res["status"] == "solved", res
    path = res["path"]
    assert len(path) >= 2
    assert np.array_equal(path[0], np.asarray(start, np.float32))
    assert np.array_equal(path[-1], np.asarray(goal, np.float32))
    import oracle
    orc, osc = oracle.OracleRobot(model), oracle.OracleScene(scene)
    # planner_states rebuilds the checked states in float64 from the waypoints; each lies within
    # ~1e-6 m (sphere centres) of the float32 state the planner checked, so a state on a contact
    # boundary can flip: the contract is checked with spheres 1e-5 m smaller.  A wrong tree link
    # shows up as a run of states deep in collision, which this still catches.
    shrunk = copy.deepcopy(model)
    shrunk.sphere_radius = (np.asarray(model.sphere_radius, np.float64) - 1e-5).astype(np.float32)
    v, _, _, _ = oracle.check(oracle.OracleRobot(shrunk), osc, planner_states(path, resolution))
    assert v.all(), "path states in collision: %d of %d" % ((v == 0).sum(), len(v))
    ok, bad = oracle.validate_path(orc, osc, path, 0.005)
    return ok


def assert_valid_many(items, threads=16):
    """assert_valid over (model, scene, result, start, goal) items on a thread pool (the oracle
    releases the GIL); returns the number of paths valid at 0.005 rad."""
    import concurrent.futures as cf
    with cf.ThreadPoolExecutor(threads) as ex:
        return sum(ex.map(lambda it: assert_valid(*it), items))
