"""Seeded synthetic workloads for the planning hot path.

  c1  FR3, 20 yaw-only boxes (half extents U[0.02,0.2]) centred in a 0.3-0.9 m shell around the
      base, collision-free start/goal, one RRT-Connect query.
  c2  2^20 uniform configurations of FR3; 64 obstacles in a 1.5 m cube: 40 oriented boxes,
      12 cylinders/capsules (radius U[0.02,0.1], length U[0.1,0.6]), 12 spheres (r U[0.03,0.15]).
  c3  Fetch arm+torso, a 3-shelf unit of 0.02 m boards (gap U[0.25,0.4], depth 0.35) plus 30
      random boxes; start outside the shelf, goal with the gripper between two boards.
  c4  256 independent queries, FR3 / Fetch mixed, obstacle count log-uniform in [5,200],
      half the start/goal pairs joined by a collision-free straight line, half blocked.
  c5  Poisson arrivals at rate lambda; scenes reused with probability 0.8.

Random numbers come from numpy's PCG64 (default_rng) seeded per workload.  Start/goal
rejection sampling takes a `check(model, scene, q) -> uint8 verdicts` callable, so the same
generator runs against the GPU checker or the CPU oracle.
"""
import math

import numpy as np

from .scene import Scene, box_record, capsule_record, cylinder_record, random_rotation, rpy_matrix


BASE_KEEPOUT = 0.15  # m: obstacles stay clear of a vertical column around the robot base

# obstacle region and base keep-out per robot: the keep-out covers the spheres of links that no
# active revolute joint moves (FR3 link0: 0.14 m; Fetch base/torso/head: 0.36 m) plus 2 cm
WORKSPACE = {
    "fr3": ((-1.1, -1.1, 0.0), (1.1, 1.1, 1.5), 0.16),
    "fetch": ((-1.5, -1.5, 0.0), (1.5, 1.5, 1.9), 0.38),
}


def workspace(model):
    return WORKSPACE.get(model.name, ((-1.1, -1.1, 0.0), (1.1, 1.1, 1.5), BASE_KEEPOUT))


def clear_of_base(centre, extent, keepout=BASE_KEEPOUT):
    """True if an obstacle of circumscribed radius `extent` at `centre` stays outside the column."""
    return math.hypot(centre[0], centre[1]) - extent > keepout


def uniform_configs(model, n, rng):
    lo, hi = model.active_lower.astype(np.float64), model.active_upper.astype(np.float64)
    return (lo + rng.random((n, model.dof)) * (hi - lo)).astype(np.float32)


# ------------------------------------------------------------------------------------------- c1
def c1_scene(seed=0, n_boxes=20):
    rng = np.random.default_rng(seed)
    boxes = []
    for _ in range(n_boxes):
        while True:
            half = rng.uniform(0.02, 0.2, 3)
            d = rng.normal(size=3)
            d[2] = abs(d[2])
            d /= np.linalg.norm(d)
            c = d * rng.uniform(0.3, 0.9)
            if clear_of_base(c, np.linalg.norm(half)):
                break
        yaw = rng.uniform(0.0, np.pi)
        boxes.append(box_record(c, rpy_matrix(0.0, 0.0, yaw), half))
    return Scene(boxes=np.array(boxes))


# ------------------------------------------------------------------------------------------- c2
def c2_scene(seed=1, n_box=40, n_cyl=6, n_cap=6, n_sph=12, lo=(-0.75, -0.75, 0.0), hi=(0.75, 0.75, 1.5)):
    rng = np.random.default_rng(seed)
    lo, hi = np.asarray(lo), np.asarray(hi)

    def pos(extent):
        while True:
            c = lo + rng.random(3) * (hi - lo)
            if clear_of_base(c, extent):
                return c
    boxes = []
    for _ in range(n_box):
        half = rng.uniform(0.02, 0.2, 3)
        boxes.append(box_record(pos(np.linalg.norm(half)), random_rotation(rng), half))
    cyls, caps = [], []
    for _ in range(n_cyl):
        rad, ln = rng.uniform(0.02, 0.1), rng.uniform(0.1, 0.6)
        cyls.append(cylinder_record(pos(math.hypot(rad, ln / 2)), random_rotation(rng), rad, ln))
    for _ in range(n_cap):
        rad, ln = rng.uniform(0.02, 0.1), rng.uniform(0.1, 0.6)
        caps.append(capsule_record(pos(rad + ln / 2), random_rotation(rng), rad, ln))
    sph = []
    for _ in range(n_sph):
        rad = rng.uniform(0.03, 0.15)
        sph.append(np.concatenate([pos(rad), [rad]]).astype(np.float32))
    return Scene(spheres=np.array(sph).reshape(-1, 4), boxes=np.array(boxes).reshape(-1, 18),
                 cylinders=np.array(cyls).reshape(-1, 14), capsules=np.array(caps).reshape(-1, 7))


def c2_configs(model, n=1 << 20, seed=1):
    return uniform_configs(model, n, np.random.default_rng(seed + 1000))


# ------------------------------------------------------------------------------------------- c3
def c3_scene(seed=0, n_random=30):
    rng = np.random.default_rng(seed)
    I = np.eye(3)
    boxes = []
    x0 = rng.uniform(0.75, 0.9)          # front face of the shelf unit
    depth, width, t = 0.35, 0.8, 0.02
    y0 = rng.uniform(-0.15, 0.15)
    z = rng.uniform(0.35, 0.5)           # lowest board
    gaps = rng.uniform(0.25, 0.4, 3)
    levels = [z]
    for g in gaps:
        levels.append(levels[-1] + g)
    xc = x0 + depth / 2
    for lz in levels:                    # 4 boards -> 3 shelves
        boxes.append(box_record([xc, y0, lz], I, [depth / 2, width / 2, t / 2]))
    top = levels[-1]
    hz = top / 2
    for side in (-1, 1):                 # side walls
        boxes.append(box_record([xc, y0 + side * (width / 2 + t / 2), hz], I, [depth / 2, t / 2, hz]))
    boxes.append(box_record([x0 + depth + t / 2, y0, hz], I, [t / 2, width / 2 + t, hz]))  # back
    shelf = {"x0": x0, "depth": depth, "y0": y0, "width": width, "levels": levels}
    for _ in range(n_random):            # clutter away from the robot base and the shelf
        while True:
            c = np.array([rng.uniform(-1.5, 1.5), rng.uniform(-1.5, 1.5), rng.uniform(0.0, 1.8)])
            if np.hypot(c[0], c[1]) < 0.7:
                continue
            if x0 - 0.3 < c[0] < x0 + depth + 0.3 and abs(c[1] - y0) < width / 2 + 0.3:
                continue
            break
        boxes.append(box_record(c, rpy_matrix(0, 0, rng.uniform(0, np.pi)), rng.uniform(0.02, 0.2, 3)))
    return Scene(boxes=np.array(boxes)), shelf


def gripper_centres(model, centres, link="gripper_link"):
    idx = model.links.index(link)
    sel = model.sphere_link == idx
    return centres[:, sel, :].mean(axis=1)


def c3_queries(model, scene, shelf, n, check, fk, seed=0, batch=1 << 14, max_rounds=200):
    """n (start, goal) pairs: start with the gripper outside the shelf, goal with the gripper
    inside a compartment; both collision-free.  `fk(model, q) -> centres[n,S,3]`."""
    rng = np.random.default_rng(seed + 77)
    starts, goals = [], []
    x0, depth, y0, width, lv = shelf["x0"], shelf["depth"], shelf["y0"], shelf["width"], shelf["levels"]
    for _ in range(max_rounds):
        q = uniform_configs(model, batch, rng)
        ok = check(model, scene, q).astype(bool)
        g = gripper_centres(model, fk(model, q))
        inside = (g[:, 0] > x0 + 0.05) & (g[:, 0] < x0 + depth - 0.02) & (np.abs(g[:, 1] - y0) < width / 2 - 0.05)
        between = np.zeros(len(q), bool)
        for a, b in zip(lv[:-1], lv[1:]):
            between |= (g[:, 2] > a + 0.05) & (g[:, 2] < b - 0.05)
        inside &= between
        outside = g[:, 0] < x0 - 0.1
        goals += list(q[ok & inside])
        starts += list(q[ok & outside])
        if len(goals) >= n and len(starts) >= n:
            break
    if len(goals) < n or len(starts) < n:
        raise RuntimeError("c3: could not sample enough start/goal pairs (%d/%d)" % (len(starts), len(goals)))
    return np.array(starts[:n]), np.array(goals[:n])


# --------------------------------------------------------------------------------------- c4 / c5
def random_scene(rng, n_obstacles, lo=(-1.1, -1.1, 0.0), hi=(1.1, 1.1, 1.5), keepout=BASE_KEEPOUT):
    """Mixed boxes / cylinders / spheres (60/20/20) with the c2 size ranges, clear of a vertical
    keep-out column of radius `keepout` around the robot base."""
    lo, hi = np.asarray(lo), np.asarray(hi)
    boxes, cyls, sph = [], [], []
    for _ in range(n_obstacles):
        u = rng.random()
        if u < 0.6:
            size = rng.uniform(0.02, 0.2, 3)
            extent = np.linalg.norm(size)
        elif u < 0.8:
            size = (rng.uniform(0.02, 0.1), rng.uniform(0.1, 0.6))
            extent = math.hypot(size[0], size[1] / 2)
        else:
            size = rng.uniform(0.03, 0.15)
            extent = size
        while True:
            c = lo + rng.random(3) * (hi - lo)
            if clear_of_base(c, extent, keepout):
                break
        if u < 0.6:
            boxes.append(box_record(c, random_rotation(rng), size))
        elif u < 0.8:
            cyls.append(cylinder_record(c, random_rotation(rng), size[0], size[1]))
        else:
            sph.append(np.concatenate([c, [size]]).astype(np.float32))
    return Scene(spheres=np.array(sph).reshape(-1, 4), boxes=np.array(boxes).reshape(-1, 18),
                 cylinders=np.array(cyls).reshape(-1, 14))


def log_uniform_count(rng, lo=5, hi=200):
    return int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))


def sample_pair(model, scene, rng, check, check_edge, want_free_line, batch=4096, max_rounds=64):
    """Collision-free (start, goal); want_free_line selects pairs whose straight segment is
    collision-free (True) or blocked (False).  Returns None if not found."""
    for _ in range(max_rounds):
        q = uniform_configs(model, batch, rng)
        ok = q[check(model, scene, q).astype(bool)]
        if len(ok) < 2:
            continue
        m = len(ok) // 2
        a, b = ok[:m], ok[m:2 * m]
        free = check_edge(model, scene, a, b).astype(bool)
        sel = np.nonzero(free == want_free_line)[0]
        if len(sel):
            return a[sel[0]], b[sel[0]]
    return None


def c4_batch(models, n, check, check_edge, seed=0, batch=4096):
    """n queries: (model_index, scene, start, goal, easy).  Queries whose scene admits no
    collision-free pair of the requested kind are redrawn with a new scene."""
    rng = np.random.default_rng(seed + 4)
    out = []
    while len(out) < n:
        k = int(rng.integers(len(models)))
        lo, hi, keep = workspace(models[k])
        scene = random_scene(rng, log_uniform_count(rng), lo, hi, keep)
        easy = (len(out) % 2) == 0
        pair = sample_pair(models[k], scene, rng, check, check_edge, easy, batch=batch)
        if pair is None:
            continue
        out.append((k, scene, pair[0], pair[1], easy))
    return out


def poisson_arrivals(rate, duration, seed=0):
    rng = np.random.default_rng(seed + 5)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return np.array(out)
        out.append(t)


def sample_pairs(model, scene, rng, check, check_edge, n, batch=4096, max_rounds=16):
    """Up to n collision-free (start, goal) pairs in one scene from batched checks, alternating
    pairs whose straight segment is collision-free and blocked ones (either kind fills in when
    the scene runs short of the other)."""
    easy, hard = [], []
    for _ in range(max_rounds):
        q = uniform_configs(model, batch, rng)
        ok = q[check(model, scene, q).astype(bool)]
        m = len(ok) // 2
        if m == 0:
            continue
        a, b = ok[:m], ok[m:2 * m]
        free = check_edge(model, scene, a, b).astype(bool)
        easy += list(zip(a[free], b[free]))
        hard += list(zip(a[~free], b[~free]))
        if len(easy) >= (n + 1) // 2 and len(hard) >= n // 2:
            break
    out = []
    while len(out) < n and (easy or hard):
        src = easy if (len(out) % 2 == 0 and easy) or not hard else hard
        out.append(src.pop())
    return out


def c5_library(models, n_scenes, pairs_per_scene, check, check_edge, seed=0, batch=2048):
    """Scenes for served planning: (robot index, scene, pairs), obstacle counts log-uniform in
    [5, 200] over box / cylinder / sphere, each with a pool of collision-free start/goal pairs of
    uneven difficulty."""
    rng = np.random.default_rng(seed + 55)
    lib = []
    while len(lib) < n_scenes:
        k = int(rng.integers(len(models)))
        lo, hi, keep = workspace(models[k])
        scene = random_scene(rng, log_uniform_count(rng), lo, hi, keep)
        pairs = sample_pairs(models[k], scene, rng, check, check_edge, pairs_per_scene, batch=batch)
        if len(pairs) >= max(2, pairs_per_scene // 2):
            lib.append((k, scene, pairs))
    return lib


def c5_stream(lib, rate, seconds, p_reuse=0.8, resident_max=256, seed=0):
    """Poisson arrivals for served planning over a scene library.  Each arrival is (t, scene,
    pair, upload): with probability p_reuse the query goes to a scene already resident on the
    device (uniformly), otherwise it brings the library's next scene, which is uploaded at its
    arrival (upload = True); the resident set keeps the last resident_max scenes brought in.
    Each query takes its scene's next start/goal pair (cycling through the pool)."""
    rng = np.random.default_rng(seed + 9)
    times = poisson_arrivals(rate, seconds, seed=seed)
    resident, nxt, used = [], 0, {}
    out = []
    for t in times:
        if resident and rng.random() < p_reuse:
            s = resident[int(rng.integers(len(resident)))]
            up = False
        else:
            s = nxt % len(lib)
            nxt += 1
            up = True
            if s in resident:
                resident.remove(s)
            resident.append(s)
            if len(resident) > resident_max:
                resident.pop(0)
        p = used.get(s, 0)
        used[s] = p + 1
        out.append((float(t), s, p % len(lib[s][2]), up))
    return out
