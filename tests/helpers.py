"""Small synthetic robots and geometry references used by the tests."""
import numpy as np

from isaac_sim_ros2_humble_amd.model import RobotModel


def point_robot(radius=0.05, lim=2.0):
    """Three prismatic joints x, y, z carrying one sphere: configuration = sphere centre."""
    L = 4
    T = np.tile(np.eye(4).reshape(16), (L, 1))
    return RobotModel(
        name="point", links=["base", "lx", "ly", "lz"],
        joint_types=np.array([0, 2, 2, 2], np.int32), joint_origins=T,
        joint_axes=np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float64),
        parents=np.array([-1, 0, 1, 2], np.int32), active=np.array([0, 1, 1, 1], np.int32),
        lower=np.array([0, -lim, -lim, -lim], np.float32), upper=np.array([0, lim, lim, lim], np.float32),
        defaults=np.zeros(4, np.float32), sphere_link=np.array([3], np.int32),
        sphere_pos=np.zeros((1, 3), np.float32), sphere_radius=np.array([radius], np.float32),
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
