"""Obstacle scene container shared by the product wrapper, the workloads and the tests.

Obstacle records (float32, world = robot base frame):
  spheres   [ns, 4]  x, y, z, radius                  (EnvConstraintSphere inputs, CPRRTC Sphere POD)
  boxes     [nb, 18] pos(3), R(9 row-major), min(3), max(3)
                                                       (EnvConstraintCuboid(name, pos, R, max, min))
  cylinders [nc, 14] pos(3), R(9), radius, height     (EnvConstraintCylinder; axis = local z)
  capsules  [ncap, 7] p1(3), p2(3), radius            (pRRTC Capsule, utils.cuh:113-127)
"""
from dataclasses import dataclass, field

import numpy as np


def _empty(w):
    return np.zeros((0, w), np.float32)


@dataclass
class Scene:
    spheres: np.ndarray = field(default_factory=lambda: _empty(4))
    boxes: np.ndarray = field(default_factory=lambda: _empty(18))
    cylinders: np.ndarray = field(default_factory=lambda: _empty(14))
    capsules: np.ndarray = field(default_factory=lambda: _empty(7))
    # joint bounds of the checks on this scene: None for the robot's limits, "none" for no
    # joint-bound check (a space without a BoundaryConstraint), or (lower, upper) over the active
    # joints (a BoundaryConstraint's own bounds; +-inf leaves a side open)
    joint_bounds: object = None

    def __post_init__(self):
        self.spheres = np.ascontiguousarray(self.spheres, np.float32).reshape(-1, 4)
        self.boxes = np.ascontiguousarray(self.boxes, np.float32).reshape(-1, 18)
        self.cylinders = np.ascontiguousarray(self.cylinders, np.float32).reshape(-1, 14)
        self.capsules = np.ascontiguousarray(self.capsules, np.float32).reshape(-1, 7)

    @property
    def counts(self):
        return (len(self.spheres), len(self.boxes), len(self.cylinders), len(self.capsules))

    @property
    def num_obstacles(self):
        return sum(self.counts)


def rpy_matrix(roll, pitch, yaw):
    """Rz(yaw) Ry(pitch) Rx(roll), row-major (obstacle_generator.hpp:78-104), float64."""
    cx, sx = np.cos(roll), np.sin(roll)
    cy, sy = np.cos(pitch), np.sin(pitch)
    cz, sz = np.cos(yaw), np.sin(yaw)
    return np.array([
        [cz * cy, cz * sy * sx - sz * cx, cz * sy * cx + sz * sx],
        [sz * cy, sz * sy * sx + cz * cx, sz * sy * cx - cz * sx],
        [-sy, cy * sx, cy * cx]])


def random_rotation(rng):
    """Uniform random rotation (quaternion method), float64 3x3."""
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def box_record(pos, R, half):
    half = np.asarray(half, np.float64)
    return np.concatenate([np.asarray(pos, np.float64), np.asarray(R, np.float64).reshape(9), -half, half]).astype(np.float32)


def cylinder_record(pos, R, radius, height):
    return np.concatenate([np.asarray(pos, np.float64), np.asarray(R, np.float64).reshape(9), [radius, height]]).astype(np.float32)


def capsule_record(pos, R, radius, length):
    axis = np.asarray(R, np.float64)[:, 2]
    p1 = np.asarray(pos, np.float64) - axis * (length / 2)
    p2 = np.asarray(pos, np.float64) + axis * (length / 2)
    return np.concatenate([p1, p2, [radius]]).astype(np.float32)
