"""Robot model data (RobotInfo layout, robot_info.hpp:13-220) loaded from the JSON files in robots/.

The arrays here are exactly what ``SingleArmSpace`` / ``CPRRTC::RobotSolver`` take
(cudampl/CUDAMPLib/include/spaces/SingleArmSpace.h:24-40, cudampl/CPRRTC/include/RobotSolver.h:84-101):
joint types, row-major 4x4 joint origins, joint axes, parent-link map, active-joint map,
bounds and defaults (full-joint indexed), collision spheres and the link-pair enable map.
"""
import json
import os
from dataclasses import dataclass

import numpy as np

ROBOT_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "robots")


@dataclass
class RobotModel:
    name: str
    links: list
    joint_types: np.ndarray      # int32 [L]
    joint_origins: np.ndarray    # float64 [L,16]
    joint_axes: np.ndarray       # float64 [L,3]
    parents: np.ndarray          # int32 [L]
    active: np.ndarray           # int32 [L]
    lower: np.ndarray            # float32 [L]
    upper: np.ndarray            # float32 [L]
    defaults: np.ndarray         # float32 [L]
    sphere_link: np.ndarray      # int32 [S]
    sphere_pos: np.ndarray       # float32 [S,3]
    sphere_radius: np.ndarray    # float32 [S]
    enabled: np.ndarray          # uint8 [L,L]
    joint_names: list = None     # name of each link's parent joint ("" for the root)

    @property
    def L(self):
        return len(self.links)

    @property
    def S(self):
        return len(self.sphere_radius)

    @property
    def dof(self):
        return int(self.active.sum())

    @property
    def active_lower(self):
        return self.lower[self.active.astype(bool)]

    @property
    def active_upper(self):
        return self.upper[self.active.astype(bool)]

    def without_links(self, link_names):
        """Copy of the model with the collision spheres of the named links removed."""
        drop = {self.links.index(n) for n in link_names}
        keep = np.array([l not in drop for l in self.sphere_link])
        m = RobotModel(**{k: getattr(self, k) for k in self.__dataclass_fields__})
        m.sphere_link = self.sphere_link[keep].copy()
        m.sphere_pos = self.sphere_pos[keep].copy()
        m.sphere_radius = self.sphere_radius[keep].copy()
        m.name = self.name + "_no_" + "_".join(link_names)
        return m


def load_robot(name: str) -> RobotModel:
    with open(os.path.join(ROBOT_DIR, name + ".json")) as f:
        d = json.load(f)
    return RobotModel(
        name=d["name"],
        links=list(d["links"]),
        joint_types=np.asarray(d["joint_types"], np.int32),
        joint_origins=np.asarray(d["joint_origins"], np.float64).reshape(-1, 16),
        joint_axes=np.asarray(d["joint_axes"], np.float64).reshape(-1, 3),
        parents=np.asarray(d["parents"], np.int32),
        active=np.asarray(d["active"], np.int32),
        lower=np.asarray(d["lower"], np.float32),
        upper=np.asarray(d["upper"], np.float32),
        defaults=np.asarray(d["defaults"], np.float32),
        sphere_link=np.asarray(d["sphere_link"], np.int32),
        sphere_pos=np.asarray(d["sphere_pos"], np.float32).reshape(-1, 3),
        sphere_radius=np.asarray(d["sphere_radius"], np.float32),
        enabled=np.asarray(d["collision_enabled"], np.uint8),
        joint_names=list(d["joint_names"]),
    )
