// Obstacle PODs of the CPRRTC API (layout of cudampl/CPRRTC/include/Obstacle.h:4-42).
// Orientation convention: R = Rz(yaw) Ry(pitch) Rx(roll) (obstacle_generator.hpp:78-104).
#pragma once
namespace CPRRTC {
struct Cuboid {
    float x_min, x_max, y_min, y_max, z_min, z_max;  // extents in the box frame
    float x, y, z, roll, pitch, yaw;                  // pose
};
struct Sphere {
    float x, y, z, radius;
};
struct Cylinder {
    float x, y, z, roll, pitch, yaw;
    float radius, height;  // axis = local z, centred
};
}  // namespace CPRRTC
