"""The reference's ROS node (scripts/gvom_ros.py) without ROS: its parameters, input
conversions and published payloads around a `Gvom` mapper.

`GvomNode` keeps the node's structure -- cb_odom (gvom_ros.py:79-80), cb_lidar (82-109),
cb_timer (113-198) -- so a rospy wrapper only forwards messages and publishes what
cb_timer returns:

- ego position from the latest Odometry (pose.pose.position), as a 3-tuple;
- the lidar-to-odom transform from a tf translation + quaternion (x, y, z, w), built
  as tf.TransformerROS.fromTranslationRotation does (`transform_from_translation_rotation`);
- the PointCloud2 message bytes go to the device as they are and x, y, z are gathered
  there (`Gvom.process_pointcloud2`), in place of ros_numpy's pointcloud2_to_xyz_array;
- each timer tick runs combine_maps and returns the OccupancyGrid payloads (int8, made
  on the device) and grid info, plus the three debug clouds as numpy record arrays with
  the node's field names.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

import gvom

# Parameter names and defaults of the node (gvom_ros.py:23-41).
DEFAULTS = dict(odom_frame="/camera_init", xy_resolution=0.40, z_resolution=0.2, width=256, height=64,
                buffer_size=4, min_point_distance=1.0, positive_obstacle_threshold=0.50,
                negative_obstacle_threshold=0.5, density_threshold=50, slope_obsacle_threshold=0.3,
                min_roughness=-10, max_roughness=0, robot_height=2.0, robot_radius=4.0,
                ground_to_lidar_height=1.0, freq=10., xy_eigen_dist=1, z_eigen_dist=1)

# Topics published from one timer tick (gvom_ros.py:61-67, 137-164); the ground
# certainty payload goes to two topics.
GRID_TOPICS = ("hard_obstacle_map", "soft_obstacle_map", "ground_certainty_map", "all_ground_certainty_map",
               "negative_obstacle_map", "roughness_map")

_QUAT_EPS = np.finfo(np.float64).eps * 4.0


def transform_from_translation_rotation(translation, rotation):
    """4x4 float64 homogeneous transform from a translation and a unit quaternion
    (x, y, z, w), as tf.TransformerROS.fromTranslationRotation (gvom_ros.py:104): the
    quaternion is normalised through a scale by sqrt(2 / |q|^2) and the rotation is
    built from the products of its scaled components (a quaternion with |q|^2 below
    4 * machine epsilon gives the identity rotation)."""
    t = np.asarray(translation, dtype=np.float64).reshape(3)
    q = np.array(rotation, dtype=np.float64).reshape(4)
    T = np.identity(4)
    nq = float(np.dot(q, q))
    if nq >= _QUAT_EPS:
        q = q * np.sqrt(2.0 / nq)
        x, y, z, w = q
        xx, yy, zz = x * x, y * y, z * z
        xy, xz, yz = x * y, x * z, y * z
        xw, yw, zw = x * w, y * w, z * w
        T[0, 0], T[0, 1], T[0, 2] = 1.0 - yy - zz, xy - zw, xz + yw
        T[1, 0], T[1, 1], T[1, 2] = xy + zw, 1.0 - xx - zz, yz - xw
        T[2, 0], T[2, 1], T[2, 2] = xz - yw, yz + xw, 1.0 - xx - yy
    T[:3, 3] = t
    return T


def _position(odom):
    """Odometry message (pose.pose.position) or a 3-sequence -> 3-tuple of floats."""
    if hasattr(odom, "pose"):
        p = odom.pose.pose.position
        return (float(p.x), float(p.y), float(p.z))
    x, y, z = odom
    return (float(x), float(y), float(z))


@dataclass
class GridInfo:
    """nav_msgs/MapMetaData as the node fills it (gvom_ros.py:124-135)."""
    frame_id: str
    resolution: float
    width: int
    height: int
    origin_position: tuple
    origin_orientation: tuple = (0.0, 0.0, 0.0, 1.0)


@dataclass
class TimerOutput:
    info: GridInfo
    grids: dict                      # topic -> int8 payload (width * width, x fastest)
    debug: dict = field(default_factory=dict)   # "voxel" / "height_map" / "inferred_height_map" records


class GvomNode:
    """VoxelMapper of gvom_ros.py:14-198 with the ROS transport taken out."""

    def __init__(self, **params):
        unknown = set(params) - set(DEFAULTS)
        if unknown:
            raise TypeError(f"unknown parameters: {sorted(unknown)}")
        self.p = dict(DEFAULTS, **params)
        p = self.p
        self.odom_data = None
        self.voxel_mapper = gvom.Gvom(p["xy_resolution"], p["z_resolution"], p["width"], p["height"],
                                      p["buffer_size"], p["min_point_distance"], p["positive_obstacle_threshold"],
                                      p["negative_obstacle_threshold"], p["slope_obsacle_threshold"],
                                      p["robot_height"], p["robot_radius"], p["ground_to_lidar_height"],
                                      p["xy_eigen_dist"], p["z_eigen_dist"])

    def cb_odom(self, odom):
        self.odom_data = _position(odom)

    def cb_lidar(self, cloud, translation, rotation, device_data=None):
        """One PointCloud2 with the tf lookup's translation and rotation (x, y, z, w).
        Returns False (after printing "no odom") when no odometry has arrived yet."""
        if self.odom_data is None:
            print("no odom")
            return False
        tf_matrix = transform_from_translation_rotation(translation, rotation)
        # An organised cloud (height = rings, width = azimuths) keeps its rows on the device,
        # non-finite returns included (the scan stage drops them), so the ray cast can trace
        # ring patches; the order does not change the counts.
        height, width = int(getattr(cloud, "height", 1)), int(getattr(cloud, "width", 0))
        stride = width if height >= 64 and width > 0 else 0
        if stride != getattr(self, "_ray_stride", None):
            self.voxel_mapper.set_ray_rings(stride, 64)
            self._ray_stride = stride
        self.voxel_mapper.process_pointcloud2(cloud, self.odom_data, tf_matrix, device_data=device_data)
        return True

    def cb_timer(self, debug=True):
        """combine_maps and the payloads of every published topic, or None when the
        buffer is empty."""
        p = self.p
        m = self.voxel_mapper
        # combine on the device; only the int8 payloads come back to the host
        if m.combine_maps_device() is None:
            print("ERROR: No data in buffer")
            print("map_data is None. returning.")
            return None
        out = m.occupancy_grids(p["density_threshold"], p["min_roughness"], p["max_roughness"])
        origin = out.pop("origin")
        info = GridInfo(p["odom_frame"], p["xy_resolution"], int(p["width"]), int(p["width"]),
                        (float(origin[0]), float(origin[1]), 0.0))
        grids = {t: out["ground_certainty_map" if t == "all_ground_certainty_map" else t] for t in GRID_TOPICS}
        res = TimerOutput(info, grids)
        if debug:
            res.debug = self.debug_clouds()
        return res

    def debug_clouds(self):
        """The node's three debug PointCloud2 payloads as record arrays
        (gvom_ros.py:168-195)."""
        m = self.voxel_mapper
        out = {}
        v = m.make_debug_voxel_map()
        if v is not None:
            out["voxel"] = np.rec.fromarrays([v[:, k] for k in range(8)],
                                             names="x,y,z,solid factor,count,eigen_line,eigen_surface,eigen_point")
        h = m.make_debug_height_map()
        if h is not None:
            obs = m.maps_state()["pos"]
            out["height_map"] = np.rec.fromarrays([h[:, k] for k in range(7)] + [obs.flatten("F")],
                                                  names="x,y,z,roughness,slope_x,slope_y,slope,obstacles")
        i = m.make_debug_inferred_height_map()
        if i is not None:
            out["inferred_height_map"] = np.rec.fromarrays([i[:, k] for k in range(3)], names="x,y,z")
        return out
