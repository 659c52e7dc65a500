"""CPU restatement of the reference ROS node's input and output conversions --
TEST INFRASTRUCTURE ONLY (imported by tests, never by the product).

- `pointcloud2_to_xyz`: ros_numpy.point_cloud2.pointcloud2_to_xyz_array as the node calls
  it (gvom_ros.py:108; ros_numpy is a ROS package dependency, package.xml:53, unpinned
  and not vendored): a structured numpy view of the message bytes (one record per
  point_step bytes, field dtypes from the PointField codes, byte order from
  is_bigendian), x / y / z copied into float64, and rows with any non-finite
  coordinate removed.
- `from_translation_rotation`: tf.TransformerROS.fromTranslationRotation
  (gvom_ros.py:104; tf, package.xml:56, unpinned): numpy.dot of the translation matrix
  and tf.transformations.quaternion_matrix (quaternion x, y, z, w; |q|^2 below
  4 * eps -> identity; else q scaled by sqrt(2 / |q|^2) and R from numpy.outer(q, q)).
- `occupancy_payloads`: the int8 OccupancyGrid data of gvom_ros.py:141-164 from the
  combine_maps outputs, with the node's numpy expressions.
"""
from __future__ import annotations

import math

import numpy as np

# sensor_msgs/PointField datatype code -> numpy type
PF_TYPES = {1: np.int8, 2: np.uint8, 3: np.int16, 4: np.uint16, 5: np.int32, 6: np.uint32, 7: np.float32,
            8: np.float64}


def pointcloud2_to_xyz(data, height, width, point_step, fields, is_bigendian=False):
    """fields: {name: (offset, datatype)}.  Returns float64 (M, 3)."""
    order = ">" if is_bigendian else "<"
    names, formats, offsets = [], [], []
    for name, (off, dt) in fields.items():
        names.append(name)
        formats.append(np.dtype(PF_TYPES[dt]).newbyteorder(order))
        offsets.append(off)
    dtype = np.dtype({"names": names, "formats": formats, "offsets": offsets, "itemsize": point_step})
    rec = np.frombuffer(bytes(data), dtype=dtype, count=height * width)
    rec = rec.reshape(height, width)
    mask = np.isfinite(rec["x"]) & np.isfinite(rec["y"]) & np.isfinite(rec["z"])
    rec = rec[mask]
    pts = np.zeros(rec.shape + (3,), dtype=np.float64)
    pts[..., 0] = rec["x"]
    pts[..., 1] = rec["y"]
    pts[..., 2] = rec["z"]
    return pts


def from_translation_rotation(translation, rotation):
    T = np.identity(4)
    T[:3, 3] = np.asarray(translation, np.float64)[:3]
    q = np.array(rotation[:4], dtype=np.float64, copy=True)
    nq = np.dot(q, q)
    if nq < np.finfo(float).eps * 4.0:
        R = np.identity(4)
    else:
        q *= math.sqrt(2.0 / nq)
        q = np.outer(q, q)
        R = np.array([[1.0 - q[1, 1] - q[2, 2], q[0, 1] - q[2, 3], q[0, 2] + q[1, 3], 0.0],
                      [q[0, 1] + q[2, 3], 1.0 - q[0, 0] - q[2, 2], q[1, 2] - q[0, 3], 0.0],
                      [q[0, 2] - q[1, 3], q[1, 2] + q[0, 3], 1.0 - q[0, 0] - q[1, 1], 0.0],
                      [0.0, 0.0, 0.0, 1.0]])
    return np.dot(T, R)


def occupancy_payloads(pos, neg, rough, vis, density_threshold=50, min_roughness=-10, max_roughness=0):
    """pos / neg / vis int32 [x, y], rough float64 [x, y] -> dict of int8 payloads."""
    with np.errstate(invalid="ignore", divide="ignore"):
        hard = np.reshape(np.maximum(100 * (pos > density_threshold), neg), -1, order="F").astype(np.int8)
        soft = np.reshape(100 * (pos <= density_threshold) * (pos > 0), -1, order="F").astype(np.int8)
        cert = np.reshape(vis * 100, -1, order="F").astype(np.int8)
        negp = np.reshape(neg, -1, order="F").astype(np.int8)
        r = ((np.maximum(np.minimum(rough, max_roughness), min_roughness) + min_roughness)
             / (max_roughness - min_roughness)) * 100
        rgh = np.reshape(r, -1, order="F").astype(np.int8)
    return {"hard_obstacle_map": hard, "soft_obstacle_map": soft, "ground_certainty_map": cert,
            "negative_obstacle_map": negp, "roughness_map": rgh}
