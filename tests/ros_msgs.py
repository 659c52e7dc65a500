"""Minimal stand-ins for the ROS messages the node consumes (sensor_msgs/PointField,
sensor_msgs/PointCloud2, nav_msgs/Odometry), with the same attribute names, and a
builder that packs points into PointCloud2 bytes."""
from __future__ import annotations

from dataclasses import dataclass, field
from types import SimpleNamespace

import numpy as np

INT8, UINT8, INT16, UINT16, INT32, UINT32, FLOAT32, FLOAT64 = range(1, 9)
_NP = {INT8: np.int8, UINT8: np.uint8, INT16: np.int16, UINT16: np.uint16, INT32: np.int32, UINT32: np.uint32,
       FLOAT32: np.float32, FLOAT64: np.float64}


@dataclass
class PointField:
    name: str
    offset: int
    datatype: int
    count: int = 1


@dataclass
class PointCloud2:
    height: int
    width: int
    fields: list
    is_bigendian: bool
    point_step: int
    row_step: int
    data: bytes
    is_dense: bool = False
    extra: dict = field(default_factory=dict)


def odometry(x, y, z):
    return SimpleNamespace(pose=SimpleNamespace(pose=SimpleNamespace(position=SimpleNamespace(x=x, y=y, z=z))))


def pack_cloud(points, height, width, layout, big=False, row_pad=0, intensity=True):
    """points: (height*width, 3) values; layout: {"x": (offset, datatype), ...} for x, y, z
    (plus an optional float32 intensity field after them); point_step = end of the last
    field rounded up to 8 bytes + 8 spare bytes; row_pad extra bytes per row."""
    order = ">" if big else "<"
    fields = [PointField(n, off, dt) for n, (off, dt) in layout.items()]
    end = max(off + np.dtype(_NP[dt]).itemsize for off, dt in layout.values())
    if intensity:
        fields.append(PointField("intensity", end, FLOAT32))
        end += 4
    point_step = ((end + 7) // 8) * 8 + 8
    row_step = width * point_step + row_pad
    buf = np.zeros(height * row_step, np.uint8)
    pts = np.asarray(points)
    for k, name in enumerate(("x", "y", "z")):
        off, dt = layout[name]
        t = np.dtype(_NP[dt]).newbyteorder(order)
        vals = pts[:, k].astype(t)
        raw = vals.view(np.uint8).reshape(-1, t.itemsize)
        for r in range(height):
            base = r * row_step + off
            idx = base + np.arange(width)[:, None] * point_step + np.arange(t.itemsize)[None, :]
            buf[idx] = raw[r * width:(r + 1) * width]
    if intensity:
        off = fields[-1].offset
        t = np.dtype(np.float32).newbyteorder(order)
        raw = np.full(height * width, 7.5, t).view(np.uint8).reshape(-1, 4)
        for r in range(height):
            idx = r * row_step + off + np.arange(width)[:, None] * point_step + np.arange(4)[None, :]
            buf[idx] = raw[r * width:(r + 1) * width]
    return PointCloud2(height, width, fields, big, point_step, row_step, buf.tobytes())
