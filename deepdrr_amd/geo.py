"""Projective-geometry subset used by the DRR path.

deepdrr re-exports these names from the external ``killeengeo`` package
(``deepdrr/geo/__init__.py:19-50``).  Only what the projector path consumes is
provided here: homogeneous points/vectors, rigid ``FrameTransform``s, the pinhole
``CameraIntrinsicTransform`` and the ``CameraProjection`` whose
``world_from_index`` the kernel receives (``projector.py:803``).

Ray-scale convention.  The projector slices rows 0-2 of a 4x3
``world_from_index`` (``projector.py:803``).  The 4x3 shape is the
pseudo-inverse of ``K [I | 0]`` composed with ``world_from_camera3d``:
``[[R K^-1], [0 0 0]]``.  A pixel's ray therefore has unit depth along the
camera z axis, so ``ray_length`` (``project_kernel.cu:230``) is about 1 mm and
the march starts about 1 mm in front of the source.  ``RAY_SCALE`` names this
choice; with ``"source_to_pixel"`` the projector rescales the rays to end on
the detector plane.
"""
from __future__ import annotations

import math
from typing import Iterable, Optional, Sequence, Tuple, Union

import numpy as np
from scipy.spatial.transform import Rotation

# "unit_depth": world_from_index = R K^-1 (unit camera-z depth).
# "source_to_pixel": the projector scales R K^-1 by the source-to-detector
# distance so that each ray ends on the detector plane.
RAY_SCALE = "unit_depth"


class HomogeneousObject:
    data: np.ndarray

    def __array__(self, dtype=None, copy=None):
        return np.asarray(self.data[:-1], dtype=dtype)

    def __getitem__(self, i):
        return float(self.data[i])

    def __len__(self):
        return self.dim

    @property
    def dim(self) -> int:
        return self.data.shape[0] - 1

    @property
    def x(self) -> float:
        return float(self.data[0])

    @property
    def y(self) -> float:
        return float(self.data[1])

    @property
    def z(self) -> float:
        return float(self.data[2])

    def tolist(self):
        return self.data[:-1].tolist()


class Point(HomogeneousObject):
    def __init__(self, data: np.ndarray):
        self.data = np.asarray(data, dtype=np.float64)

    def __repr__(self):
        return f"Point({self.data[:-1].tolist()})"

    def __sub__(self, other):
        if isinstance(other, Point):
            return Vector(np.concatenate([self.data[:-1] - other.data[:-1], [0.0]]))
        if isinstance(other, Vector):
            return Point(np.concatenate([self.data[:-1] - other.data[:-1], [1.0]]))
        return NotImplemented

    def __add__(self, other):
        if isinstance(other, Vector):
            return Point(np.concatenate([self.data[:-1] + other.data[:-1], [1.0]]))
        return NotImplemented

    def lerp(self, other: "Point", t: float) -> "Point":
        return point((1 - t) * np.array(self) + t * np.array(other))


class Vector(HomogeneousObject):
    def __init__(self, data: np.ndarray):
        self.data = np.asarray(data, dtype=np.float64)

    def __repr__(self):
        return f"Vector({self.data[:-1].tolist()})"

    def __add__(self, other):
        if isinstance(other, Vector):
            return Vector(np.concatenate([self.data[:-1] + other.data[:-1], [0.0]]))
        if isinstance(other, Point):
            return other + self
        return NotImplemented

    def __sub__(self, other):
        if isinstance(other, Vector):
            return Vector(np.concatenate([self.data[:-1] - other.data[:-1], [0.0]]))
        return NotImplemented

    def __neg__(self):
        return Vector(np.concatenate([-self.data[:-1], [0.0]]))

    def __mul__(self, s):
        return Vector(np.concatenate([self.data[:-1] * float(s), [0.0]]))

    __rmul__ = __mul__

    def __truediv__(self, s):
        return Vector(np.concatenate([self.data[:-1] / float(s), [0.0]]))

    def norm(self) -> float:
        return float(np.linalg.norm(self.data[:-1]))

    def hat(self) -> "Vector":
        return self / self.norm()

    def dot(self, other: "Vector") -> float:
        return float(np.dot(self.data[:-1], other.data[:-1]))

    def cross(self, other: "Vector") -> "Vector":
        return vector(np.cross(self.data[:-1], other.data[:-1]))

    def angle(self, other: "Vector") -> float:
        c = self.dot(other) / (self.norm() * other.norm())
        return float(math.acos(max(-1.0, min(1.0, c))))


Point3D = Point
Vector3D = Vector


def point(*args) -> Point:
    if len(args) == 1:
        a = args[0]
        if isinstance(a, Point):
            return a
        if isinstance(a, Vector):
            raise TypeError("cannot convert a Vector to a Point")
        a = np.asarray(a, dtype=np.float64).reshape(-1)
    else:
        a = np.asarray(args, dtype=np.float64)
    return Point(np.concatenate([a, [1.0]]))


def vector(*args) -> Vector:
    if len(args) == 1:
        a = args[0]
        if isinstance(a, Vector):
            return a
        if isinstance(a, Point):
            a = a.data[:-1]
        a = np.asarray(a, dtype=np.float64).reshape(-1)
    else:
        a = np.asarray(args, dtype=np.float64)
    return Vector(np.concatenate([a, [0.0]]))


p = point
v = vector


class Transform:
    """A projective transform stored as a dense matrix (any shape)."""

    def __init__(self, data: np.ndarray, _inv: Optional[np.ndarray] = None):
        self.data = np.asarray(data, dtype=np.float64)
        self._inv = None if _inv is None else np.asarray(_inv, dtype=np.float64)

    def __array__(self, dtype=None, copy=None):
        return np.asarray(self.data, dtype=dtype)

    def toarray(self) -> np.ndarray:
        return np.array(self.data)

    def __getitem__(self, idx):
        return self.data[idx]

    @property
    def shape(self):
        return self.data.shape

    @property
    def inv(self) -> "Transform":
        if self._inv is not None:
            return Transform(self._inv, _inv=self.data)
        if self.data.shape[0] == self.data.shape[1]:
            return Transform(np.linalg.inv(self.data), _inv=self.data)
        return Transform(np.linalg.pinv(self.data), _inv=self.data)

    def inverse(self) -> "Transform":
        return self.inv

    def __matmul__(self, other):
        if isinstance(other, (Point, Vector)):
            out = self.data @ other.data
            if isinstance(other, Point):
                if out.shape[0] == other.data.shape[0]:
                    return Point(out / out[-1])
                return Point(out / out[-1])
            return Vector(out)
        if isinstance(other, Transform):
            return Transform(self.data @ other.data)
        return NotImplemented


class FrameTransform(Transform):
    """Rigid/affine 4x4 homogeneous transform between 3-D frames."""

    def __init__(self, data: np.ndarray, _inv: Optional[np.ndarray] = None):
        data = np.asarray(data, dtype=np.float64)
        if data.shape != (4, 4):
            raise ValueError(f"FrameTransform needs a 4x4 matrix, got {data.shape}")
        super().__init__(data, _inv)

    @classmethod
    def identity(cls, dim: int = 3) -> "FrameTransform":
        return cls(np.eye(dim + 1))

    @classmethod
    def from_rt(cls, rotation=None, translation=None) -> "FrameTransform":
        m = np.eye(4)
        if rotation is not None:
            if isinstance(rotation, Rotation):
                rotation = rotation.as_matrix()
            m[:3, :3] = np.asarray(rotation, dtype=np.float64)
        if translation is not None:
            m[:3, 3] = np.asarray(translation, dtype=np.float64).reshape(-1)[:3]
        return cls(m)

    @classmethod
    def from_scaling(cls, scaling, translation=None) -> "FrameTransform":
        return cls.from_rt(np.diag(np.asarray(scaling, dtype=np.float64).reshape(-1)[:3]), translation)

    @classmethod
    def from_translation(cls, translation) -> "FrameTransform":
        return cls.from_rt(None, translation)

    @property
    def R(self) -> np.ndarray:
        return self.data[:3, :3].copy()

    @property
    def t(self) -> np.ndarray:
        return self.data[:3, 3].copy()

    @property
    def inv(self) -> "FrameTransform":
        if self._inv is not None:
            return FrameTransform(self._inv, _inv=self.data)
        return FrameTransform(np.linalg.inv(self.data), _inv=self.data)

    def inverse(self) -> "FrameTransform":
        return self.inv

    def copy(self) -> "FrameTransform":
        return FrameTransform(self.data.copy())

    def __matmul__(self, other):
        if isinstance(other, FrameTransform):
            return FrameTransform(self.data @ other.data)
        return super().__matmul__(other)

    def __repr__(self):
        return f"FrameTransform({self.data.tolist()})"


F = FrameTransform
f = FrameTransform


def frame_transform(x=None) -> FrameTransform:
    if x is None:
        return FrameTransform.identity()
    if isinstance(x, FrameTransform):
        return x
    if isinstance(x, Rotation):
        return FrameTransform.from_rt(x.as_matrix())
    a = np.asarray(x, dtype=np.float64)
    if a.shape == (4, 4):
        return FrameTransform(a)
    if a.shape == (3, 3):
        return FrameTransform.from_rt(a)
    if a.shape == (3,):
        return FrameTransform.from_translation(a)
    raise TypeError(f"cannot make a FrameTransform from shape {a.shape}")


class CameraIntrinsicTransform(Transform):
    """Pinhole intrinsics ``K`` (3x3) plus the sensor size in pixels."""

    def __init__(self, data: np.ndarray, sensor_height: Optional[int] = None,
                 sensor_width: Optional[int] = None):
        data = np.asarray(data, dtype=np.float64)
        if data.shape != (3, 3):
            raise ValueError("intrinsic matrix must be 3x3")
        super().__init__(data)
        self._sensor_height = sensor_height
        self._sensor_width = sensor_width

    @classmethod
    def from_parameters(cls, optical_center, focal_length=1.0, sensor_size=None):
        if np.isscalar(focal_length):
            focal_length = (focal_length, focal_length)
        if sensor_size is None:
            sensor_size = (int(2 * optical_center[0]), int(2 * optical_center[1]))
        elif np.isscalar(sensor_size):
            sensor_size = (int(sensor_size), int(sensor_size))
        k = np.array([[focal_length[0], 0, optical_center[0]],
                      [0, focal_length[1], optical_center[1]],
                      [0, 0, 1]], dtype=np.float64)
        return cls(k, sensor_width=int(sensor_size[0]), sensor_height=int(sensor_size[1]))

    @classmethod
    def from_sizes(cls, sensor_size, pixel_size, source_to_detector_distance):
        """``sensor_size`` is (width, height) in pixels, ``pixel_size`` in mm."""
        if np.isscalar(sensor_size):
            sensor_size = (int(sensor_size), int(sensor_size))
        if np.isscalar(pixel_size):
            pixel_size = (float(pixel_size), float(pixel_size))
        fx = source_to_detector_distance / pixel_size[0]
        fy = source_to_detector_distance / pixel_size[1]
        cx = sensor_size[0] / 2
        cy = sensor_size[1] / 2
        return cls.from_parameters((cx, cy), (fx, fy), sensor_size)

    @property
    def fx(self) -> float:
        return float(self.data[0, 0])

    @property
    def fy(self) -> float:
        return float(self.data[1, 1])

    @property
    def cx(self) -> float:
        return float(self.data[0, 2])

    @property
    def cy(self) -> float:
        return float(self.data[1, 2])

    @property
    def sensor_width(self) -> int:
        return int(self._sensor_width if self._sensor_width is not None else round(2 * self.cx))

    @property
    def sensor_height(self) -> int:
        return int(self._sensor_height if self._sensor_height is not None else round(2 * self.cy))

    @property
    def sensor_size(self) -> Tuple[int, int]:
        """(width, height), the order ``Projector.initialize_output_arrays`` uses."""
        return (self.sensor_width, self.sensor_height)


class CameraProjection(Transform):
    """``index_from_world = K [I|0] camera3d_from_world`` (3x4)."""

    def __init__(self, intrinsic: CameraIntrinsicTransform,
                 extrinsic: Optional[FrameTransform] = None):
        if not isinstance(intrinsic, CameraIntrinsicTransform):
            intrinsic = CameraIntrinsicTransform(np.asarray(intrinsic))
        self.intrinsic = intrinsic
        self.extrinsic = frame_transform(extrinsic)
        proj = intrinsic.data @ np.hstack([np.eye(3), np.zeros((3, 1))]) @ self.extrinsic.data
        super().__init__(proj)

    @property
    def index_from_camera2d(self) -> CameraIntrinsicTransform:
        return self.intrinsic

    @property
    def camera3d_from_world(self) -> FrameTransform:
        return self.extrinsic

    @property
    def world_from_camera3d(self) -> FrameTransform:
        return self.extrinsic.inv

    @property
    def index_from_world(self) -> Transform:
        return Transform(self.data)

    @property
    def world_from_index(self) -> Transform:
        """4x3 map from homogeneous pixel index to a world-space ray vector."""
        kinv = np.linalg.inv(self.intrinsic.data)
        m = np.zeros((4, 3))
        m[:3, :] = self.world_from_camera3d.data[:3, :3] @ kinv
        return Transform(m)

    @property
    def center_in_world(self) -> Point:
        return point(self.world_from_camera3d.data[:3, 3])

    @property
    def principle_ray_in_world(self) -> Vector:
        return vector(self.world_from_camera3d.data[:3, 2]).hat()

    @property
    def sensor_width(self) -> int:
        return self.intrinsic.sensor_width

    @property
    def sensor_height(self) -> int:
        return self.intrinsic.sensor_height

    @property
    def sensor_size(self) -> Tuple[int, int]:
        return self.intrinsic.sensor_size


RAS_from_LPS = FrameTransform(np.diag([-1.0, -1.0, 1.0, 1.0]))
LPS_from_RAS = RAS_from_LPS


def get_data(x) -> np.ndarray:
    return np.asarray(x.data if hasattr(x, "data") else x)
