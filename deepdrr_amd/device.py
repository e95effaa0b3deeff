"""Imaging devices that turn a pose into a ``CameraProjection``.

Follows ``deepdrr/device/device.py:29-130`` (frame chain and
``get_camera_projection``), ``deepdrr/device/simple_device.py:10-165``
(``camera_intrinsics`` and ``set_view``) and
``deepdrr/device/mobile_carm.py:57-275`` (defaults, ``device_from_arm``,
``camera3d_from_device``).
"""
from __future__ import annotations

import math
from typing import Optional

import numpy as np
from scipy.spatial.transform import Rotation

from . import geo


class Device:
    sensor_height: int
    sensor_width: int
    pixel_size: float
    source_to_detector_distance: float
    world_from_device: geo.FrameTransform

    @property
    def detector_height(self) -> float:
        return self.sensor_height * self.pixel_size

    @property
    def detector_width(self) -> float:
        return self.sensor_width * self.pixel_size

    @property
    def device_from_world(self) -> geo.FrameTransform:
        return self.world_from_device.inverse()

    @property
    def camera3d_from_device(self) -> geo.FrameTransform:
        return self.device_from_camera3d.inverse()

    @property
    def world_from_camera3d(self) -> geo.FrameTransform:
        return self.world_from_device @ self.device_from_camera3d

    @property
    def camera3d_from_world(self) -> geo.FrameTransform:
        return self.camera3d_from_device @ self.device_from_world

    def get_camera_projection(self) -> geo.CameraProjection:
        return geo.CameraProjection(self.camera_intrinsics, self.camera3d_from_world)

    @property
    def index_from_world(self) -> geo.CameraProjection:
        return self.get_camera_projection()

    @property
    def source_in_world(self) -> geo.Point3D:
        return self.world_from_camera3d @ geo.point(0, 0, 0)

    @property
    def principle_ray_in_world(self) -> geo.Vector3D:
        return (self.world_from_camera3d @ geo.vector(0, 0, 1)).hat()


def _rotvec_between(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a = a / np.linalg.norm(a)
    b = b / np.linalg.norm(b)
    c = np.cross(a, b)
    ang = math.acos(max(-1.0, min(1.0, float(np.dot(a, b)))))
    return c / np.linalg.norm(c) * ang


class SimpleDevice(Device):
    def __init__(self, sensor_height: int = 384, sensor_width: int = 384, pixel_size: float = 1.0,
                 source_to_detector_distance: float = 1000.0,
                 world_from_device: Optional[geo.FrameTransform] = None):
        self.sensor_height = sensor_height
        self.sensor_width = sensor_width
        self.pixel_size = pixel_size
        self.source_to_detector_distance = source_to_detector_distance
        self.world_from_device = geo.frame_transform(world_from_device)
        self.set_view([0, 0, 0], [0, 0, 1], [0, -1, 0])

    def scale_sensor(self, detector_size: float) -> None:
        if self.sensor_height < self.sensor_width:
            self.pixel_size = detector_size / self.sensor_height
        else:
            self.pixel_size = detector_size / self.sensor_width

    @property
    def camera_intrinsics(self) -> geo.CameraIntrinsicTransform:
        f = self.source_to_detector_distance / self.pixel_size
        k = np.array([[f, 0, self.sensor_width / 2], [0, f, self.sensor_height / 2], [0, 0, 1]])
        return geo.CameraIntrinsicTransform(k, sensor_height=self.sensor_height,
                                            sensor_width=self.sensor_width)

    def set_view(self, point=None, direction=None, up=None,
                 source_to_point_distance: Optional[float] = None,
                 source_to_point_fraction: float = 0.5) -> None:
        """Source-detector midpoint-style pose (simple_device.py:80-165)."""
        if source_to_point_distance is None:
            source_to_point_distance = self.source_to_detector_distance * source_to_point_fraction
        if point is None:
            p_dev = np.array(self.device_from_camera3d @ geo.point(0, 0, source_to_point_distance))
        else:
            p_dev = np.array(self.device_from_world @ geo.point(point))
        if direction is None:
            d_dev = np.array(self.device_from_camera3d @ geo.vector(0, 0, 1))
        else:
            d_dev = np.array(self.device_from_world @ geo.vector(direction))
        if up is None:
            up_dev = np.array([0.0, -1.0, 0.0])
        else:
            up_dev = np.array(self.device_from_world @ geo.vector(up))

        z = np.array([0.0, 0.0, 1.0])
        rv = np.cross(z, d_dev)
        if np.linalg.norm(rv) < 1e-6:
            rv = np.array([np.pi, 0, 0]) if np.dot(z, d_dev) < 0 else np.zeros(3)
        else:
            rv = _rotvec_between(z, d_dev)
        device_from_ray = geo.FrameTransform.from_rt(Rotation.from_rotvec(rv).as_matrix(), p_dev)

        neg_y = np.array([0.0, -1.0, 0.0])
        up_ray = np.array(device_from_ray.inverse() @ geo.vector(up_dev))
        up_plane = np.array([up_ray[0], up_ray[1], 0.0])
        rv = np.cross(neg_y, up_plane)
        if np.linalg.norm(rv) < 1e-6:
            rv = np.zeros(3)
        else:
            rv = _rotvec_between(neg_y, up_plane)
        ray_from_ray_up = geo.FrameTransform.from_rt(Rotation.from_rotvec(rv).as_matrix())
        ray_up_from_camera3d = geo.FrameTransform.from_rt(
            translation=[0.0, 0.0, -source_to_point_distance])
        self._device_from_camera3d = device_from_ray @ ray_from_ray_up @ ray_up_from_camera3d

    @property
    def device_from_camera3d(self) -> geo.FrameTransform:
        return self._device_from_camera3d


class MobileCArm(Device):
    """Orbital (alpha) / angulation (beta) C-arm with the Cios-Fusion defaults."""

    def __init__(self, world_from_device=None, isocenter=(0, 0, 0), alpha: float = 0,
                 beta: float = 0, gamma: float = 0, degrees: bool = True,
                 source_to_detector_distance: float = 1020,
                 source_to_isocenter_vertical_distance: float = 530,
                 source_to_isocenter_horizontal_offset: float = 0,
                 sensor_height: int = 1536, sensor_width: int = 1536, pixel_size: float = 0.194,
                 rotate_camera_left: bool = True, **kwargs):
        rad = (lambda a: math.radians(a)) if degrees else (lambda a: float(a))
        self.world_from_device = geo.frame_transform(world_from_device)
        self.isocenter = np.asarray(isocenter, dtype=np.float64)
        self.alpha, self.beta, self.gamma = rad(alpha), rad(beta), rad(gamma)
        self.source_to_detector_distance = source_to_detector_distance
        self.source_to_isocenter_vertical_distance = source_to_isocenter_vertical_distance
        self.source_to_isocenter_horizontal_offset = source_to_isocenter_horizontal_offset
        self.sensor_height = sensor_height
        self.sensor_width = sensor_width
        self.pixel_size = pixel_size
        self.rotate_camera_left = rotate_camera_left
        self.camera_intrinsics = geo.CameraIntrinsicTransform.from_sizes(
            sensor_size=(sensor_width, sensor_height), pixel_size=pixel_size,
            source_to_detector_distance=source_to_detector_distance)

    def move_to(self, isocenter=None, alpha=None, beta=None, gamma=None, degrees=True):
        rad = (lambda a: math.radians(a)) if degrees else (lambda a: float(a))
        if isocenter is not None:
            self.isocenter = np.asarray(isocenter, dtype=np.float64)
        if alpha is not None:
            self.alpha = rad(alpha)
        if beta is not None:
            self.beta = rad(beta)
        if gamma is not None:
            self.gamma = rad(gamma)

    @property
    def device_from_arm(self) -> geo.FrameTransform:
        rot = Rotation.from_euler("xy", [self.alpha, self.beta]).as_matrix()
        return geo.FrameTransform.from_rt(rotation=rot, translation=self.isocenter)

    @property
    def arm_from_device(self) -> geo.FrameTransform:
        return self.device_from_arm.inverse()

    @property
    def camera3d_from_device(self) -> geo.FrameTransform:
        c_from_arm = geo.FrameTransform.from_rt(translation=[
            0, -self.source_to_isocenter_horizontal_offset, self.source_to_isocenter_vertical_distance])
        if self.rotate_camera_left:
            c_from_arm = geo.frame_transform(Rotation.from_euler("z", 90, degrees=True)) @ c_from_arm
        g = geo.frame_transform(Rotation.from_euler("z", self.gamma, degrees=False))
        return g @ c_from_arm @ self.arm_from_device

    @property
    def device_from_camera3d(self) -> geo.FrameTransform:
        return self.camera3d_from_device.inverse()
