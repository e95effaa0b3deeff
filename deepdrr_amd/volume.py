"""CT volume: density, material segmentation and the IJK <-> world pose.

Follows ``deepdrr/vol/volume.py`` (``Volume.__init__`` 152-192,
``from_parameters`` 213-278, ``from_hu`` 280-302,
``_convert_hounsfield_to_density`` 338-351, ``_format_materials`` 955-992,
``shape`` 994, ``get_center`` 1001) and ``deepdrr/vol/renderable.py:58-109``
(``world_from_IJK``, ``IJK_from_world``, ``center_in_world``), plus the HU
threshold segmentation of ``deepdrr/load_dicom.py:132-143``.
"""
from __future__ import annotations

from typing import Dict, List, Optional, Sequence, Tuple, Union

import numpy as np

from . import geo


def conv_hu_to_density(hu_values: np.ndarray) -> np.ndarray:
    """rho = max(min(0.001029 HU + 1.030, 0.0005886 HU + 1.03), 0), FP64 (load_dicom.py:117-129)."""
    hu = np.asarray(hu_values)
    return np.maximum(np.minimum(0.001029 * hu + 1.030, 0.0005886 * hu + 1.03), 0)


def conv_hu_to_materials_thresholding(hu_values: np.ndarray) -> Dict[str, np.ndarray]:
    """air <= -800 < soft tissue <= 350 < bone (load_dicom.py:132-143)."""
    return {
        "air": hu_values <= -800,
        "soft tissue": (-800 < hu_values) * (hu_values <= 350),
        "bone": 350 < hu_values,
    }


def segment_by_thresholds(hu_values: np.ndarray,
                          classes: Sequence[Tuple[str, float]]) -> Dict[str, np.ndarray]:
    """Generalised thresholding: ``classes = [(name, upper_HU), ...]`` ascending;
    class i takes ``lower_i < HU <= upper_i`` (the last upper may be ``inf``)."""
    out = {}
    lo = -np.inf
    for name, hi in classes:
        out[name] = (lo < hu_values) & (hu_values <= hi) if np.isfinite(lo) else hu_values <= hi
        lo = hi
    return out


# conv_hu_to_materials_thresholding as classes for segment_by_thresholds
THRESHOLD_CLASSES = (
    ("air", -800.0),
    ("soft tissue", 350.0),
    ("bone", np.inf),
)

FIVE_MATERIAL_CLASSES = (
    ("air", -900.0),
    ("lung", -200.0),
    ("soft tissue", 300.0),
    ("bone", 3000.0),
    ("titanium", np.inf),
)


class VolumePose:
    """Pose and grid size of a volume without its voxel data."""

    def __init__(self, anatomical_from_IJK, shape, world_from_anatomical=None, enabled=True):
        self.anatomical_from_IJK = geo.frame_transform(anatomical_from_IJK)
        self.world_from_anatomical = geo.frame_transform(world_from_anatomical)
        self._shape = tuple(int(s) for s in shape)
        self.enabled = enabled

    @property
    def shape(self):
        return self._shape

    @property
    def world_from_IJK(self) -> geo.FrameTransform:
        return self.world_from_anatomical @ self.anatomical_from_IJK

    @property
    def IJK_from_world(self) -> geo.FrameTransform:
        return self.world_from_IJK.inverse()

    def get_center(self) -> geo.Point3D:
        return self.anatomical_from_IJK @ geo.point(np.array(self.shape) / 2)

    @property
    def center_in_world(self) -> geo.Point3D:
        return self.world_from_anatomical @ self.get_center()


class Volume:
    def __init__(
        self,
        data: np.ndarray,
        materials: Union[Dict[str, np.ndarray], Tuple[Dict[str, int], np.ndarray]],
        anatomical_from_IJK: Optional[geo.FrameTransform] = None,
        world_from_anatomical: Optional[geo.FrameTransform] = None,
        anatomical_coordinate_system: Optional[str] = None,
        enabled: bool = True,
        anatomical_from_ijk: Optional[geo.FrameTransform] = None,
        **kwargs,
    ) -> None:
        if anatomical_from_ijk is not None:
            anatomical_from_IJK = anatomical_from_ijk
        self.anatomical_from_IJK = geo.frame_transform(anatomical_from_IJK)
        self.world_from_anatomical = (geo.FrameTransform.identity(3) if world_from_anatomical is None
                                      else geo.frame_transform(world_from_anatomical))
        hu_source = kwargs.pop("_hu_source", None)
        self._hu = None
        self._hu_classes = None
        if hu_source is not None:
            # HU held as int16: density and labels are derived when first read
            # on the host, or on the GPU by the projector (drr_volume_create_from_hu)
            self._hu, self._hu_classes = hu_source
            self._data = None
            self._materials = None
            self._material_ids = {name: i for i, (name, _) in enumerate(self._hu_classes)}
        else:
            data = np.asarray(data)
            assert data.ndim == 3, "Volume data must be 3D."
            self._data = np.array(data).astype(np.float32)
            if isinstance(materials, tuple):
                self._materials = materials[0], np.asarray(materials[1]).astype(np.uint16)
            else:
                self._materials = self._format_materials(materials)
            self._material_ids = self._materials[0]
        self.anatomical_coordinate_system = anatomical_coordinate_system
        assert self.anatomical_coordinate_system in ["LPS", "RAS", None]
        self.enabled = enabled

    # -- voxel data (derived from HU on first host access) ---------------------
    def _materialize(self) -> None:
        hu = self._hu
        self._data = self._convert_hounsfield_to_density(hu).astype(np.float32)
        self._materials = self._format_materials(segment_by_thresholds(hu, self._hu_classes))

    @property
    def data(self) -> np.ndarray:
        if self._data is None:
            self._materialize()
        return self._data

    @data.setter
    def data(self, value) -> None:
        if self._data is None and self._hu is not None:
            self._materialize()
        self._data = np.asarray(value, dtype=np.float32)

    @property
    def materials(self):
        if self._materials is None:
            self._materialize()
        return self._materials

    @materials.setter
    def materials(self, value) -> None:
        if self._data is None and self._hu is not None:
            self._materialize()
        self._materials = value
        self._material_ids = value[0]

    @property
    def material_ids(self) -> Dict[str, int]:
        """Material name -> id of ``materials`` (without deriving the labels)."""
        return self._material_ids

    def hu_source(self):
        """(int16 HU [i, j, k], ((name, upper HU), ...)) while the density and
        labels have not been derived on the host (so nothing can have changed
        them), else None."""
        if self._hu is None or self._data is not None:
            return None
        return self._hu, self._hu_classes

    # -- construction -------------------------------------------------------
    @classmethod
    def from_parameters(cls, data, materials, origin, spacing=(1, 1, 1),
                        anatomical_coordinate_system: Optional[str] = None,
                        world_from_anatomical=None, **kwargs) -> "Volume":
        origin = np.asarray(geo.point(origin))
        spacing = np.asarray(geo.vector(spacing))
        if anatomical_coordinate_system is None or anatomical_coordinate_system == "none":
            raise NotImplementedError
        if anatomical_coordinate_system == "LPS":
            rotation = [[spacing[0], 0, 0], [0, 0, spacing[2]], [0, -spacing[1], 0]]
            anatomical_from_IJK = geo.FrameTransform.from_rt(rotation=rotation, translation=origin)
        elif anatomical_coordinate_system == "RAS":
            raise NotImplementedError("conversion from RAS (not hard, look at LPS example)")
        else:
            raise ValueError()
        return cls(data=data, materials=materials, anatomical_from_IJK=anatomical_from_IJK,
                   world_from_anatomical=world_from_anatomical,
                   anatomical_coordinate_system=anatomical_coordinate_system, **kwargs)

    @classmethod
    def from_hu(cls, hu_values, origin, use_thresholding: bool = True, spacing=(1, 1, 1),
                anatomical_coordinate_system: Optional[str] = None,
                world_from_anatomical=None, material_classes=None, **kwargs) -> "Volume":
        """vol/volume.py:280-302.  An int16 HU array (the on-disk CT format) is
        kept as is: density and labels are derived on first host access, or
        on the GPU when a projector uploads the volume."""
        if material_classes is None:
            if not use_thresholding:
                raise NotImplementedError("network segmentation is not part of this engine")
            classes = THRESHOLD_CLASSES
        else:
            classes = tuple((str(n), float(u)) for n, u in material_classes)
        hu = np.asarray(hu_values)
        uppers = [u for _, u in classes]
        partition = all(a < b for a, b in zip(uppers, uppers[1:])) and uppers[-1] == np.inf
        if hu.ndim == 3 and hu.dtype == np.int16 and partition:
            return cls.from_parameters(None, None, origin=origin, spacing=spacing,
                                       anatomical_coordinate_system=anatomical_coordinate_system,
                                       world_from_anatomical=world_from_anatomical,
                                       _hu_source=(np.ascontiguousarray(hu), classes), **kwargs)
        data = cls._convert_hounsfield_to_density(hu)
        materials = segment_by_thresholds(hu, classes)
        return cls.from_parameters(data, materials, origin=origin, spacing=spacing,
                                   anatomical_coordinate_system=anatomical_coordinate_system,
                                   world_from_anatomical=world_from_anatomical, **kwargs)

    @staticmethod
    def _convert_hounsfield_to_density(hu_values: np.ndarray, smooth_air: bool = False):
        if smooth_air:
            hu_values = np.array(hu_values)
            hu_values[hu_values <= -900] = -1000
        return conv_hu_to_density(hu_values)

    @staticmethod
    def _format_materials(materials: Dict[str, np.ndarray]) -> Tuple[dict, np.ndarray]:
        """Dict order gives ids; later masks win; unlabelled voxels get id 0."""
        combined = None
        material_dict = {}
        for mat_id, mat in enumerate(materials):
            mask = np.asarray(materials[mat]) > 0
            if combined is None:
                combined = np.zeros(mask.shape, dtype=np.uint16)
            combined[mask] = mat_id
            material_dict[mat] = mat_id
        return material_dict, combined

    # -- pose ---------------------------------------------------------------
    def set_enabled(self, enabled: bool) -> None:
        self.enabled = enabled

    @property
    def world_from_IJK(self) -> geo.FrameTransform:
        return self.world_from_anatomical @ self.anatomical_from_IJK

    world_from_ijk = world_from_IJK

    @property
    def IJK_from_world(self) -> geo.FrameTransform:
        return self.world_from_IJK.inverse()

    ijk_from_world = IJK_from_world

    @property
    def shape(self) -> Tuple[int, int, int]:
        return tuple(self._hu.shape) if self._data is None else self._data.shape

    @property
    def spacing(self) -> geo.Vector3D:
        return geo.vector(np.abs(np.array(self.anatomical_from_IJK.R)).max(axis=0))

    def __array__(self, dtype=None, copy=None) -> np.ndarray:
        return np.asarray(self.data, dtype=dtype)

    def get_center(self) -> geo.Point3D:
        return self.anatomical_from_IJK @ geo.point(np.array(self.shape) / 2)

    @property
    def center_in_world(self) -> geo.Point3D:
        return self.world_from_anatomical @ self.get_center()

    def get_bounding_box_in_world(self):
        n = np.array(self.shape, dtype=np.float64)
        corners = np.array([[x, y, z] for x in (-0.5, n[0] - 0.5) for y in (-0.5, n[1] - 0.5)
                            for z in (-0.5, n[2] - 0.5)])
        w = np.array([np.array(self.world_from_IJK @ geo.point(c)) for c in corners])
        return geo.point(w.min(0)), geo.point(w.max(0))

    def place_center(self, x) -> None:
        """Move the volume so its centre lands on world point ``x``."""
        c = np.array(self.center_in_world)
        t = np.array(geo.point(x)) - c
        self.world_from_anatomical = geo.FrameTransform.from_translation(t) @ self.world_from_anatomical

    translate_center_to = place_center
