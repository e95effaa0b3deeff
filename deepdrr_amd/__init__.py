"""MI355X-native DRR engine with deepdrr's Projector API.

    from deepdrr_amd import Volume, SimpleDevice, Projector
    with Projector(volume, device=device) as projector:
        image = projector()
"""
from . import geo, utils
from .device import Device, MobileCArm, SimpleDevice
from .material import Material
from .projector import DeprecationError, Projector
from .volume import Volume

__all__ = ["geo", "utils", "Device", "MobileCArm", "SimpleDevice", "Material", "Projector",
           "DeprecationError", "Volume"]
