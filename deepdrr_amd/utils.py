"""Host-side image epilogue of the DRR path.

``neglog`` follows ``deepdrr/utils/image_utils.py:18-59`` (float32 NumPy, as
the reference computes it); ``add_noise`` follows
``deepdrr/projector/analytic_generators.py:6-18``.
"""
from __future__ import annotations

import logging

import numpy as np
import scipy.signal

log = logging.getLogger(__name__)

KERNEL_SHOT_NOISE = np.array([[0.03, 0.06, 0.02], [0.11, 0.98, 0.11], [0.02, 0.06, 0.03]])
FACTOR_ELECTRONIC_NOISE = 0
KERNEL_ELECTRONIC_NOISE = np.array([[0.0, 0.0, 0.0, 0.0, 0.0], [0.0, 0.0, 0.0, 0.0, 0.0],
                                    [0.01, 0.01, 0.98, 0.01, 0.01], [0.0, 0.0, 0.0, 0.0, 0.0],
                                    [0.0, 0.0, 0.0, 0.0, 0.0]])


NEGLOG_EPSILON = 0.01  # image_utils.py:18 default


def neglog(image: np.ndarray, epsilon: float = NEGLOG_EPSILON) -> np.ndarray:
    """x += min(x) + eps; y = -log(x); min-max normalise per image (0 if constant)."""
    image = np.array(image)
    shape = image.shape
    if len(shape) == 2:
        image = image[np.newaxis, :, :]
    image += image.min(axis=(1, 2), keepdims=True) + epsilon
    image = -np.log(image)
    image_min = image.min(axis=(1, 2), keepdims=True)
    image_max = image.max(axis=(1, 2), keepdims=True)
    if np.any(image_max == image_min):
        log.debug("mapping constant image to 0")
        image[:] = 0
    else:
        image = (image - image_min) / (image_max - image_min)
    if np.any(np.isnan(image)):
        log.warning("got NaN values from negative log transform.")
    return image[0] if len(shape) == 2 else image


def add_noise(input_image: np.ndarray, photon_prob: np.ndarray, photon_number: float,
              rng: np.random.Generator | None = None) -> np.ndarray:
    """Poisson shot noise correlated by a 3x3 kernel, in float64, clipped to [0, 1e31]."""
    rng = rng if rng is not None else np.random.default_rng()
    out = np.zeros(input_image.shape)
    for i in range(input_image.shape[0]):
        lam = photon_prob[i] * photon_number
        shot = (rng.poisson(lam) - lam) * input_image[i] / lam
        shot = scipy.signal.convolve2d(shot, KERNEL_SHOT_NOISE, mode="same")
        elec = rng.standard_normal(input_image.shape[1:]) * FACTOR_ELECTRONIC_NOISE
        elec = scipy.signal.convolve2d(elec, KERNEL_ELECTRONIC_NOISE, mode="same")
        out[i] = np.clip(input_image[i] + shot + elec, a_min=0.0, a_max=10e30)
    return out


def listify(x):
    return list(x) if isinstance(x, (list, tuple)) else [x]
