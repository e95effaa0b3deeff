"""X-ray spectra ``[energy_eV, photons]`` (deepdrr/projector/spectral_data.py).

The three named MC-GPU spectra are loaded from ``data/spectra.npz`` (extracted
from the reference's literal arrays by ``data/make_spectra.py``).  Each ends in
a negative last bin, as the reference's do (``spectral_data.py:461``).

``synthetic_120kv`` is the 20-bin smooth 120 kV spectrum of the benchmark
workloads: E = 20, 25, ..., 115 keV with weight E (120 - E)^2 exp(-E / 40).
"""
from __future__ import annotations

import os
from typing import Union

import numpy as np

_NPZ = os.path.join(os.path.dirname(__file__), "data", "spectra.npz")
with np.load(_NPZ) as _z:
    spectrums = {k: np.array(_z[k]) for k in _z.files}

SPECTRUM60KV_AL35 = spectrums["60KV_AL35"]
SPECTRUM90KV_AL40 = spectrums["90KV_AL40"]
SPECTRUM120KV_AL43 = spectrums["120KV_AL43"]


def synthetic_120kv(n_bins: int = 20) -> np.ndarray:
    e_keV = 20.0 + 5.0 * np.arange(n_bins, dtype=np.float64)
    w = e_keV * (120.0 - e_keV) ** 2 * np.exp(-e_keV / 40.0)
    return np.stack([e_keV * 1000.0, w], axis=1)


spectrums["SYNTH120KV_20"] = synthetic_120kv()


def get_spectrum(spectrum: Union[np.ndarray, str]) -> np.ndarray:
    """projector.py:260-279: an ``[B, 2]`` array passes through, a name is looked up."""
    if isinstance(spectrum, np.ndarray):
        return spectrum
    if isinstance(spectrum, str):
        if spectrum not in spectrums:
            raise KeyError(f"unrecognized spectrum: {spectrum}")
        return spectrums[spectrum]
    raise TypeError(f"unrecognized spectrum type: {type(spectrum)}")


def spectrum_tables(spectrum_arr: np.ndarray):
    """Energies in keV (float32) and the normalised pdf (float32), projector.py:1660-1673."""
    energies = np.ascontiguousarray(spectrum_arr[:, 0].copy() / 1000, dtype=np.float32)
    pdf = spectrum_arr[:, 1] / np.sum(spectrum_arr[:, 1])
    pdf = np.ascontiguousarray(pdf.copy(), dtype=np.float32)
    return energies, pdf
