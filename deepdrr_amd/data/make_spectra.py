"""Extract the three MC-GPU X-ray spectra into ``spectra.npz``.

The spectra are physical data (MC-GPU spectrum files, Badal & Badano 2009) that
deepdrr keeps as literal arrays in ``deepdrr/projector/spectral_data.py:7-467``.
This script reads that file as *text* (it is never imported or executed),
parses the ``[energy_eV, photons]`` pairs of each array and stores them as
float64 arrays under the same names the reference's ``spectrums`` dict uses
(``spectral_data.py:463-467``).

Run once, in a checkout that has the reference sources:

    python deepdrr_amd/data/make_spectra.py /path/to/deepdrr/projector/spectral_data.py
"""
import re
import sys
from pathlib import Path

import numpy as np

NAMES = {
    "SPECTRUM60KV_AL35": "60KV_AL35",
    "SPECTRUM90KV_AL40": "90KV_AL40",
    "SPECTRUM120KV_AL43": "120KV_AL43",
}


def parse(text: str) -> dict:
    out = {}
    for var, key in NAMES.items():
        m = re.search(var + r"\s*=\s*np\.array\(\[(.*?)\]\)", text, re.S)
        if m is None:
            raise ValueError(f"{var} not found")
        pairs = re.findall(r"\[\s*([-+0-9.eE]+)\s*,\s*([-+0-9.eE]+)\s*\]", m.group(1))
        out[key] = np.array([[float(a), float(b)] for a, b in pairs], dtype=np.float64)
    return out


def main() -> None:
    src = Path(sys.argv[1])
    arrays = parse(src.read_text())
    dst = Path(__file__).with_name("spectra.npz")
    np.savez(dst, **arrays)
    for k, v in arrays.items():
        print(k, v.shape, v[0], v[-1])


if __name__ == "__main__":
    main()
