"""Extracts a small realistic fixture from the reference's example data.

Reads the four example tilt series (HAADF + Al/Si/Yb EDX sinograms, 36 tilts
x 60 slices x 433 bins, example/*.mrc + .rawtlt of kbredies/graptor, GPLv3)
with graptor_amd.io and keeps slices 28..29 of every channel, so GPU tests
can run the solver on real electron-tomography data without the reference
on the machine.  Run in the build container only:

    python tests/golden/make_example_fixture.py /root/reference/example
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from graptor_amd.io import read_tilt_series  # noqa: E402

CHANNELS = ("HAADF_lrsino", "Al_EDXsino", "Si_EDXsino", "Yb_EDXsino")


def main(src):
    data, angles = [], None
    for name in CHANNELS:
        d, a = read_tilt_series(os.path.join(src, name + ".mrc"))
        data.append(d[:, 28:30, :].astype(np.float32))       # (tilts, 2 slices, bins)
        angles = a if angles is None else angles
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "example_2slices.npz")
    np.savez_compressed(out, sino=np.stack(data), angles=angles,
                        channels=np.array(CHANNELS), slices=np.array([28, 29]))
    print(out, np.stack(data).shape)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/example")
