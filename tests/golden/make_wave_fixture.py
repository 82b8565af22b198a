"""Extracts a small wavenumber fixture from the reference's own data files.

InertialFriction/post_process/kx.txt and ky.txt are the reference's dump of
Mesh::kx / Mesh::ky (written by wave_visual in EnergySpec.cu:152-153) for a
768 x 768 mesh with Lx = Ly = 33*2*pi, printed with 6 significant digits,
one row per ky index j, Nxh = 385 columns (comma-terminated).

Run in a container that holds the reference checkout:
    python tests/golden/make_wave_fixture.py /root/reference
Writes tests/golden/wave_768_L66pi.json (rows 0,1,2,383,384,385,386,767 and
columns 0,1,2,191,383,384).
"""
import json
import os
import sys

ROWS = [0, 1, 2, 383, 384, 385, 386, 767]
COLS = [0, 1, 2, 191, 383, 384]


def read(path):
    out = {}
    with open(path) as fh:
        for j, line in enumerate(fh):
            if j in ROWS:
                vals = [v for v in line.strip().split(",") if v != ""]
                out[j] = [float(vals[i]) for i in COLS]
    return out


def main(ref_root):
    pp = os.path.join(ref_root, "InertialFriction", "post_process")
    kx, ky = read(os.path.join(pp, "kx.txt")), read(os.path.join(pp, "ky.txt"))
    fx = {"source": "InertialFriction/post_process/{kx,ky}.txt", "N": 768,
          "L": "33*2*pi", "rows": ROWS, "cols": COLS, "significant_digits": 6,
          "kx": [kx[j] for j in ROWS], "ky": [ky[j] for j in ROWS]}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "wave_768_L66pi.json")
    with open(dst, "w") as fh:
        json.dump(fx, fh, indent=1)
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
