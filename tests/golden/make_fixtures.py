"""Writes tests/golden/ref_2d_70k.npz from the reference's input point cloud
build/geometries_2d/2d_70k.h5 (datasets x, y, llGrainIDs; float32 attribute
`volume` on llGrainIDs), read with the h5dump tool of the image's HDF5.

Run in the build container only (it reads /root/reference); the GPU runs use
the committed .npz.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

REF = "/root/reference/build/geometries_2d/2d_70k.h5"
H5DUMP = "/opt/conda/bin/h5dump"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_2d_70k.npz")


def dump(ds, dtype):
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "a.bin")
        subprocess.check_call([H5DUMP, "-d", ds, "-b", "LE", "-o", f, REF], stdout=subprocess.DEVNULL)
        return np.fromfile(f, dtype=dtype)


def main():
    x = dump("/x", "<f8")
    y = dump("/y", "<f8")
    g = dump("/llGrainIDs", "<u8")
    txt = subprocess.check_output([H5DUMP, "-a", "/llGrainIDs/volume", REF]).decode()
    vol = np.float32(float(txt.split("(0):")[1].split()[0]))
    assert len(x) == len(y) == len(g)
    np.savez_compressed(OUT, x=x, y=y, llGrainIDs=g, volume=np.array([vol], dtype=np.float32))
    print(OUT, len(x), vol)


if __name__ == "__main__":
    sys.exit(main())
