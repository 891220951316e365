"""Writes tests/golden/geometry_ref.npz: S / oid / bmap of small scenes computed by the
reference's own geometry code (parsing3D.h CPU path + filterABC.h build_boundary_map,
compiled into oracle/_ref by oracle/build_ref.sh). Run where the reference exists:

    python tests/golden/make_geometry_golden.py
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))
from oracle import oracle as O  # noqa: E402

BOX_FACES = [("west", "4 8 5 1"), ("east", "2 6 7 3"), ("south", "1 5 6 2"), ("north", "3 7 8 4"),
             ("floor", "1 2 3 4"), ("ceiling", "5 8 7 6")]


def write_box(path, nx, ny, nz, h):
    h = np.float32(h)
    X, Y, Z = [np.float32(n) * h for n in (nx, ny, nz)]
    V = [(0, 0, 0), (X, 0, 0), (X, Y, 0), (0, Y, 0), (0, 0, Z), (X, 0, Z), (X, Y, Z), (0, Y, Z)]
    with open(path, "w") as f:
        f.write("o Box\n")
        for v in V:
            f.write("v %.9g %.9g %.9g\n" % tuple(float(a) for a in v))
        for name, q in BOX_FACES:
            f.write("g %s\nf %s\n" % (name, q))


CASES = [
    ("box_30x20x25_h0.034", "box", (30, 20, 25), 0.034),
    ("box_17x13x11_h0.05", "box", (17, 13, 11), 0.05),
    ("room_modes_h0.2", "scenes/room_modes_shoebox.obj", None, 0.2),
    ("default_room_h0.15", "scenes/default_project_room.obj", None, 0.15),
]


def main():
    out = {}
    for name, src, box, h in CASES:
        if src == "box":
            path = "/tmp/_fases_golden_box.obj"
            write_box(path, *box, h)
        else:
            path = str(HERE / src)
        dims, S, oid = O.ref_scene(path, np.float32(h))
        bmap, nB = O.ref_boundary_map(oid)
        out[name + "/dims"] = np.array(dims)
        out[name + "/S"] = S
        out[name + "/oid"] = oid
        out[name + "/bmap"] = bmap
        out[name + "/h"] = np.float32(h)
        out[name + "/box"] = np.array(box if box else (0, 0, 0))
        print(name, dims, nB)
    np.savez_compressed(HERE / "geometry_ref.npz", **out)


if __name__ == "__main__":
    main()
