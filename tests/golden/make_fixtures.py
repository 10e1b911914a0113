#!/usr/bin/env python3
"""Write the golden fixtures under tests/golden/ from the reference checkout.

Reads reference data files as text (nothing is compiled, imported or run):

* ``src/lib/dataN01.h .. dataN08.h`` -- NUDG-generated tables for order N:
  nodes p_r/p_s, Vandermonde p_Vand and its inverse p_IVand (N=1 only), nodal
  differentiation p_Dr/p_Ds, the strong-form lift p_LIFT and NUDG's face
  masks p_Fmask.  Stored as ``nudg_N<N>.npz`` (plain arrays, no pickle).
* five Gmsh v2.2 meshes from ``examples/unstructuredDG`` copied verbatim into
  ``meshes/`` (data the reference ships with its example decks), and six
  larger ones gzip-compressed (``.msh.gz``, the text unchanged): those of the
  Euler shock decks, and those of the isentropic-vortex and
  transverse-magnetic decks (grid2.msh, square3.msh).

Usage: python tests/golden/make_fixtures.py [/root/reference]
"""
import gzip
import os
import re
import shutil
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
NUM = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?"

MESHES = {
    "advection_test.msh": "examples/unstructuredDG/advection/test.msh",
    "small_square_4bc.msh": "examples/unstructuredDG/multifluid/collisionalRadiative/small_square_4bc.msh",
    "vortex_grid.msh": "examples/unstructuredDG/euler/isentropicVortex/grid.msh",
    "tm_square.msh": "examples/unstructuredDG/maxwell/transverseMagnetic/square.msh",
    "ffs_grid.msh": "examples/unstructuredDG/euler/forwardFacingStep/grid.msh",
}

GZ_MESHES = {
    "pulse_grid2.msh.gz": "examples/unstructuredDG/euler/eulerPulse/grid2.msh",
    "scramjet_grid.msh.gz": "examples/unstructuredDG/euler/scramjet/grid.msh",
    "explosion_gridBox.msh.gz": "examples/unstructuredDG/euler/explosionTest/gridBox.msh",
    "bfs_grid.msh.gz": "examples/unstructuredDG/euler/backwardFacingStep/grid.msh",
    "vortex_grid2.msh.gz": "examples/unstructuredDG/euler/isentropicVortex/grid2.msh",
    "tm_square3.msh.gz": "examples/unstructuredDG/maxwell/transverseMagnetic/square3.msh",
}


def parse_array(src, name):
    m = re.search(r"(?:double|int)\s+%s((?:\[\d+\])+)\s*=\s*\{(.*?)\};" % name, src, re.S)
    if not m:
        return None
    dims = [int(d) for d in re.findall(r"\[(\d+)\]", m.group(1))]
    vals = [float(v) for v in re.findall(NUM, m.group(2))]
    assert len(vals) == int(np.prod(dims)), (name, dims, len(vals))
    return np.array(vals).reshape(dims)


def main():
    for N in range(1, 9):
        src = open(os.path.join(REF, "src/lib/dataN%02d.h" % N)).read()
        out = {}
        for name in ("p_r", "p_s", "p_Vand", "p_IVand", "p_Dr", "p_Ds", "p_LIFT", "p_Fmask"):
            a = parse_array(src, name)
            if a is None:
                # only dataN01.h carries p_Vand / p_IVand
                assert name in ("p_Vand", "p_IVand"), (N, name)
                continue
            out[name[2:]] = a.astype(np.int32) if name == "p_Fmask" else a
        np.savez(os.path.join(HERE, "nudg_N%d.npz" % N), **out)
    os.makedirs(os.path.join(HERE, "meshes"), exist_ok=True)
    for dst, src in MESHES.items():
        shutil.copyfile(os.path.join(REF, src), os.path.join(HERE, "meshes", dst))
    for dst, src in GZ_MESHES.items():
        with open(os.path.join(REF, src), "rb") as fi:
            data = fi.read()
        with open(os.path.join(HERE, "meshes", dst), "wb") as fo:
            with gzip.GzipFile(filename="", mode="wb", fileobj=fo, mtime=0) as gz:
                gz.write(data)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
