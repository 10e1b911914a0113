"""Gmsh v2.2 ASCII reader producing the engine's mesh input.

Follows what Apollo's mesh path hands to the RHS (src/solvers/apsolver.cc:
296-359, DMPlexCreateGmsh): cells are the triangles (element type 2) in file
order, vertices are the $Nodes in file order, each cell's cone is its node
list as written, and boundary faces carry the physical tag of the line
element (type 1) on the same vertex pair as their "Face Sets" label.
A path ending in ".gz" is read through gzip.
"""
import gzip

import numpy as np

from .engine import MeshData


def read_gmsh(path):
    opener = gzip.open if str(path).endswith(".gz") else open
    with opener(path, "rt") as fh:
        lines = [ln.strip() for ln in fh]
    i = lines.index("$MeshFormat")
    version = lines[i + 1].split()[0]
    if not version.startswith("2"):
        raise ValueError("only Gmsh format 2.x is supported, got %s" % version)
    i = lines.index("$Nodes")
    nn = int(lines[i + 1])
    ids, VX, VY = {}, np.zeros(nn), np.zeros(nn)
    for j in range(nn):
        parts = lines[i + 2 + j].split()
        ids[int(parts[0])] = j
        VX[j], VY[j] = float(parts[1]), float(parts[2])
    i = lines.index("$Elements")
    ne = int(lines[i + 1])
    tris, edges = [], {}
    for j in range(ne):
        parts = [int(p) for p in lines[i + 2 + j].split()]
        etype, ntags = parts[1], parts[2]
        tags = parts[3:3 + ntags]
        nodes = [ids[p] for p in parts[3 + ntags:]]
        phys = tags[0] if ntags > 0 else 0
        if etype == 2:
            tris.append(nodes[:3])
        elif etype == 1:
            edges[(min(nodes[0], nodes[1]), max(nodes[0], nodes[1]))] = phys
    EToV = np.array(tris, dtype=np.int32)
    face_tag = np.zeros_like(EToV)
    for k, t in enumerate(tris):
        for f, (a, b) in enumerate(((0, 1), (1, 2), (2, 0))):
            key = (min(t[a], t[b]), max(t[a], t[b]))
            face_tag[k, f] = edges.get(key, 0)
    return MeshData(VX, VY, EToV, None, face_tag)
