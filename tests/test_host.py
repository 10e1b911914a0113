"""The engine's C-ABI library and its host-side logic (CPU only: no compute
call needs a GPU here)."""
import ctypes
import os
import re

import numpy as np
import pytest

from apollo_amd import engine as E
from oracle import oracle as O

from cases import GOLDEN

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "apollo_dg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(apl_\w+)\s*\(", text)))


def test_library_exports_every_header_symbol():
    L = E.lib()
    names = header_functions()
    assert len(names) >= 18
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(E.EXPORTS)


def test_status_strings():
    for s in range(12):
        assert E.status_string(s)
    assert E.status_string(0) == "ok"


def test_create_validates_before_touching_the_device():
    mesh = E.structured_mesh(3, 0, 1, 0, 1)
    bad = [
        (E.Params(9, E.EQN_EULER4), 2),                                      # order
        (E.Params(3, 7), 1),                                                 # eqn
        (E.Params(3, E.EQN_EULER4, flux=E.FLUX_UPWIND), 1),                  # upwind: TM only
        (E.Params(3, E.EQN_MAXWELL_TM3, chi=1.0), 1),                        # TM needs chi=0
        # the limiter: Euler only, N <= 4, no isentropic-vortex boundary
        (E.Params(3, E.EQN_MAXWELL_TM3, limiter=(1.4, 0.0, 0.0)), 1),
        (E.Params(5, E.EQN_EULER5, limiter=(1.4, 0.0, 0.0)), 11),
        (E.Params(3, E.EQN_EULER5, limiter=(1.4, 0.0, 0.0),
                  bcs=[(E.BC_ISENTROPIC_VORTEX, [1, 0, 0, 0, 1.4, 5])]), 11),
        # EULER4 carries no z momentum
        (E.Params(3, E.EQN_EULER4, bcs=[(E.BC_EULER_DIRICHLET, [1, 0, 0, 0.5, 1.4, 1])]), 1),
    ]
    for params, code in bad:
        with pytest.raises(E.ApolloError) as ei:
            E.Engine(mesh, params)
        assert ei.value.status == code
    # clockwise element -> J <= 0 (wxnodaldggeometry2d.cc:428-434)
    cw = E.MeshData(mesh.VX, mesh.VY, mesh.EToV[:, ::-1].copy(), None, None)
    with pytest.raises(E.ApolloError) as ei:
        E.Engine(cw, E.Params(3, E.EQN_MAXWELL_TM3, bcs=[(E.BC_PEC, [])]))
    assert ei.value.status == 3
    # a boundary face whose tag has no boundary condition
    with pytest.raises(E.ApolloError) as ei:
        E.Engine(mesh, E.Params(3, E.EQN_MAXWELL_TM3))
    assert ei.value.status == 5


def test_host_operators_match_oracle():
    for N in range(1, 9):
        o = O.Ops(N)
        for which, ref in ((0, o.Vc), (1, o.Drc), (2, o.Dsc), (3, o.If), (4, o.VVT), (5, o.cw),
                           (8, o.r), (9, o.s)):
            got = E.host_operator(N, which)
            assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max()) * N, (N, which)
        # the Gauss rule follows Apollo's gauleg bit for bit
        assert np.array_equal(E.host_operator(N, 6), o.gz)
        assert np.array_equal(E.host_operator(N, 7), o.gw)


@pytest.mark.parametrize("periodic", [True, False])
def test_host_connectivity_matches_oracle(periodic):
    mesh = E.structured_mesh(9, -5, 5, -5, 5, 0.15, 3, periodic)
    a = E.host_connectivity(mesh)
    b = O.facepair(mesh.EToV, mesh.Nv, mesh.vert_id, mesh.face_tag)
    assert np.array_equal(a, b)
    if periodic:
        assert (a >= 0).all()
    else:
        assert (a < 0).sum() == 2 * 4 * 9     # 36 boundary faces, two entries each


def test_negative_face_tag_is_rejected():
    """FacePair2d stores -tag on boundary faces, so a negative tag would read
    as an interior neighbour (|tag|, |tag|); both entry points refuse it."""
    mesh = E.structured_mesh(3, 0, 1, 0, 1)
    tags = mesh.face_tag.copy()
    tags.flat[np.argmax(tags.ravel() > 0)] = -2
    bad = E.MeshData(mesh.VX, mesh.VY, mesh.EToV, None, tags)
    with pytest.raises(E.ApolloError) as ei:
        E.host_connectivity(bad)
    assert ei.value.status == 1
    with pytest.raises(E.ApolloError) as ei:
        E.Engine(bad, E.Params(3, E.EQN_MAXWELL_TM3, bcs=[(E.BC_PEC, [])] * 2))
    assert ei.value.status == 1


def test_state_arguments_are_checked():
    """The wrappers refuse a state the C-ABI would read past or misread:
    wrong dtype, wrong size, non-contiguous, or host/device mixed."""
    import types
    fake = types.SimpleNamespace(K=4, Np=3, m=2, device=0)
    ptr = lambda a, name="q": E.Engine._ptr(fake, a, name)
    good = np.zeros((4, 3, 2))
    assert ptr(good)[1] == E.HOST
    for bad in (np.zeros((4, 3, 2), dtype=np.float32), np.zeros((4, 3, 3)),
                np.zeros((4, 6, 2))[:, ::2], [0.0] * 24):
        with pytest.raises(ValueError):
            ptr(bad)
    torch = pytest.importorskip("torch")
    assert ptr(torch.zeros(24, dtype=torch.float64))[1] == E.HOST
    for bad in (torch.zeros(24, dtype=torch.float32), torch.zeros(25, dtype=torch.float64),
                torch.zeros(48, dtype=torch.float64)[::2]):
        with pytest.raises(ValueError):
            ptr(bad)


@pytest.mark.parametrize("name", ["vortex_64_N3", "tm_cavity_256_N5", "vortex_512_N4",
                                  "adjoint_vortex_256_N3", "tm_fourier_384_N8"])
def test_workload_meshes_full_size(name):
    """Each benchmark workload's mesh at full size: K, the boundary-face
    count, and the engine's FacePair bit for bit against the oracle's."""
    from apollo_amd import workloads as W
    w = W.get(name)
    mesh = w.mesh()
    assert mesh.K == w.K == 2 * w.n * w.n
    et = E.host_connectivity(mesh)
    assert np.array_equal(et, O.facepair(mesh.EToV, mesh.Nv, mesh.vert_id, mesh.face_tag))
    nb = int((et[:, 1::2] < 0).sum())
    assert nb == (0 if w.periodic else 4 * w.n)
    # interior vertices are jittered by at most jitter * h, boundary ones not
    h = w.L / w.n
    i, j = np.meshgrid(np.arange(w.n + 1), np.arange(w.n + 1), indexing="xy")
    dx = mesh.VX - (w.box[0] + h * i.ravel())
    assert np.abs(dx).max() <= w.jitter * h * (1 + 1e-12) and np.abs(dx).max() > 0.9 * w.jitter * h
    edge = (i.ravel() == 0) | (i.ravel() == w.n) | (j.ravel() == 0) | (j.ravel() == w.n)
    assert np.all(dx[edge] == 0.0)


@pytest.mark.parametrize("name", ["advection_test.msh", "small_square_4bc.msh",
                                  "vortex_grid.msh", "tm_square.msh", "ffs_grid.msh"])
def test_host_connectivity_gmsh_fixtures(name):
    from apollo_amd.gmsh import read_gmsh
    mesh = read_gmsh(os.path.join(GOLDEN, "meshes", name))
    assert np.array_equal(E.host_connectivity(mesh),
                          O.facepair(mesh.EToV, mesh.Nv, None, mesh.face_tag))


@pytest.mark.parametrize("name", ["bfs", "explosion", "pulse", "scramjet"])
def test_shock_deck_meshes(name):
    """The meshes of Apollo's Euler shock decks: the triangle count, every
    triangle counter-clockwise, every boundary edge (an edge of one triangle
    only) tagged with an entry of the deck's boundary list and every entry
    used, and the engine's face pairing equal to the oracle's."""
    from apollo_amd.gmsh import read_gmsh
    from cases import DECKS
    fixture, ntri, _, bcs, _, _ = DECKS[name]
    m = read_gmsh(os.path.join(GOLDEN, "meshes", fixture))
    assert m.K == ntri
    x, y = m.VX[m.EToV], m.VY[m.EToV]
    area = (x[:, 1] - x[:, 0]) * (y[:, 2] - y[:, 0]) - (x[:, 2] - x[:, 0]) * (y[:, 1] - y[:, 0])
    assert (area > 0).all()
    count = {}
    for k in range(m.K):
        for f in range(3):
            a, b = m.EToV[k, f], m.EToV[k, (f + 1) % 3]
            key = (min(a, b), max(a, b))
            count[key] = count.get(key, 0) + 1
    tags = set()
    for k in range(m.K):
        for f in range(3):
            a, b = m.EToV[k, f], m.EToV[k, (f + 1) % 3]
            if count[(min(a, b), max(a, b))] == 1:
                assert 1 <= m.face_tag[k, f] <= len(bcs), (k, f, m.face_tag[k, f])
                tags.add(int(m.face_tag[k, f]))
    assert tags == set(range(1, len(bcs) + 1))
    assert np.array_equal(E.host_connectivity(m), O.facepair(m.EToV, m.Nv, None, m.face_tag))


def test_gmsh_reader():
    from apollo_amd.gmsh import read_gmsh
    mesh = read_gmsh(os.path.join(GOLDEN, "meshes", "advection_test.msh"))
    assert mesh.K == 16 and mesh.Nv == 13
    assert sorted(set(mesh.face_tag.ravel()) - {0}) == [1, 2, 3, 4]
    # every triangle of Apollo's example meshes is counter-clockwise
    for name in ("advection_test.msh", "vortex_grid.msh", "tm_square.msh", "ffs_grid.msh"):
        m = read_gmsh(os.path.join(GOLDEN, "meshes", name))
        x, y = m.VX[m.EToV], m.VY[m.EToV]
        area = (x[:, 1] - x[:, 0]) * (y[:, 2] - y[:, 0]) - (x[:, 2] - x[:, 0]) * (y[:, 1] - y[:, 0])
        assert (area > 0).all()


# --- synthetic meshes -----------------------------------------------------------

class MT19937_64:
    """Reference MT19937-64 (Matsumoto & Nishimura), independent of libstdc++."""

    def __init__(self, seed):
        self.mt = [0] * 312
        self.mt[0] = seed & (2 ** 64 - 1)
        for i in range(1, 312):
            prev = self.mt[i - 1]
            self.mt[i] = (6364136223846793005 * (prev ^ (prev >> 62)) + i) & (2 ** 64 - 1)
        self.idx = 312

    def __call__(self):
        if self.idx >= 312:
            for i in range(312):
                x = (self.mt[i] & 0xFFFFFFFF80000000) | (self.mt[(i + 1) % 312] & 0x7FFFFFFF)
                xa = x >> 1
                if x & 1:
                    xa ^= 0xB5026F5AA96619E9
                self.mt[i] = self.mt[(i + 156) % 312] ^ xa
            self.idx = 0
        x = self.mt[self.idx]
        self.idx += 1
        x ^= (x >> 29) & 0x5555555555555555
        x ^= (x << 17) & 0x71D67FFFEDA60000
        x ^= (x << 37) & 0xFFF7EEE000000000
        x ^= x >> 43
        return x & (2 ** 64 - 1)


def test_mt19937_64_reference_value():
    g = MT19937_64(5489)
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042   # required by the C++ standard [rand.predef]


@pytest.mark.parametrize("periodic", [True, False])
def test_structured_mesh_definition(periodic):
    n, a, seed = 6, 0.15, 2
    x0, x1, y0, y1 = -10.0, 10.0, -10.0, 10.0
    mesh = E.structured_mesh(n, x0, x1, y0, y1, a, seed, periodic)
    assert mesh.K == 2 * n * n and mesh.Nv == (n + 1) ** 2
    g = MT19937_64(seed)
    h = (x1 - x0) / n
    VX = np.zeros(mesh.Nv)
    VY = np.zeros(mesh.Nv)
    for j in range(n + 1):
        for i in range(n + 1):
            x, y = x0 + i * h, y0 + j * h
            if i == n:
                x = x1
            if j == n:
                y = y1
            if 0 < i < n and 0 < j < n:
                x += (2.0 * (g() >> 11) * 2.0 ** -53 - 1.0) * a * h
                y += (2.0 * (g() >> 11) * 2.0 ** -53 - 1.0) * a * h
            VX[j * (n + 1) + i], VY[j * (n + 1) + i] = x, y
    assert np.array_equal(mesh.VX, VX) and np.array_equal(mesh.VY, VY)
    # element order: row-major quads, T0 = (v00, v10, v11), T1 = (v00, v11, v01)
    q = 2 * n + 3
    j, i = divmod(q, n)
    v00 = j * (n + 1) + i
    assert list(mesh.EToV[2 * q]) == [v00, v00 + 1, v00 + n + 2]
    assert list(mesh.EToV[2 * q + 1]) == [v00, v00 + n + 2, v00 + n + 1]
    xs, ys = mesh.VX[mesh.EToV], mesh.VY[mesh.EToV]
    area = (xs[:, 1] - xs[:, 0]) * (ys[:, 2] - ys[:, 0]) - (xs[:, 2] - xs[:, 0]) * (ys[:, 1] - ys[:, 0])
    assert (area > 0).all()
    if periodic:
        assert (mesh.face_tag == 0).all()
        assert mesh.vert_id[n] == 0 and mesh.vert_id[n * (n + 1)] == 0
    else:
        assert (mesh.face_tag > 0).sum() == 4 * n


def test_structured_mesh_rejects_bad_arguments():
    with pytest.raises(E.ApolloError):
        E.structured_mesh(2, 0, 1, 0, 1, 0.1, 0, True)      # periodic needs n >= 3
    with pytest.raises(E.ApolloError):
        E.structured_mesh(4, 0, 1, 0, 1, 0.6, 0, False)     # jitter must stay < 0.5


def test_cpp_host_interface_compiles(tmp_path):
    """apollo_amd/host/wx_nodal_dg2d_hip.hpp compiles and links against the C-ABI."""
    import subprocess
    exe = tmp_path / "demo"
    src = os.path.join(ROOT, "tests", "cpp", "subsolver_demo.cpp")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-Wall", "-Werror", "-o", str(exe), src,
                           "-L" + os.path.join(ROOT, "apollo_amd"), "-lapollo_dg"])
    assert exe.exists()


# --- per-frame VTU output ---------------------------------------------------------

def test_vtu_frame_layout(tmp_path):
    """The layout Apollo's readers expect (scripts/wxunsdgdata2.py:38-84):
    cell array 0 is "Rank", array j + 1 is dof j = node * m + variable, the
    cells' first three points are the element corners, and the nodes
    rebuilt from them with Apollo's affine map are the engine's nodes."""
    from apollo_amd import vtu
    from apollo_amd.gmsh import read_gmsh
    mesh = read_gmsh(os.path.join(GOLDEN, "meshes", "ffs_grid.msh"))
    N, m = 3, 5
    ops = O.Ops(N)
    x, y, _ = O.node_coords(ops, mesh.VX, mesh.VY, mesh.EToV)
    rng = np.random.default_rng(0)
    q = rng.standard_normal((mesh.K, ops.Np, m))
    path = tmp_path / vtu.frame_name("ffstep", 7)
    assert path.name == "ffstep_7.vtu"
    vtu.write_frame(str(path), mesh, q)
    pts, cells, arrays = vtu.read_frame(str(path))
    assert arrays[0][0] == "Rank" and len(arrays) == 1 + ops.Np * m
    assert np.array_equal(cells, mesh.EToV)
    for node in range(ops.Np):
        for v in range(m):
            assert np.array_equal(arrays[m * node + v + 1][1], q[:, node, v])
    p1, p2, p3 = pts[cells[:, 0]], pts[cells[:, 1]], pts[cells[:, 2]]
    r, s = ops.r[None, :], ops.s[None, :]
    X = 0.5 * (-p1[:, :1] * (r + s) + p2[:, :1] * (1. + r) + p3[:, :1] * (1. + s))
    Y = 0.5 * (-p1[:, 1:2] * (r + s) + p2[:, 1:2] * (1. + r) + p3[:, 1:2] * (1. + s))
    assert np.abs(X - x).max() < 1e-14 and np.abs(Y - y).max() < 1e-14


@pytest.mark.parametrize("name", ["advection_test.msh", "small_square_4bc.msh", "vortex_grid.msh",
                                  "tm_square.msh", "ffs_grid.msh"])
def test_cpp_gmsh_reader_matches_python(name, tmp_path):
    """The C++ host's Gmsh reader (apollo_amd/host/gmsh_v2.hpp) and the
    Python one give the same vertices, cones and face tags."""
    import subprocess
    from apollo_amd.gmsh import read_gmsh
    exe = os.path.join(ROOT, "tests", "cpp", "gmsh_dump")
    assert os.path.exists(exe), "build tests/cpp first (make -C tests/cpp)"
    path = os.path.join(GOLDEN, "meshes", name)
    out = tmp_path / "m.bin"
    assert subprocess.run([exe, path, str(out)], timeout=60).returncode == 0
    raw = np.fromfile(str(out), dtype=np.float64)
    K, Nv = int(raw[0]), int(raw[1])
    VX, VY = raw[2:2 + Nv], raw[2 + Nv:2 + 2 * Nv]
    cones = raw[2 + 2 * Nv:2 + 2 * Nv + 3 * K].astype(np.int64).reshape(K, 3)
    tags = raw[2 + 2 * Nv + 3 * K:2 + 2 * Nv + 6 * K].astype(np.int64).reshape(K, 3)
    m = read_gmsh(path)
    assert np.array_equal(VX, m.VX) and np.array_equal(VY, m.VY)
    assert np.array_equal(cones, m.EToV) and np.array_equal(tags, m.face_tag)
