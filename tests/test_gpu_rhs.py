"""GPU parity of the RHS and the stepper against the CPU oracle.

Bar: FP64 fields within 1e-12 relative L2 of the oracle; integer maps
bit-exact.  The FP32-MFMA variant is checked against the oracle with the
tolerance stated in each test."""
import numpy as np
import pytest

from apollo_amd import engine as E
from oracle import oracle as O

from cases import (Case, euler_case, ffs_case, ffs_state, fp64_parity, maxwell_case, rel_l2,
                   vortex_bc, walled_unit_square)

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def _rhs(case, q, t=0.0, precision=None, algorithm=None):
    params = case.params
    if precision is not None or algorithm is not None:
        params = E.Params(params.order, params.eqn, flux=params.flux,
                          precision=params.precision if precision is None else precision,
                          gas_gamma=params.gas_gamma, c0=params.c0, chi=params.chi,
                          gamma_ph=params.gamma_ph, bcs=params.bcs,
                          algorithm=params.algorithm if algorithm is None else algorithm)
    eng = E.Engine(case.mesh, params)
    out = np.zeros_like(q)
    eng.rhs(t, q, out)
    eng.close()
    return out


@pytest.mark.parametrize("m", [4, 5])
def test_euler_periodic_rhs_fp64(m):
    case = euler_case(n=16, N=3, m=m)
    q = case.state()
    fp64_parity(case, _rhs(case, q), q)


def test_euler_vortex_bc_rhs_fp64():
    case = euler_case(n=10, N=4, m=4, periodic=False)
    q = case.state()
    for t in (0.0, 0.37):
        fp64_parity(case, _rhs(case, q, t), q, t)


@pytest.mark.parametrize("bc", [E.BC_MAXWELL_TM, E.BC_PEC])
@pytest.mark.parametrize("m", [3, 8])
def test_maxwell_rhs_fp64(bc, m):
    case = maxwell_case(n=10, N=5, m=m, bc=bc)
    q = case.state()
    fp64_parity(case, _rhs(case, q), q)


@pytest.mark.parametrize("N", [2, 5, 8])
@pytest.mark.parametrize("m", [3, 8])
def test_maxwell_cubature_and_nodal_forms_agree(N, m):
    """Maxwell runs the nodal linear-flux kernel by default; Apollo's cubature
    form (APL_ALGO_CUBATURE) must give the same RHS, and both match the oracle."""
    for bc, flux in ((E.BC_MAXWELL_TM, E.FLUX_LF), (E.BC_PEC, E.FLUX_UPWIND if m == 3 else E.FLUX_LF)):
        case = maxwell_case(n=6, N=N, m=m, bc=bc, flux=flux)
        q = case.state()
        nodal = _rhs(case, q)
        cub = _rhs(case, q, algorithm=E.ALGO_CUBATURE)
        fp64_parity(case, nodal, q)
        fp64_parity(case, cub, q)
        assert rel_l2(nodal, cub) <= 1e-12


def test_phmaxwell_chi_gamma_fp64():
    case = maxwell_case(n=8, N=4, m=8, bc=E.BC_PEC, chi=1.3, gph=0.7)
    q = case.state()
    fp64_parity(case, _rhs(case, q), q)


def test_tm_upwind_fp64():
    case = maxwell_case(n=10, N=5, m=3, bc=E.BC_PEC, flux=E.FLUX_UPWIND)
    q = case.state()
    fp64_parity(case, _rhs(case, q), q)


@pytest.mark.parametrize("flux", [E.FLUX_HLL, E.FLUX_ROE, E.FLUX_WAVE])
@pytest.mark.parametrize("m", [4, 5])
@pytest.mark.parametrize("N", [1, 4, 8])
def test_euler_riemann_fluxes_fp64(flux, m, N):
    """Apollo's HLL, Roe and "Wave" fluxes (Roe with its back-rotation as
    written, Wave as LF of the face-frame traces), periodic and with the
    isentropic-vortex boundary."""
    for periodic in (True, False):
        case = euler_case(n=6 if N == 8 else 8, N=N, m=m, periodic=periodic, flux=flux)
        q = case.state()
        fp64_parity(case, _rhs(case, q, 0.21), q, 0.21)


@pytest.mark.parametrize("flux", [E.FLUX_HLL, E.FLUX_ROE, E.FLUX_WAVE])
def test_euler_riemann_fluxes_reduced_precision(flux):
    case = euler_case(n=8, N=4, m=4, flux=flux)
    q = case.state()
    ref, _ = case.oracle_rhs(q)
    e32 = rel_l2(_rhs(case, q, precision=E.PREC_FP32_MFMA), ref)
    e3 = rel_l2(_rhs(case, q, precision=E.PREC_BF16X3), ref)
    print("flux %d: FP32 %.3e BF16x3 %.3e" % (flux, e32, e3))
    assert e32 <= 2e-5 and e3 <= 1e-3


@pytest.mark.parametrize("flux", [E.FLUX_LF, E.FLUX_HLL])
@pytest.mark.parametrize("N", [1, 3, 6])
def test_euler_wall_inflow_outflow_rhs_fp64(flux, N):
    """Apollo's forward-facing-step mesh with its inflow, wall and
    zero-gradient boundaries, and a Dirichlet-walled square."""
    for m in (4, 5):
        case = ffs_case(N=N, m=m, flux=flux, limiter=False)
        q = ffs_state(case, "smooth", seed=N)
        fp64_parity(case, _rhs(case, q), q)
    sq = euler_case(n=6, N=N, m=5, periodic=False, flux=flux)
    case = Case(sq.mesh, E.Params(N, E.EQN_EULER5, flux=flux,
                                  bcs=[(E.BC_EULER_DIRICHLET, [1.0, 0.5, 0.1, 0.05, 1.4, 1.0])]))
    q = sq.state()
    fp64_parity(case, _rhs(case, q), q)


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6, 7, 8])
def test_all_orders_fp64(N):
    for case in (euler_case(n=6, N=N, m=4), maxwell_case(n=5, N=N, m=3)):
        q = case.state()
        fp64_parity(case, _rhs(case, q), q)


@pytest.mark.parametrize("N", [3, 5, 8])
def test_fp32_mfma_rhs(N):
    """FP32-MFMA contractions: relative L2 against the FP64 oracle <= 2e-5."""
    for case in (euler_case(n=6, N=N, m=4), maxwell_case(n=5, N=N, m=3)):
        q = case.state()
        ref, _ = case.oracle_rhs(q)
        err = rel_l2(_rhs(case, q, precision=E.PREC_FP32_MFMA), ref)
        print("FP32 N=%d %s rel L2 %.3e" % (N, case.name, err))
        assert err <= 2e-5


BF16X3_TOL = 1e-3


@pytest.mark.parametrize("N", [1, 3, 4, 5, 8])
def test_bf16x3_rhs(N):
    """Split-BF16x3 contractions: each operand is split as x = hi + lo with
    hi = bf16(x), lo = bf16(x - hi), so |x - hi - lo| <= 2^-18 |x|, and the
    dropped lo.lo term is <= 2^-18 |ab|: operand rounding is 2^6 times that
    of FP32 (2^-24).  The RHS error carries the operators' cancellation
    factor in both variants, so the bar is stated against the FP32-MFMA
    error of the same case: BF16x3 <= 2^8 x FP32 (one BF16 product alone
    would be ~2^15 x), and <= 1e-3 relative L2 against the FP64 oracle."""
    cases = (euler_case(n=6, N=N, m=4), euler_case(n=5, N=N, m=4, periodic=False),
             maxwell_case(n=5, N=N, m=3, bc=E.BC_PEC, flux=E.FLUX_UPWIND))
    for case in cases:
        q = case.state()
        ref, _ = case.oracle_rhs(q)
        algos = [None] if case.name == "euler" else [E.ALGO_AUTO, E.ALGO_CUBATURE]
        for algo in algos:
            e32 = rel_l2(_rhs(case, q, precision=E.PREC_FP32_MFMA, algorithm=algo), ref)
            ebf = rel_l2(_rhs(case, q, precision=E.PREC_BF16X3, algorithm=algo), ref)
            print("N=%d %s algo=%s rel L2: FP32 %.3e BF16x3 %.3e (ratio %.0f)"
                  % (N, case.name, algo, e32, ebf, ebf / e32))
            assert ebf <= BF16X3_TOL
            assert ebf <= 2.0 ** 8 * e32


def test_bf16x3_lserk4_graph():
    """LSERK4 steps with BF16x3 contractions against the FP64 engine."""
    for case in (euler_case(n=10, N=4, m=4), maxwell_case(n=8, N=5, m=3, bc=E.BC_PEC)):
        q0 = case.state()
        out = {}
        for prec in (E.PREC_FP64, E.PREC_BF16X3):
            p = case.params
            eng = E.Engine(case.mesh, E.Params(p.order, p.eqn, flux=p.flux, precision=prec,
                                               bcs=p.bcs))
            q = q0.copy()
            eng.lserk4(0.0, 0.002, 10, q)
            out[prec] = q
            eng.close()
        err = rel_l2(out[E.PREC_BF16X3] - q0, out[E.PREC_FP64] - q0)
        print("BF16x3 %s LSERK4 increment rel L2 vs FP64 %.3e" % (case.name, err))
        assert err <= BF16X3_TOL


def test_ragged_gmsh_mesh():
    """250-triangle Gmsh mesh from Apollo's vortex deck: K not a multiple of 16."""
    from apollo_amd.gmsh import read_gmsh
    import os
    from cases import GOLDEN
    mesh = read_gmsh(os.path.join(GOLDEN, "meshes", "vortex_grid.msh"))
    assert mesh.K % 16 != 0
    params = E.Params(3, E.EQN_EULER5, bcs=[vortex_bc(x0=5.0)])
    case = Case(mesh, params)
    x, y = case.omesh.x, case.omesh.y
    q = np.ascontiguousarray(E.vortex_state(x, y, 0.0, x0=5.0, m=5))
    ref, ms = case.oracle_rhs(q, 0.1)
    eng = E.Engine(mesh, params)
    out = np.zeros_like(q)
    dt = eng.rhs(0.1, q, out, cfl=0.2)
    fp64_parity(case, out, q, 0.1)
    assert abs(dt - O.suggested_dt(case.omesh, 0.2, ms)) <= 1e-14 * dt


def test_index_maps_bit_exact():
    case = maxwell_case(n=7, N=4, m=3)
    eng = E.Engine(case.mesh, case.params)
    et, fm, vp = eng.index_maps()
    assert np.array_equal(et, case.omesh.etetf)
    assert np.array_equal(fm, case.ops.fmask)
    # vmapP: own face node i coincides with the neighbour node it names
    g, x, y = eng.geometry()
    K, Nfp = case.mesh.K, case.ops.Nfp
    xf, yf = x.ravel(), y.ravel()
    for k in range(K):
        for f in range(3):
            for i in range(Nfp):
                j = vp[k, f, i]
                own = k * case.ops.Np + fm[f, i]
                if et[k, 2 * f + 1] < 0:
                    assert j == -1
                else:
                    assert abs(xf[j] - xf[own]) < 1e-12 and abs(yf[j] - yf[own]) < 1e-12


def test_lserk4_graph_matches_oracle():
    case = euler_case(n=12, N=3, m=4)
    q0 = case.state()
    eng = E.Engine(case.mesh, case.params)
    dt = 0.01
    q = q0.copy()
    eng.lserk4(0.0, dt, 7, q)
    ref = O.lserk4(case.omesh, case.ophys, 0.0, dt, 7, __import__("cases").to_oracle(case.params.eqn, q0))
    ref = __import__("cases").from_oracle(case.params.eqn, ref)
    err = rel_l2(q - q0, ref - q0)
    print("lserk4 increment rel L2 %.3e" % err)
    assert err <= 1e-11
    assert rel_l2(q, ref) <= 1e-13


@pytest.mark.parametrize("N,n,flux,steps", [(1, 41, E.FLUX_LF, 3), (2, 41, E.FLUX_UPWIND, 3),
                                             (3, 41, E.FLUX_UPWIND, 3), (4, 41, E.FLUX_LF, 3),
                                             (5, 41, E.FLUX_UPWIND, 3), (5, 161, E.FLUX_UPWIND, 2)])
def test_nodal_lserk4_matches_oracle(N, n, flux, steps):
    """LSERK4 steps of the nodal form (Maxwell TM, PEC walls) against the
    oracle's LSERK4:
    K = 2 n^2 = 3362 and 51842, both with a partial last batch (2 elements),
    the larger with several batches per wave.  FP64 increment within 1e-11;
    FP32 within 2e-5 of the FP64 increment (the single-RHS FP32 bar)."""
    from cases import from_oracle, to_oracle
    case = maxwell_case(n=n, N=N, m=3, bc=E.BC_PEC, flux=flux)
    assert case.mesh.K % 16 == 2
    x, y = case.omesh.x, case.omesh.y
    # the cavity mode with 1% white noise: a smooth state alone has a tiny
    # increment on the fine mesh, whose rounding floor would set the bar
    q0 = E.tm_cavity_mode(x, y, 0.013, m=3)
    q0 = np.ascontiguousarray(q0 + 1e-2 * np.random.default_rng(N).standard_normal(q0.shape))
    dt = 0.2 / n / (2 * N + 1)
    print("nodal LSERK4: N=%d K=%d" % (N, case.mesh.K), flush=True)
    # against the oracle with the elements' exact affine geometric factors,
    # as fp64_parity does (Apollo recomputes them at the cubature points, a
    # rounding of its own of the same size as the bar at N = 5)
    with O.exact_geometry():
        ref = from_oracle(case.params.eqn, O.lserk4(case.omesh, case.ophys, 0.0, dt, steps,
                                                    to_oracle(case.params.eqn, q0)))
    out = {}
    for prec in (E.PREC_FP64, E.PREC_FP32_MFMA):
        p = E.Params(N, E.EQN_MAXWELL_TM3, flux=flux, precision=prec, bcs=[(E.BC_PEC, [])])
        eng = E.Engine(case.mesh, p)
        q = q0.copy()
        eng.lserk4(0.0, dt, steps, q)
        eng.close()
        out[prec] = q
    e64 = rel_l2(out[E.PREC_FP64] - q0, ref - q0)
    e32 = rel_l2(out[E.PREC_FP32_MFMA] - q0, out[E.PREC_FP64] - q0)
    print("  increment rel L2: FP64 vs oracle %.3e, FP32 vs FP64 %.3e" % (e64, e32), flush=True)
    assert e64 <= 1e-11
    assert e32 <= 2e-5


def test_lserk4_device_pointer_and_time_dependent_bc():
    torch = pytest.importorskip("torch")
    case = euler_case(n=8, N=3, m=4, periodic=False)
    q0 = case.state()
    eng = E.Engine(case.mesh, case.params)
    qd = torch.from_numpy(q0.copy()).cuda()
    eng.lserk4(0.25, 0.005, 6, qd)
    from cases import to_oracle, from_oracle
    ref = from_oracle(case.params.eqn, O.lserk4(case.omesh, case.ophys, 0.25, 0.005, 6,
                                                to_oracle(case.params.eqn, q0)))
    assert rel_l2(qd.cpu().numpy() - q0, ref - q0) <= 1e-11


def test_rhs_error_codes():
    """Apollo exits on rho <= 0 (wxeulereqn.cc:838-845); the engine returns a status."""
    case = euler_case(n=4, N=2, m=4)
    q = case.state()
    q[3, 2, 0] = -1.0
    eng = E.Engine(case.mesh, case.params)
    with pytest.raises(E.ApolloError) as ei:
        eng.rhs(0.0, q, np.zeros_like(q))
    assert ei.value.status in (8, 9, 7)


def test_cpp_subsolver_interface(tmp_path):
    """The Apollo-shaped C++ interface (apollo_amd/host) driven from C++:
    setup/init/step/advance/solve on the Maxwell TM cavity, checked against
    the oracle and the Python binding."""
    import os
    import subprocess
    from cases import Case, walled_unit_square
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "subsolver_demo")
    assert os.path.exists(exe), "build tests/cpp first (make -C tests/cpp)"
    out = tmp_path / "demo.bin"
    n, N, steps = 6, 4, 5
    r = subprocess.run([exe, str(n), str(N), str(steps), str(out)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    raw = np.fromfile(str(out), dtype=np.float64)
    dt0, dt, length = raw[0], raw[1], int(raw[2])
    q0, rhs, qT, qF = (raw[3 + i * length:3 + (i + 1) * length] for i in range(4))
    case = Case(walled_unit_square(n), E.Params(N, E.EQN_MAXWELL_TM3, bcs=[(E.BC_MAXWELL_TM, [])]))
    q0 = q0.reshape(case.shape)
    fp64_parity(case, rhs.reshape(case.shape), q0)
    cfl = 0.5 / (2 * N + 1)
    assert abs(dt0 - O.initial_dt(case.omesh, case.ophys, __import__("cases").to_oracle(case.params.eqn, q0), cfl)) <= 1e-14 * dt0
    assert abs(dt - O.suggested_dt(case.omesh, cfl, 1.0)) <= 1e-14 * dt
    eng = E.Engine(case.mesh, case.params)
    q = q0.copy()
    eng.lserk4(0.0, dt, steps, q)
    assert rel_l2(qT.reshape(case.shape), q) <= 1e-14
    # solve(): two frames of Apollo's stepping, each restarted at dt0
    from cases import to_oracle, from_oracle
    qo = to_oracle(case.params.eqn, q0)
    tsize = 2 * steps * dt / 2
    for i in range(2):
        qo, _ = O.ssp_advance(case.omesh, case.ophys, i * tsize, (i + 1) * tsize, dt0, cfl, qo)
    qo = from_oracle(case.params.eqn, qo)
    assert rel_l2(qF.reshape(case.shape) - q0, qo - q0) <= 1e-11
