"""The vector-ALU cubature kernel (APL_ALGO_CUBATURE_VALU, N <= 3) against
the oracle, and against the MFMA cubature kernel on the same input.

Bars as for the MFMA kernel: FP64 within 1e-12 relative L2 of the oracle
(or 4x the oracle's own rounding floor, whichever is larger); FP32 within
2e-5.  Meshes are sized so that the last 64-element workgroup is partial
(K = 72, 200, 242, 16200, 33800); meshes of fewer than 512 groups run
eight-wave workgroups, larger ones four-wave workgroups."""
import numpy as np
import pytest

from apollo_amd import engine as E
from oracle import oracle as O

from cases import (Case, euler_case, ffs_case, ffs_state, fp64_parity, from_oracle, maxwell_case,
                   rel_l2, to_oracle)

pytestmark = pytest.mark.gpu

VALU, MFMA = E.ALGO_CUBATURE_VALU, E.ALGO_CUBATURE_MFMA


def _params(case, algorithm, precision=None):
    p = case.params
    return E.Params(p.order, p.eqn, flux=p.flux,
                    precision=p.precision if precision is None else precision,
                    gas_gamma=p.gas_gamma, c0=p.c0, chi=p.chi, gamma_ph=p.gamma_ph, bcs=p.bcs,
                    algorithm=algorithm)


def _rhs(case, q, algorithm, t=0.0, precision=None, cfl=None):
    eng = E.Engine(case.mesh, _params(case, algorithm, precision))
    out = np.zeros_like(q)
    dt = eng.rhs(t, q, out, cfl=1.0 if cfl is None else cfl)
    eng.close()
    return (out, dt) if cfl is not None else out


@pytest.mark.parametrize("N", [1, 2, 3])
@pytest.mark.parametrize("m", [4, 5])
def test_valu_euler_lf_fp64(N, m):
    """Periodic and with the isentropic-vortex boundary (time-dependent
    exterior state); K = 72 and 200 (a partial last workgroup)."""
    for n, periodic in ((6, True), (10, False)):
        case = euler_case(n=n, N=N, m=m, periodic=periodic)
        assert case.mesh.K % 64 != 0
        q = case.state()
        for t in ((0.0,) if periodic else (0.0, 0.37)):
            got = _rhs(case, q, VALU, t)
            fp64_parity(case, got, q, t)
            # the MFMA kernel on the same input: same result to rounding
            assert rel_l2(got, _rhs(case, q, MFMA, t)) <= 1e-13


@pytest.mark.parametrize("N", [1, 2, 3])
def test_valu_four_wave_groups_fp64(N):
    """K = 33800: 529 groups of 64 elements, run as four-wave workgroups
    (the small meshes above run eight-wave ones)."""
    case = euler_case(n=130, N=N, m=5, periodic=False)
    q = case.state()
    got = _rhs(case, q, VALU, 0.1)
    fp64_parity(case, got, q, 0.1)
    assert rel_l2(got, _rhs(case, q, MFMA, 0.1)) <= 1e-13


@pytest.mark.parametrize("flux", [E.FLUX_HLL, E.FLUX_ROE, E.FLUX_WAVE])
@pytest.mark.parametrize("N", [1, 3])
def test_valu_euler_riemann_fluxes_fp64(flux, N):
    for m, periodic in ((4, True), (5, False)):
        case = euler_case(n=8, N=N, m=m, periodic=periodic, flux=flux)
        q = case.state()
        fp64_parity(case, _rhs(case, q, VALU, 0.21), q, 0.21)


@pytest.mark.parametrize("flux", [E.FLUX_LF, E.FLUX_HLL])
@pytest.mark.parametrize("N", [1, 2, 3])
def test_valu_wall_inflow_outflow_fp64(flux, N):
    """Apollo's forward-facing-step mesh: inflow, walls and zero-gradient
    boundaries, several boundary faces per workgroup."""
    case = ffs_case(N=N, m=5, flux=flux, limiter=False)
    q = ffs_state(case, "smooth", seed=N)
    fp64_parity(case, _rhs(case, q, VALU), q)


@pytest.mark.parametrize("N", [1, 2, 3])
def test_valu_maxwell_cubature_fp64(N):
    """The cubature form of the linear Maxwell fluxes on the VALU kernel:
    TM (3 variables) with upwind flux and PEC walls, phMaxwell (8) with LF."""
    for m, bc, flux in ((3, E.BC_PEC, E.FLUX_UPWIND), (8, E.BC_MAXWELL_TM, E.FLUX_LF)):
        case = maxwell_case(n=6, N=N, m=m, bc=bc, flux=flux)
        q = case.state()
        fp64_parity(case, _rhs(case, q, VALU), q)


@pytest.mark.parametrize("N", [1, 2, 3])
def test_valu_fp32(N):
    """FP32 arithmetic on the VALU kernel: relative L2 against the FP64
    oracle <= 2e-5, as for FP32-MFMA."""
    for case in (euler_case(n=11, N=N, m=4), euler_case(n=8, N=N, m=5, periodic=False),
                 maxwell_case(n=6, N=N, m=3)):
        q = case.state()
        ref, _ = case.oracle_rhs(q)
        e_valu = rel_l2(_rhs(case, q, VALU, precision=E.PREC_FP32_MFMA), ref)
        e_mfma = rel_l2(_rhs(case, q, MFMA, precision=E.PREC_FP32_MFMA), ref)
        print("N=%d m=%d: FP32 VALU %.3e, FP32 MFMA %.3e" % (N, q.shape[-1], e_valu, e_mfma),
              flush=True)
        assert e_valu <= 2e-5


def test_valu_suggested_dt_matches_mfma():
    """The per-batch LF speed maxima feed the suggested dt: both kernels see
    the same fastest wave."""
    case = euler_case(n=10, N=3, m=4, periodic=False)
    q = case.state()
    _, dt_v = _rhs(case, q, VALU, cfl=0.5)
    _, dt_m = _rhs(case, q, MFMA, cfl=0.5)
    assert abs(dt_v - dt_m) <= 1e-14 * dt_m


@pytest.mark.parametrize("N,n", [(1, 130), (3, 90)])
def test_valu_lserk4_matches_oracle(N, n):
    """The fused LSERK4 stage (captured graph): K = 2 x 130^2 = 33800 runs
    529 four-wave workgroups, K = 2 x 90^2 = 16200 254 eight-wave ones
    (fewer than 512 groups); the last group holds 8 elements in both."""
    case = euler_case(n=n, N=N, m=4)
    assert case.mesh.K % 64 != 0
    q0 = case.state()
    eng = E.Engine(case.mesh, _params(case, VALU))
    dt = 0.002
    q = q0.copy()
    eng.lserk4(0.0, dt, 4, q)
    eng.close()
    with O.exact_geometry():
        ref = from_oracle(case.params.eqn, O.lserk4(case.omesh, case.ophys, 0.0, dt, 4,
                                                    to_oracle(case.params.eqn, q0)))
    err = rel_l2(q - q0, ref - q0)
    print("VALU LSERK4 N=%d K=%d: increment rel L2 %.3e" % (N, case.mesh.K, err), flush=True)
    assert err <= 1e-11


def test_valu_ssp_advance_matches_oracle():
    """Apollo's TSSSP rks2 with its dt control (the SSP combination epilogue)."""
    case = euler_case(n=12, N=2, m=5, periodic=False)
    q0 = case.state()
    cfl = 1.0 / (2 * 3 - 1)
    eng = E.Engine(case.mesh, _params(case, VALU))
    q = q0.copy()
    dts = eng.ssp_advance(0.0, 0.05, 0.01, cfl, q)
    eng.close()
    qo, dto = O.ssp_advance(case.omesh, case.ophys, 0.0, 0.05, 0.01, cfl,
                            to_oracle(case.params.eqn, q0), max_steps=1000)
    qo = from_oracle(case.params.eqn, qo)
    assert len(dts) == len(dto)
    assert np.abs(dts - dto).max() <= 1e-11 * dto.max()
    assert rel_l2(q - q0, qo - q0) <= 1e-11


def test_valu_error_flags_and_unsupported():
    case = euler_case(n=4, N=2, m=4)
    q = case.state()
    q[3, 2, 0] = -1.0
    eng = E.Engine(case.mesh, _params(case, VALU))
    with pytest.raises(E.ApolloError) as ei:
        eng.rhs(0.0, q, np.zeros_like(q))
    assert ei.value.status in (7, 8, 9)
    eng.close()
    for N, prec in ((4, E.PREC_FP64), (2, E.PREC_BF16X3)):
        c = euler_case(n=4, N=N, m=4)
        with pytest.raises(E.ApolloError) as ei:
            E.Engine(c.mesh, _params(c, VALU, prec))
        assert ei.value.status == 11   # APL_ERR_UNSUPPORTED


@pytest.mark.parametrize("n,periodic", [(2, False), (3, True), (3, False), (8, True)])
def test_valu_small_meshes_fp64(n, periodic):
    """A single partial workgroup (K = 8, 18; periodic meshes start at
    n = 3) and exactly two full ones (K = 128), FP64 against the oracle and
    the MFMA kernel."""
    for N in (1, 3):
        case = euler_case(n=n, N=N, m=5, periodic=periodic)
        q = case.state()
        got = _rhs(case, q, VALU, 0.2)
        fp64_parity(case, got, q, 0.2)
        assert rel_l2(got, _rhs(case, q, MFMA, 0.2)) <= 1e-13


# --- the nodal form of the TM fluxes on the vector ALU ---------------------

NVALU, NMFMA = E.ALGO_NODAL_VALU, E.ALGO_NODAL_MFMA


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6, 7, 8])
def test_nodal_valu_tm_fp64(N):
    """TM with PEC walls (upwind) and the TM boundary state (LF), K = 72 and
    200 (partial last workgroup): FP64 against the oracle and the MFMA nodal
    kernel."""
    for n, bc, flux in ((6, E.BC_PEC, E.FLUX_UPWIND), (10, E.BC_MAXWELL_TM, E.FLUX_LF)):
        case = maxwell_case(n=n, N=N, m=3, bc=bc, flux=flux)
        q = case.state()
        got = _rhs(case, q, NVALU)
        fp64_parity(case, got, q)
        assert rel_l2(got, _rhs(case, q, NMFMA)) <= 1e-13


@pytest.mark.parametrize("N", [1, 3, 5, 8])
def test_nodal_valu_tm_fp32(N):
    case = maxwell_case(n=8, N=N, m=3, bc=E.BC_PEC, flux=E.FLUX_UPWIND)
    q = case.state()
    ref, _ = case.oracle_rhs(q)
    e = rel_l2(_rhs(case, q, NVALU, precision=E.PREC_FP32_MFMA), ref)
    print("nodal VALU FP32 N=%d: %.3e" % (N, e), flush=True)
    assert e <= 2e-5


@pytest.mark.parametrize("N,n", [(2, 90), (5, 41)])
def test_nodal_valu_lserk4_matches_oracle(N, n):
    """Fused LSERK4 stages (captured graph) of the VALU nodal kernel against
    the oracle: K = 16200 (254 groups, 8 in the last) and 3362."""
    case = maxwell_case(n=n, N=N, m=3, bc=E.BC_PEC, flux=E.FLUX_UPWIND)
    x, y = case.omesh.x, case.omesh.y
    q0 = E.tm_cavity_mode(x, y, 0.013, m=3)
    q0 = np.ascontiguousarray(q0 + 1e-2 * np.random.default_rng(N).standard_normal(q0.shape))
    dt = 0.2 / n / (2 * N + 1)
    p = E.Params(N, E.EQN_MAXWELL_TM3, flux=E.FLUX_UPWIND, bcs=[(E.BC_PEC, [])], algorithm=NVALU)
    eng = E.Engine(case.mesh, p)
    q = q0.copy()
    eng.lserk4(0.0, dt, 3, q)
    eng.close()
    with O.exact_geometry():
        ref = from_oracle(case.params.eqn, O.lserk4(case.omesh, case.ophys, 0.0, dt, 3,
                                                    to_oracle(case.params.eqn, q0)))
    err = rel_l2(q - q0, ref - q0)
    print("nodal VALU LSERK4 N=%d K=%d: %.3e" % (N, case.mesh.K, err), flush=True)
    assert err <= 1e-11


def test_nodal_valu_ssp_and_dt_match_mfma():
    """Apollo's SSP stepping with its dt control: the same dt sequence and
    state as the MFMA nodal kernel."""
    case = maxwell_case(n=12, N=3, m=3, bc=E.BC_MAXWELL_TM, flux=E.FLUX_LF)
    q0 = case.state(kind="cavity")
    out = {}
    for algo in (NVALU, NMFMA):
        eng = E.Engine(case.mesh, _params(case, algo))
        q = q0.copy()
        dts = eng.ssp_advance(0.0, 0.05, 0.01, 0.5 / 7, q)
        eng.close()
        out[algo] = (q, dts)
    assert np.array_equal(out[NVALU][1], out[NMFMA][1])
    assert rel_l2(out[NVALU][0] - q0, out[NMFMA][0] - q0) <= 1e-12


def test_nodal_valu_unsupported():
    c = maxwell_case(n=4, N=3, m=8)
    with pytest.raises(E.ApolloError) as ei:
        E.Engine(c.mesh, _params(c, NVALU))
    assert ei.value.status == 11   # APL_ERR_UNSUPPORTED: phMaxwell
    e = euler_case(n=4, N=3, m=4)
    with pytest.raises(E.ApolloError) as ei:
        E.Engine(e.mesh, _params(e, NVALU))
    assert ei.value.status == 1    # APL_ERR_ARG: the nodal form needs a linear flux
