"""GPU parity of Apollo's per-frame stepping (apl_ssp_advance: PETSc TSSSP
rks2 under ApSolver's dt control) against the oracle's restatement.

Bar: the same number of steps, each step's dt within 1e-12 relative (the
suggested dt comes from a maximum of FP64 face speeds), and the state
increment over the frame within 1e-11 relative L2 (FP64)."""
import numpy as np
import pytest

from apollo_amd import engine as E
from oracle import oracle as O

from cases import euler_case, from_oracle, maxwell_case, rel_l2, to_oracle

pytestmark = pytest.mark.gpu


def _frame(case, t0, t1, dt0, cfl, stages=5, q=None, engine_q=None):
    q0 = case.state() if q is None else q
    eng = E.Engine(case.mesh, case.params)
    qg = q0.copy() if engine_q is None else engine_q
    dts = eng.ssp_advance(t0, t1, dt0, cfl, qg, stages=stages)
    eng.close()
    qo, dto = O.ssp_advance(case.omesh, case.ophys, t0, t1, dt0, cfl,
                            to_oracle(case.params.eqn, q0), stages=stages)
    return q0, qg, dts, from_oracle(case.params.eqn, qo), dto


@pytest.mark.parametrize("name", ["euler_periodic", "euler_vortex_bc", "maxwell_pec_upwind",
                                  "euler_hll_vortex_bc", "euler5_hll_periodic"])
def test_ssp_frame_matches_oracle(name):
    if name == "euler_periodic":
        case, cfl = euler_case(n=8, N=3, m=4), 0.5
    elif name == "euler_hll_vortex_bc":
        case, cfl = euler_case(n=8, N=3, m=4, periodic=False, flux=E.FLUX_HLL), 0.5
    elif name == "euler5_hll_periodic":
        case, cfl = euler_case(n=8, N=3, m=5, flux=E.FLUX_HLL), 0.5
    elif name == "euler_vortex_bc":
        case, cfl = euler_case(n=8, N=3, m=4, periodic=False), 0.5
    else:
        case, cfl = maxwell_case(n=6, N=4, m=3, bc=E.BC_PEC, flux=E.FLUX_UPWIND), 0.5 / 9
    q0 = case.state()
    _, ms = case.oracle_rhs(q0)
    s0 = O.suggested_dt(case.omesh, cfl, ms)
    t0 = 0.05
    q0, qg, dts, qo, dto = _frame(case, t0, t0 + 9.5 * s0, 2.0 * s0, cfl, q=q0)
    print("%s: %d steps, dt %.6e .. %.6e" % (name, len(dts), dts[0], dts[-1]))
    assert len(dts) == len(dto) and len(dts) >= 9
    assert np.abs(dts - dto).max() <= 1e-12 * dto.max()
    assert abs(t0 + dts.sum() - (t0 + 9.5 * s0)) <= 1e-14
    err = rel_l2(qg - q0, qo - q0)
    print("%s: increment rel L2 vs oracle %.3e" % (name, err))
    assert err <= 1e-11


def test_ssp_roe_as_written_loses_positivity():
    """Apollo's Roe flux puts the y-momentum flux in the x-momentum slot and
    leaves the y-momentum slot at 0 (wxeulereqn.cc:1259-1260), so its surface
    term is inconsistent and a frame of the vortex ends in a positivity error
    in the oracle and in the engine alike."""
    case = euler_case(n=8, N=3, m=5, flux=E.FLUX_ROE)
    q0 = case.state()
    _, ms = case.oracle_rhs(q0)
    s0 = O.suggested_dt(case.omesh, 0.5, ms)
    with pytest.raises(ValueError):
        O.ssp_advance(case.omesh, case.ophys, 0.0, 3 * s0, s0, 0.5, to_oracle(case.params.eqn, q0))
    eng = E.Engine(case.mesh, case.params)
    with pytest.raises(E.ApolloError) as ei:
        eng.ssp_advance(0.0, 3 * s0, s0, 0.5, q0.copy())
    eng.close()
    assert ei.value.status in (8, 9)      # APL_ERR_NEG_DENSITY / APL_ERR_NEG_PRESSURE


def test_ssp_frame_device_pointer_and_stage_count():
    torch = pytest.importorskip("torch")
    case = euler_case(n=6, N=2, m=4)
    q0 = case.state()
    qd = torch.from_numpy(q0.copy()).cuda()
    _, _, dts, qo, dto = _frame(case, 0.0, 0.03, 0.004, 0.5, stages=3, q=q0, engine_q=qd)
    assert len(dts) == len(dto)
    assert rel_l2(qd.cpu().numpy() - q0, qo - q0) <= 1e-11


def test_ssp_frame_fp32_against_fp64():
    """FP32-MFMA contractions: the same dt sequence (speeds are FP64) and the
    state within 1e-5 relative of the FP64 run."""
    case = maxwell_case(n=6, N=5, m=3, bc=E.BC_PEC, flux=E.FLUX_UPWIND)
    q0 = case.state()
    p = case.params
    out = {}
    for prec in (E.PREC_FP64, E.PREC_FP32_MFMA):
        eng = E.Engine(case.mesh, E.Params(p.order, p.eqn, flux=p.flux, precision=prec, bcs=p.bcs))
        q = q0.copy()
        out[prec] = (q, eng.ssp_advance(0.0, 0.01, 0.002, 0.05, q))
        eng.close()
    assert np.array_equal(out[E.PREC_FP64][1], out[E.PREC_FP32_MFMA][1])
    assert rel_l2(out[E.PREC_FP32_MFMA][0] - q0, out[E.PREC_FP64][0] - q0) <= 1e-5


def test_ssp_frame_rejects_too_many_steps():
    case = euler_case(n=4, N=2, m=4)
    eng = E.Engine(case.mesh, case.params)
    q = case.state()
    with pytest.raises(E.ApolloError) as ei:
        eng.ssp_advance(0.0, 1.0, 0.01, 0.5, q, max_steps=3)
    assert ei.value.status == 1     # APL_ERR_ARG
