"""GPU discrete adjoint: adjoint RHS, reverse LSERK4 with checkpoints, the
disc objective and a torch.autograd.Function wrapper (harness only).

Pins: the oracle's CPU adjoint (itself pinned by dot-product and finite-
difference tests in test_oracle.py), the GPU dot-product test, and central
finite differences of the GPU forward model (eps = 1e-6)."""
import numpy as np
import pytest

from apollo_amd import engine as E
from oracle import oracle as O

from cases import euler_case, from_oracle, maxwell_case, rel_l2, to_oracle

pytestmark = pytest.mark.gpu

CASES = {
    "euler4_periodic": lambda: euler_case(n=6, N=3, m=4),
    "euler5_periodic": lambda: euler_case(n=5, N=2, m=5),
    "euler4_vortex_bc": lambda: euler_case(n=6, N=3, m=4, periodic=False),
    "tm3_lf": lambda: maxwell_case(n=5, N=4, m=3),
    "tm3_pec_upwind": lambda: maxwell_case(n=5, N=4, m=3, bc=E.BC_PEC, flux=E.FLUX_UPWIND),
    "ph8_pec": lambda: maxwell_case(n=4, N=3, m=8, bc=E.BC_PEC, chi=0.5, gph=0.3),
    "euler4_N8": lambda: euler_case(n=4, N=8, m=4),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_adjoint_rhs_matches_cpu_adjoint(name):
    case = CASES[name]()
    q = case.state()
    rng = np.random.default_rng(5)
    lam = rng.standard_normal(q.shape)
    eng = E.Engine(case.mesh, case.params)
    out = np.zeros_like(q)
    eng.adjoint_rhs(0.3, q, lam, out)
    with O.exact_geometry():
        ref = from_oracle(case.params.eqn, O.adjoint_rhs(case.omesh, case.ophys, 0.3,
                                                         to_oracle(case.params.eqn, q),
                                                         to_oracle(case.params.eqn, lam)))
    err = rel_l2(out, ref)
    print("%s: adjoint RHS rel L2 vs CPU adjoint %.3e" % (name, err))
    assert err <= 1e-12


def test_adjoint_dot_product_gpu():
    case = euler_case(n=6, N=4, m=4, periodic=False)
    q = case.state()
    rng = np.random.default_rng(9)
    v, w = rng.standard_normal(q.shape), rng.standard_normal(q.shape)
    eng = E.Engine(case.mesh, case.params)
    eps = 1e-6
    rp, rm, jtw = np.zeros_like(q), np.zeros_like(q), np.zeros_like(q)
    eng.rhs(0.1, q + eps * v, rp)
    eng.rhs(0.1, q - eps * v, rm)
    eng.adjoint_rhs(0.1, q, w, jtw)
    lhs, rhs = np.sum((rp - rm) / (2 * eps) * w), np.sum(v * jtw)
    assert abs(lhs - rhs) <= 1e-7 * abs(lhs), (lhs, rhs)


@pytest.mark.parametrize("ckpt", [1, 3, 100])
def test_reverse_lserk4_matches_cpu(ckpt):
    case = euler_case(n=5, N=3, m=4, periodic=False)
    q0 = case.state()
    rng = np.random.default_rng(2)
    gT = rng.standard_normal(q0.shape)
    steps, dt = 7, 0.01
    eng = E.Engine(case.mesh, case.params)
    g0, qT = np.zeros_like(q0), np.zeros_like(q0)
    eng.lserk4_adjoint(0.2, dt, steps, ckpt, q0, gT, g0, qT)
    with O.exact_geometry():
        ref = from_oracle(case.params.eqn, O.lserk4_adjoint(case.omesh, case.ophys, 0.2, dt, steps,
                                                            to_oracle(case.params.eqn, q0),
                                                            to_oracle(case.params.eqn, gT)))
        refT = from_oracle(case.params.eqn, O.lserk4(case.omesh, case.ophys, 0.2, dt, steps,
                                                     to_oracle(case.params.eqn, q0)))
    assert rel_l2(qT, refT) <= 1e-12
    err = rel_l2(g0, ref)
    print("ckpt %d: dJ/dq0 rel L2 vs CPU reverse LSERK4 %.3e" % (ckpt, err))
    assert err <= 1e-11


def test_reverse_lserk4_same_for_every_checkpoint_interval():
    """The checkpoint interval only decides which forward states are kept
    and which recomputed, from the same inputs by the same kernels: the
    gradient and q(T) are bit-identical for every interval, including
    ragged last segments (7 steps) and an interval longer than the run."""
    case = euler_case(n=5, N=3, m=4, periodic=False)
    q0 = case.state()
    gT = np.random.default_rng(5).standard_normal(q0.shape)
    eng = E.Engine(case.mesh, case.params)
    out = {}
    for ckpt in (1, 2, 3, 7, 100):
        g0, qT = np.zeros_like(q0), np.zeros_like(q0)
        eng.lserk4_adjoint(0.2, 0.01, 7, ckpt, q0, gT, g0, qT)
        out[ckpt] = (g0, qT)
    eng.close()
    for ckpt, (g0, qT) in out.items():
        assert np.array_equal(g0, out[1][0]), ckpt
        assert np.array_equal(qT, out[1][1]), ckpt


def test_objective_matches_oracle():
    case = euler_case(n=8, N=3, m=4)
    q = case.state()
    eng = E.Engine(case.mesh, case.params)
    g = np.zeros_like(q)
    J = eng.objective_disc_sq(q, 0, 0.3, -0.2, 1.0, g)
    Jr, gr = O.objective_disc_sq(case.omesh, to_oracle(case.params.eqn, q), 0, 0.3, -0.2, 1.0, 5)
    assert abs(J - Jr) <= 1e-13 * abs(Jr)
    assert rel_l2(g, from_oracle(case.params.eqn, gr)) <= 1e-13


class LSERK4Objective:
    """torch.autograd.Function (test harness only): J(q0) = disc objective of
    the density after `steps` LSERK4 steps, gradient by the engine's reverse
    LSERK4."""

    @staticmethod
    def make(torch, eng, t0, dt, steps, ckpt, disc):
        class Fn(torch.autograd.Function):
            @staticmethod
            def forward(ctx, q0):
                qT = q0.detach().clone()
                eng.lserk4(t0, dt, steps, qT)
                gT = torch.zeros_like(qT)
                J = eng.objective_disc_sq(qT, 0, disc[0], disc[1], disc[2], gT)
                ctx.save_for_backward(q0.detach().clone(), gT)
                return torch.tensor(J, dtype=q0.dtype, device=q0.device)

            @staticmethod
            def backward(ctx, gJ):
                q0, gT = ctx.saved_tensors
                g0 = torch.zeros_like(q0)
                eng.lserk4_adjoint(t0, dt, steps, ckpt, q0, gT.contiguous(), g0)
                return gJ * g0

        return Fn


def test_autograd_function_against_finite_differences():
    torch = pytest.importorskip("torch")
    case = euler_case(n=8, N=3, m=4)
    eng = E.Engine(case.mesh, case.params)
    q0 = torch.from_numpy(case.state()).cuda()
    steps, dt, disc = 12, 0.01, (0.0, 0.0, 1.0)
    Fn = LSERK4Objective.make(torch, eng, 0.0, dt, steps, 5, disc)
    q = q0.clone().requires_grad_(True)
    J = Fn.apply(q)
    J.backward()
    g = q.grad.detach()
    gen = torch.Generator(device="cpu").manual_seed(3)
    eps = 1e-6
    for _ in range(4):
        d = torch.randn(q0.shape, generator=gen, dtype=torch.float64).cuda()
        with torch.no_grad():
            Jp = Fn.apply(q0 + eps * d).item()
            Jm = Fn.apply(q0 - eps * d).item()
        fd = (Jp - Jm) / (2 * eps)
        ad = float((g * d).sum())
        print("directional derivative: FD %.12e adjoint %.12e rel %.2e" % (fd, ad, abs(fd - ad) / abs(ad)))
        assert abs(fd - ad) <= 1e-6 * abs(ad)


def test_adjoint_rejects_riemann_fluxes():
    """The discrete adjoint covers LF and the TM upwind flux; HLL, Roe and
    Wave handles refuse it with APL_ERR_UNSUPPORTED rather than returning a
    wrong adjoint."""
    for flux in (E.FLUX_HLL, E.FLUX_ROE, E.FLUX_WAVE):
        case = euler_case(n=4, N=2, m=4, flux=flux)
        q = case.state()
        eng = E.Engine(case.mesh, case.params)
        with pytest.raises(E.ApolloError) as ei:
            eng.adjoint_rhs(0.0, q, q.copy(), np.zeros_like(q))
        assert ei.value.status == 11
        eng.close()
