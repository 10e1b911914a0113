"""Parity at the full size of each of the five benchmark workloads
(apollo_amd/workloads.py): the exact mesh (n, domain, jitter, seed), order,
equation, flux and boundary condition of each, on the GPU against the CPU
oracle (Apollo's RHS restated, src/subsolvers/wxnodaldg2dmethod.cc:320-510).

Bars:
  - index maps (ETETF, Fmask) bit-exact; vmapP geometric (coinciding nodes);
  - FP64 RHS: tests/cases.py fp64_parity against the exact-geometry oracle:
    1e-12 relative L2 on a perturbed state, and max(1e-12, 4 x the FP64
    rounding floor) on the workload's own smooth state, whose RHS cancels
    (the floor is measured and printed; 7e-12 on the 512x512 N=4 vortex);
  - FP32-MFMA RHS <= 2e-5 and split-BF16x3 RHS <= 1e-3 relative L2 against
    the FP64 oracle on the perturbed state (the single-RHS bars of
    test_gpu_rhs.py), FP32 within 2 x 2^29 x the FP64 floor on the smooth
    state;
  - stepped runs: FP64 LSERK4 against the oracle's LSERK4 (increment
    <= 1e-11); FP32-MFMA within 1e-5 relative L2 of the FP64 engine run (the
    error bar the FP32 contractions are held to); split-BF16x3 within 1e-4 (it misses 1e-5 on the
    N=4 Euler and N=8 Maxwell runs: 1.5e-5 and 2.4e-5 measured);
  - adjoint: the adjoint RHS and the reverse LSERK4 against the oracle's CPU
    adjoint (<= 1e-12 / 1e-11), and dJ/dq0 against central differences.
"""
import numpy as np
import pytest

from apollo_amd import engine as E
from apollo_amd import workloads as W
from oracle import oracle as O

from cases import Case, fp64_parity, from_oracle, rel_l2, rounding_floor, to_oracle

pytestmark = pytest.mark.gpu


def _case(name, precision=E.PREC_FP64, algorithm=E.ALGO_AUTO):
    w = W.get(name)
    return w, Case(w.mesh(), w.params(precision, algorithm), name)


def _rhs_state(w, case):
    """A state with every variable live: the workload's initial state, the
    cavity mode at t = 0.03 (B != 0), or the 16-mode Ez with random smooth
    Bx, By."""
    x, y = case.omesh.x, case.omesh.y
    if w.state_kind == "cavity":
        return np.ascontiguousarray(E.tm_cavity_mode(x, y, 0.03, m=3))
    q = np.array(w.initial_state(x, y))
    if w.state_kind == "fourier":
        q[..., 1] = E.random_fourier_ez(x, y, 16, 5, m=3)[..., 0]
        q[..., 2] = E.random_fourier_ez(x, y, 16, 6, m=3)[..., 0]
    return np.ascontiguousarray(q)


def _engine_rhs(w, precision, algorithm, q, t=0.0, cfl=None):
    eng = E.Engine(w.mesh(), w.params(precision, algorithm))
    out = np.zeros_like(q)
    dt = eng.rhs(t, q, out, cfl=w.cfl if cfl is None else cfl)
    eng.close()
    return out, dt


@pytest.mark.parametrize("name", W.NAMES)
def test_index_maps_full_size(name):
    w, case = _case(name)
    eng = E.Engine(case.mesh, case.params)
    et, fm, vp = eng.index_maps()
    g, x, y = eng.geometry()
    eng.close()
    print("%s: K=%d" % (name, case.mesh.K), flush=True)
    assert et.shape == (w.K, 6)
    assert np.array_equal(et, case.omesh.etetf)
    assert np.array_equal(fm, case.ops.fmask)
    # node coordinates as the oracle computes them
    assert np.abs(x - case.omesh.x).max() <= 1e-14 * w.L
    assert np.abs(y - case.omesh.y).max() <= 1e-14 * w.L
    # vmapP: own face node i of face f coincides with the neighbour node it
    # names (up to the period on periodic meshes); -1 on boundary faces
    Np = case.ops.Np
    own = np.arange(w.K)[:, None, None] * Np + fm[None, :, :]      # [K][3][Nfp]
    bnd = et[:, 1::2] < 0                                            # [K][3]
    assert np.all(vp[bnd] == -1)
    inner = ~bnd
    j, o = vp[inner].ravel(), own[inner].ravel()
    assert j.min() >= 0 and j.max() < w.K * Np
    dx = x.ravel()[j] - x.ravel()[o]
    dy = y.ravel()[j] - y.ravel()[o]
    if w.periodic:
        dx -= w.L * np.round(dx / w.L)
        dy -= w.L * np.round(dy / w.L)
    assert np.abs(dx).max() <= 1e-12 and np.abs(dy).max() <= 1e-12
    # every interior face is paired both ways
    k = np.repeat(np.arange(w.K), 3).reshape(w.K, 3)[inner]
    f = np.tile(np.arange(3), w.K).reshape(w.K, 3)[inner]
    e2, f2 = et[:, 0::2][inner], et[:, 1::2][inner]
    assert np.array_equal(et[e2, 2 * f2], k) and np.array_equal(et[e2, 2 * f2 + 1], f)


def _perturbed(w, q):
    """The workload state with 1% white noise: multiplicative for Euler
    (keeps rho and p positive), additive N(0, 0.01 max|q|) for Maxwell.
    Its RHS is of the order of its face jumps everywhere, so the RHS does
    not cancel to near zero as a smooth state's does in its far field."""
    rng = np.random.default_rng(31)
    if w.linear:
        return np.ascontiguousarray(q + 1e-2 * np.abs(q).max() * rng.standard_normal(q.shape))
    return np.ascontiguousarray(q * (1.0 + 1e-2 * rng.standard_normal(q.shape)))


# unit roundoff of FP32 over that of FP64
FP32_OVER_FP64 = 2.0 ** 29


@pytest.mark.parametrize("name", W.NAMES)
def test_rhs_full_size(name):
    """FP64, FP32-MFMA and BF16x3 RHS against the FP64 oracle at the
    workload's full size (both forms for Maxwell), and the suggested dt, on
    two states:
      - the workload's own smooth state.  Its RHS is a cancellation of
        volume and surface terms of order |F| / h, so its rounding floor
        (cases.rounding_floor) grows as h shrinks.  Bars: FP64
        max(1e-12, 4 x floor); FP32 2 x 2^29 x floor (the same floor at
        FP32's unit roundoff); BF16x3 printed;
      - that state with 1% white noise, whose RHS does not cancel.  Bars:
        FP64 1e-12, FP32 2e-5, BF16x3 1e-3 (the single-RHS bars of
        test_gpu_rhs.py)."""
    w, case = _case(name)
    t = 0.0
    q = _rhs_state(w, case)
    qn = _perturbed(w, q)
    floor = rounding_floor(case, q, t)
    ref, ms = case.oracle_rhs(q, t)
    refn, _ = case.oracle_rhs(qn, t)
    dt_ref = O.suggested_dt(case.omesh, w.cfl, ms)
    print("%s: FP64 rounding floor of the workload state %.3e" % (name, floor), flush=True)
    algos = [E.ALGO_AUTO, E.ALGO_CUBATURE] if w.linear else [E.ALGO_AUTO]
    for algo in algos:
        got, dt = _engine_rhs(w, E.PREC_FP64, algo, q, t)
        print("%s algo=%d FP64, workload state:" % (name, algo), flush=True)
        fp64_parity(case, got, q, t, floor=floor)
        assert abs(dt - dt_ref) <= 1e-14 * dt_ref
        got, _ = _engine_rhs(w, E.PREC_FP64, algo, qn, t)
        print("%s algo=%d FP64, perturbed state:" % (name, algo), flush=True)
        fp64_parity(case, got, qn, t)
        for prec, tol in ((E.PREC_FP32_MFMA, 2e-5), (E.PREC_BF16X3, 1e-3)):
            err = rel_l2(_engine_rhs(w, prec, algo, q, t)[0], ref)
            errn = rel_l2(_engine_rhs(w, prec, algo, qn, t)[0], refn)
            print("%s algo=%d precision=%d: rel L2 vs FP64 oracle %.3e (workload state), "
                  "%.3e (perturbed)" % (name, algo, prec, err, errn), flush=True)
            if prec == E.PREC_FP32_MFMA:
                assert err <= 2.0 * FP32_OVER_FP64 * floor
            assert errn <= tol


def _run(w, precision, steps, q0, algorithm=E.ALGO_AUTO):
    """`steps` LSERK4 steps at the workload's fixed dt, on device memory."""
    import torch
    eng = E.Engine(w.mesh(), w.params(precision, algorithm))
    dt = w.fixed_dt(eng, q0)
    qd = torch.from_numpy(q0.copy()).cuda()
    eng.lserk4(0.0, dt, steps, qd)
    torch.cuda.synchronize()
    eng.close()
    return qd.cpu().numpy(), dt


def test_vortex_64_N3_lserk4_against_oracle():
    """100 FP64 LSERK4 steps (CFL 0.5) against the oracle's LSERK4 with the
    same dt; the vortex has drifted by 100 dt."""
    w, case = _case("vortex_64_N3")
    q0 = np.ascontiguousarray(w.initial_state(case.omesh.x, case.omesh.y))
    qT, dt = _run(w, E.PREC_FP64, w.steps, q0)
    _, ms = case.oracle_rhs(q0)
    assert abs(dt - O.suggested_dt(case.omesh, w.cfl, ms)) <= 1e-14 * dt
    with O.exact_geometry():
        ref = from_oracle(w.eqn, O.lserk4(case.omesh, case.ophys, 0.0, dt, w.steps,
                                          to_oracle(w.eqn, q0)))
    err_inc, err = rel_l2(qT - q0, ref - q0), rel_l2(qT, ref)
    ex = w.exact_state(case.omesh.x, case.omesh.y, w.steps * dt)
    print("vortex_64_N3: dt %.6e, %d steps: increment rel L2 vs oracle %.3e, state %.3e, "
          "vs analytic vortex %.3e" % (dt, w.steps, err_inc, err, rel_l2(qT, ex)), flush=True)
    assert err_inc <= 1e-11 and err <= 1e-12


def test_tm_cavity_256_N5_500_steps():
    """500 steps of FP64 and FP32-MFMA (nodal form) and FP64 in Apollo's
    cubature form against the analytic cavity mode and each other."""
    w, case = _case("tm_cavity_256_N5")
    x, y = case.omesh.x, case.omesh.y
    q0 = np.ascontiguousarray(w.initial_state(x, y))
    runs = {}
    for key, prec, algo in (("fp64", E.PREC_FP64, E.ALGO_AUTO),
                            ("fp64_cub", E.PREC_FP64, E.ALGO_CUBATURE),
                            ("fp32", E.PREC_FP32_MFMA, E.ALGO_AUTO),
                            ("bf16x3", E.PREC_BF16X3, E.ALGO_AUTO)):
        runs[key], dt = _run(w, prec, w.steps, q0, algo)
    ex = w.exact_state(x, y, w.steps * dt)
    errs = {k: rel_l2(v, ex) for k, v in runs.items()}
    print("tm_cavity_256_N5, T = %d dt: rel L2 vs analytic %s" % (w.steps, errs), flush=True)
    e32 = rel_l2(runs["fp32"], runs["fp64"])
    ebf = rel_l2(runs["bf16x3"], runs["fp64"])
    ecub = rel_l2(runs["fp64_cub"], runs["fp64"])
    print("  FP32 vs FP64 %.3e, BF16x3 vs FP64 %.3e, cubature vs nodal FP64 %.3e" % (e32, ebf, ecub))
    assert errs["fp64"] <= 1e-10 and errs["fp64_cub"] <= 1e-10
    assert ecub <= 1e-12
    assert e32 <= 1e-5 and errs["fp32"] <= 1e-5
    assert ebf <= 1e-4


def test_vortex_512_N4_200_steps():
    """Density after 200 steps: FP32-MFMA and BF16x3 against the FP64 run,
    and FP64 against the analytic vortex."""
    w, case = _case("vortex_512_N4")
    x, y = case.omesh.x, case.omesh.y
    q0 = np.ascontiguousarray(w.initial_state(x, y))
    runs = {}
    for key, prec in (("fp64", E.PREC_FP64), ("fp32", E.PREC_FP32_MFMA), ("bf16x3", E.PREC_BF16X3)):
        runs[key], dt = _run(w, prec, w.steps, q0)
    rho = {k: v[..., 0] for k, v in runs.items()}
    ex = w.exact_state(x, y, w.steps * dt)[..., 0]
    e64 = rel_l2(rho["fp64"], ex)
    e32, ebf = rel_l2(rho["fp32"], rho["fp64"]), rel_l2(rho["bf16x3"], rho["fp64"])
    print("vortex_512_N4: density rel L2: FP64 vs analytic %.3e, FP32 vs FP64 %.3e, "
          "BF16x3 vs FP64 %.3e" % (e64, e32, ebf), flush=True)
    assert e64 <= 1e-8
    assert e32 <= 1e-5
    assert ebf <= 1e-4


def test_tm_fourier_384_N8_300_steps():
    w, case = _case("tm_fourier_384_N8")
    q0 = np.ascontiguousarray(w.initial_state(case.omesh.x, case.omesh.y))
    runs = {}
    for key, prec in (("fp64", E.PREC_FP64), ("fp32", E.PREC_FP32_MFMA), ("bf16x3", E.PREC_BF16X3)):
        runs[key], dt = _run(w, prec, w.steps, q0)
    e32, ebf = rel_l2(runs["fp32"], runs["fp64"]), rel_l2(runs["bf16x3"], runs["fp64"])
    # energy of the TM system is conserved up to the upwind dissipation
    en = [float(np.sum(v * v)) for v in (q0, runs["fp64"])]
    print("tm_fourier_384_N8: FP32 vs FP64 %.3e, BF16x3 vs FP64 %.3e; nodal energy %.6e -> %.6e"
          % (e32, ebf, en[0], en[1]), flush=True)
    assert e32 <= 1e-5
    assert ebf <= 1e-4
    assert en[1] <= en[0] * (1 + 1e-9)


def test_adjoint_vortex_256_N3_adjoint_rhs_against_cpu():
    w, case = _case("adjoint_vortex_256_N3")
    q = np.ascontiguousarray(w.initial_state(case.omesh.x, case.omesh.y))
    lam = np.random.default_rng(11).standard_normal(q.shape)
    eng = E.Engine(case.mesh, case.params)
    out = np.zeros_like(q)
    eng.adjoint_rhs(0.0, q, lam, out)
    eng.close()
    with O.exact_geometry():
        ref = from_oracle(w.eqn, O.adjoint_rhs(case.omesh, case.ophys, 0.0, to_oracle(w.eqn, q),
                                               to_oracle(w.eqn, lam)))
    err = rel_l2(out, ref)
    print("adjoint_vortex_256_N3: adjoint RHS rel L2 vs CPU adjoint %.3e" % err, flush=True)
    assert err <= 1e-12


def test_adjoint_vortex_256_N3_reverse_lserk4_against_cpu():
    """Reverse LSERK4 over a shortened horizon (10 steps, checkpoint every 5)
    at the full mesh against the oracle's reverse LSERK4 (which stores the
    whole trajectory)."""
    w, case = _case("adjoint_vortex_256_N3")
    q0 = np.ascontiguousarray(w.initial_state(case.omesh.x, case.omesh.y))
    steps, ckpt = 10, 5
    eng = E.Engine(case.mesh, case.params)
    dt = w.fixed_dt(eng, q0)
    gT = np.zeros_like(q0)
    qT = q0.copy()
    eng.lserk4(0.0, dt, steps, qT)
    eng.objective_disc_sq(qT, 0, 0.0, 0.0, 1.0, gT)
    g0, qT2 = np.zeros_like(q0), np.zeros_like(q0)
    eng.lserk4_adjoint(0.0, dt, steps, ckpt, q0, gT, g0, qT2)
    eng.close()
    assert np.array_equal(qT, qT2)
    with O.exact_geometry():
        ref = from_oracle(w.eqn, O.lserk4_adjoint(case.omesh, case.ophys, 0.0, dt, steps,
                                                  to_oracle(w.eqn, q0), to_oracle(w.eqn, gT)))
    err = rel_l2(g0, ref)
    print("adjoint_vortex_256_N3: dJ/dq0 after %d steps rel L2 vs CPU reverse LSERK4 %.3e"
          % (steps, err), flush=True)
    assert err <= 1e-11


def test_adjoint_vortex_256_N3_finite_differences():
    """dJ/dq0 over the full horizon (200 steps, checkpoint every 20) against
    central differences along 8 Gaussian directions of two kinds: smooth
    random fields (16 Fourier modes per variable) and white noise (N(0,1)
    per degree of freedom).

    Along white noise the directional derivative g.d sums 5.2 million terms
    of random sign, so it is small next to J, and the central difference at
    eps = 1e-6 is limited by the rounding noise of J itself: its error is
    about noise(J) / (eps |g.d|) and grows as eps shrinks (measured: 3.5e-6
    at eps = 1e-6, 9.4e-5 at 1e-7).  The test measures noise(J) as the
    change of J when q0 moves by one ulp at random, predicts that error per
    direction, checks that the eps = 1e-6 error stays within a small
    multiple of the prediction, and checks the white-noise directions at
    eps = 1e-5, where rounding is 10x smaller, against 1e-6 (eps = 1e-4 is
    printed).
    Smooth directions meet 1e-6 at eps = 1e-6."""
    import torch
    w, case = _case("adjoint_vortex_256_N3")
    x, y = case.omesh.x, case.omesh.y
    q0 = np.ascontiguousarray(w.initial_state(x, y))
    eng = E.Engine(case.mesh, case.params)
    dt = w.fixed_dt(eng, q0)
    disc = (0, 0.0, 0.0, 1.0)
    q0d = torch.from_numpy(q0).cuda()

    def objective(qinit):
        q = qinit.clone()
        eng.lserk4(0.0, dt, w.steps, q)
        return eng.objective_disc_sq(q, *disc)

    qT = q0d.clone()
    eng.lserk4(0.0, dt, w.steps, qT)
    gT = torch.zeros_like(q0d)
    J0 = eng.objective_disc_sq(qT, *disc, dJdq=gT)
    grad = torch.empty_like(q0d)
    eng.lserk4_adjoint(0.0, dt, w.steps, w.ckpt, q0d, gT, grad)

    def fd_err(d, eps):
        fd = (objective(q0d + eps * d) - objective(q0d - eps * d)) / (2 * eps)
        ad = float((grad * d).sum())
        return abs(fd - ad) / abs(ad), abs(ad)

    rng = np.random.default_rng(8)
    smooth = []
    for _ in range(8):
        d = np.zeros(q0.shape)
        for v in range(q0.shape[-1]):
            for _ in range(16):
                kx, ky = rng.integers(-3, 4, size=2)
                ph = rng.uniform(0, 2 * np.pi)
                d[..., v] += rng.standard_normal() * np.cos(2 * np.pi * (kx * x + ky * y) / w.L + ph)
        smooth.append(fd_err(torch.from_numpy(d).cuda(), 1e-6)[0])
    # rounding noise of J: J after q0 moves by -1, 0 or +1 ulp per entry
    ulp = np.finfo(np.float64).eps
    noise = 0.0
    for s in (1, 2, 3):
        r = torch.from_numpy(np.random.default_rng(s).integers(-1, 2, size=q0.shape).astype(np.float64))
        noise = max(noise, abs(objective(q0d * (1.0 + ulp * r.cuda())) - J0))
    gen = torch.Generator(device="cpu").manual_seed(8)
    dirs = [torch.randn(q0.shape, generator=gen, dtype=torch.float64).cuda() for _ in range(8)]
    white = {eps: [fd_err(d, eps) for d in dirs] for eps in (1e-4, 1e-5, 1e-6)}
    eng.close()
    # predicted rounding error of the eps = 1e-6 central difference per direction
    pred = [noise / (1e-6 * gd) for _, gd in white[1e-6]]
    ratio = max(e / p for (e, _), p in zip(white[1e-6], pred))
    print("adjoint_vortex_256_N3: J %.15f, rounding noise of J %.2e; FD rel err: smooth "
          "eps=1e-6 max %.2e; white noise max %s; white noise eps=1e-6 error / predicted "
          "rounding error max %.2f (predicted max %.2e)"
          % (J0, noise, max(smooth),
             ", ".join("eps=%.0e %.2e" % (e, max(v for v, _ in white[e])) for e in white),
             ratio, max(pred)), flush=True)
    assert max(smooth) <= 1e-6
    assert max(v for v, _ in white[1e-5]) <= 1e-6
    assert ratio <= 4.0
