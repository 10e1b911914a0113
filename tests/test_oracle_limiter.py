"""Pins the oracle's restatement of Apollo's Hesthaven-Warburton Euler
limiter (src/hyperapps/euler/wxhesthavenwarburtoneulerlimiter.cc:61-484) and
the Euler boundary states it uses (CPU only).

Checks:
- the cell-average weights integrate polynomials of degree <= N exactly;
- invariants of the reconstruction that hold whatever the patch gradients
  are: a constant state is a fixed point, the conserved cell averages are
  kept (no floor active), the limited density is affine in x and y on every
  element, the floors bound the density from below;
- an element-by-element transliteration of the reference's loops, written
  here independently of the C restatement (different data layout, explicit
  per-face bookkeeping), on the forward-facing-step mesh and on a periodic
  mesh;
- the refusals: N > 4 and the isentropic-vortex boundary."""
import math

import numpy as np
import pytest

from apollo_amd import engine as E
from oracle import oracle as O

from cases import euler_case, ffs_case, ffs_state, from_oracle, rel_l2, to_oracle

from test_oracle import _tri_monomial


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_limiter_weights_are_cell_averages(N):
    ops = O.Ops(N)
    ave = O.limiter_weights(ops)
    assert abs(ave.sum() - 1.0) < 1e-14
    for a in range(N + 1):
        for b in range(N + 1 - a):
            exact = float(_tri_monomial(a, b)) / 2.0        # reference area 2
            assert abs(ave @ (ops.r ** a * ops.s ** b) - exact) < 1e-13


def test_euler_boundary_states():
    ph = O.phys(O.ORC_EULER, bcs=[(O.ORC_BC_EULER_WALL, []),
                                  (O.ORC_BC_EULER_INFLOW, [3.0, 30.0, 1.4, 5.0, 1.2]),
                                  (O.ORC_BC_EULER_DIRICHLET, [0.125, 0.3, -0.2, 0.1, 1.4, 0.1]),
                                  (O.ORC_BC_COPY, [])])
    q = np.array([1.3, 0.4, -0.7, 0.2, 4.0])
    n = np.array([0.6, 0.8])
    w = O.bc(ph, 1, np.zeros(3), n, q)
    # wall: normal momentum reversed, tangential kept, rho w dropped
    assert np.isclose(w[1] * n[0] + w[2] * n[1], -(q[1] * n[0] + q[2] * n[1]))
    assert np.isclose(-w[1] * n[1] + w[2] * n[0], -q[1] * n[1] + q[2] * n[0])
    assert w[0] == q[0] and w[3] == 0.0 and w[4] == q[4]
    inflow = O.bc(ph, 2, np.zeros(3), n, q)
    c = math.sqrt(1.4 * 5.0 / 1.2)
    u, v = 3.0 * c * math.cos(math.pi / 6), 3.0 * c * math.sin(math.pi / 6)
    assert np.allclose(inflow, [1.2, 1.2 * u, 1.2 * v, 0.0, 5.0 / 0.4 + 0.6 * (u * u + v * v)],
                       rtol=1e-15)
    d = O.bc(ph, 3, np.zeros(3), n, q)
    assert np.allclose(d, [0.125, 0.0375, -0.025, 0.0125, 0.25 + 0.0625 * 0.14], rtol=1e-15)
    assert np.array_equal(O.bc(ph, 4, np.zeros(3), n, q), q)


def _ave_of(ave, q):
    return np.einsum("n,knc->kc", ave, q)


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_limiter_invariants(N):
    case = ffs_case(N=N)
    mesh, ph = case.omesh, case.ophys
    ave = O.limiter_weights(case.ops)
    # constant state: fixed point
    q = to_oracle(case.params.eqn, ffs_state(case))
    ql = O.limit(mesh, ph, q)
    assert np.abs(ql - q).max() <= 1e-13 * np.abs(q).max()
    # smooth state: averages kept, density affine per element
    q = to_oracle(case.params.eqn, ffs_state(case, "smooth", seed=N))
    ql = O.limit(mesh, ph, q)
    assert np.abs(_ave_of(ave, ql) - _ave_of(ave, q)).max() <= 1e-13 * np.abs(q).max()
    assert rel_l2(ql, q) > 1e-4        # it did change the state
    X = np.stack([np.ones_like(mesh.x), mesh.x, mesh.y], axis=-1)
    for k in range(mesh.K):
        coef, *_ = np.linalg.lstsq(X[k], ql[k, :, 0], rcond=None)
        assert np.abs(X[k] @ coef - ql[k, :, 0]).max() <= 1e-12 * ql[k, :, 0].max()


def test_limiter_floors():
    """Floors above part of the state: a node whose reconstructed density
    stays below min_density after four halvings of the slope takes the cell
    average; one whose pressure stays below min_pressure takes the cell
    average pressure (wxhesthavenwarburtoneulerlimiter.cc:413-460)."""
    case = ffs_case(N=2)
    mesh = case.omesh
    mind, minp = 1.45, 5.2
    ph = O.phys(O.ORC_EULER, flux=O.ORC_FLUX_HLL, bcs=case.params.bcs, limiter=(1.4, mind, minp))
    q = to_oracle(case.params.eqn, ffs_state(case, "smooth", seed=3))
    ql = O.limit(mesh, ph, q)
    avg = _ave_of(O.limiter_weights(case.ops), q)
    rho = ql[..., 0]
    at_avg = rho == avg[:, None, 0]
    assert np.all((rho > mind) | at_avg) and at_avg.sum() > 10 and (rho > mind).sum() > 10
    pres = 0.4 * (ql[..., 4] - 0.5 * (ql[..., 1] ** 2 + ql[..., 2] ** 2 + ql[..., 3] ** 2) / rho)
    pavg = 0.4 * (avg[:, 4] - 0.5 * (avg[:, 1] ** 2 + avg[:, 2] ** 2 + avg[:, 3] ** 2) / avg[:, 0])
    low = pres <= minp
    assert low.sum() > 10 and np.allclose(pres[low], np.broadcast_to(pavg[:, None], pres.shape)[low],
                                          rtol=1e-12)
    want = _limit_transliterated(case.ops, mesh, case.params.bcs, (1.4, mind, minp), q)
    assert np.abs(ql - want).max() <= 1e-13 * np.abs(want).max()


def test_limiter_refusals():
    case = ffs_case(N=1)
    with pytest.raises(ValueError, match="11"):
        c5 = ffs_case(N=5)
        O.limit(c5.omesh, c5.ophys, to_oracle(c5.params.eqn, ffs_state(c5)))
    cv = euler_case(n=4, N=2, m=5, periodic=False)
    phv = O.phys(O.ORC_EULER, bcs=cv.params.bcs, limiter=(1.4, 1e-2, 1e-2))
    with pytest.raises(ValueError, match="11"):
        O.limit(cv.omesh, phv, cv.state())
    assert case.ophys.limiter == 1


# --- the transliteration ------------------------------------------------------

def _bc(kind, p, nrm, q):
    """wxeulerwallbc.cc:18-25, wxeulerinflowbc.cc:21-36,
    wxeulerdirichletbc.cc:22-30, wxeulerzerogradientbc.cc:18-25."""
    if kind == E.BC_EULER_WALL:
        d = nrm[0] * q[1] + nrm[1] * q[2]
        return [q[0], q[1] - 2. * d * nrm[0], q[2] - 2. * d * nrm[1], 0.0, q[4]]
    if kind == E.BC_EULER_INFLOW:
        mach, ang, g, pres, rho = p[:5]
        snd = math.sqrt(g * pres / rho)
        u, v = mach * snd * math.cos(ang * math.pi / 180.), mach * snd * math.sin(ang * math.pi / 180.)
        return [rho, rho * u, rho * v, 0.0, pres / (g - 1) + 0.5 * rho * (u * u + v * v)]
    if kind == E.BC_EULER_DIRICHLET:
        rho, u, v, w, g, pres = p[:6]
        return [rho, rho * u, rho * v, rho * w, pres / (g - 1) + 0.5 * rho * (u * u + v * v + w * w)]
    return list(q)


def _prim(g, c):
    return [c[0], c[1] / c[0], c[2] / c[0], c[3] / c[0],
            (g - 1.) * (c[4] - 0.5 * (c[1] * c[1] + c[2] * c[2] + c[3] * c[3]) / c[0])]


def _limit_transliterated(ops, mesh, bcs, lim, q):
    """applyLimiter read line by line, with one dict per element."""
    g, mind, minp = lim
    Np, NpF = ops.Np, ops.Nfp
    F = [int(v) for v in ops.fmask.ravel()]
    Mass = np.linalg.inv(ops.VVT)
    AVE = [sum(0.5 * Mass.ravel()[k + m * Np] for m in range(Np)) for k in range(Np)]
    drop = [[(1. if i == j else 0.) - AVE[j] for j in range(Np)] for i in range(Np)]
    X, Y, ET = mesh.x, mesh.y, mesh.etetf
    K = mesh.K

    def geo(k):
        x1, y1 = X[k][F[0]], Y[k][F[0]]
        x2, y2 = X[k][F[NpF]], Y[k][F[NpF]]
        x3, y3 = X[k][F[2 * NpF]], Y[k][F[2 * NpF]]
        J = -((x3 - x1) / 2) * ((y2 - y1) / 2) + ((x2 - x1) / 2) * ((y3 - y1) / 2)
        nr = [(y2 - y1, -(x2 - x1)), (y3 - y2, -(x3 - x2)), (y1 - y3, -(x1 - x3))]
        out = []
        for nx, ny in nr:
            sj = math.sqrt(nx * nx + ny * ny)
            out.append((nx / sj, ny / sj, sj / 2.))
        return J, out

    def cell_avg(k):
        c = [sum(AVE[n] * q[k][n][v] for n in range(Np)) for v in range(5)]
        return c, _prim(g, c)

    data = []
    for e in range(K):
        J, nrm = geo(e)
        xc, yc, A0 = [0.] * 4, [0.] * 4, [0.] * 4
        for kk in range(3):
            xc[0] += AVE[kk] * X[e][F[kk * NpF]]
            yc[0] += AVE[kk] * Y[e][F[kk * NpF]]
            A0[0] += 2. / 3. * AVE[kk] * J
        for f in range(3):
            if ET[e][2 * f] < 0:
                H1 = 2.0 * A0[0] / nrm[f][2]
                xc[f + 1] = xc[0] + 2.0 * nrm[f][0] * H1
                yc[f + 1] = yc[0] + 2.0 * nrm[f][1] * H1
            else:
                nb = ET[e][2 * f]
                Jn, _ = geo(nb)
                for node in range(3):
                    xc[f + 1] += AVE[node] * X[nb][F[node * NpF]]
                    yc[f + 1] += AVE[node] * Y[nb][F[node * NpF]]
                    A0[f + 1] += 2. / 3. * AVE[node] * Jn
        for f in range(3):
            A0[f + 1] += A0[0]
        cons, prim = [None] * 4, [None] * 4
        cons[0], prim[0] = cell_avg(e)
        QM = {(f, n): list(q[e][F[f * NpF + n]]) for f in range(3) for n in range(NpF)}
        QP = {}
        qIn = e
        for f in range(3):
            if ET[e][2 * f] < 0:
                kind, p = bcs[abs(ET[e][2 * f + 1]) - 1]
                cons[f + 1] = _bc(kind, p, nrm[f], cons[0])
                prim[f + 1] = _prim(g, cons[f + 1])
                for n in range(NpF):
                    QP[(f, n)] = _bc(kind, p, nrm[f], list(q[qIn][F[f * NpF + n]]))
            else:
                qIn = ET[e][2 * f]
                cons[f + 1], prim[f + 1] = cell_avg(qIn)
                fn = ET[e][2 * f + 1]
                for n in range(NpF):
                    QP[(f, n)] = list(q[qIn][F[fn * NpF + NpF - 1 - n]])
        flat = [0.5 * (QP[(pp // NpF, pp % NpF)][v] + QM[(pp // NpF, pp % NpF)][v])
                for pp in range(3 * NpF) for v in range(5)]
        ends = [F[0], F[NpF - 1], F[NpF], F[2 * NpF - 1], F[2 * NpF], F[3 * NpF - 1]]
        fP = [_prim(g, flat[5 * end:5 * end + 5]) for end in ends]
        xv1, yv1 = X[e][ends[0]], Y[e][ends[0]]
        xv2, yv2 = X[e][ends[2]], Y[e][ends[2]]
        xv3, yv3 = X[e][ends[4]], Y[e][ends[4]]
        gr = {}
        for c in range(5):
            E1 = (0.5 * ((prim[1][c] - prim[0][c]) * (yv2 - yv1) + (fP[0][c] - fP[1][c]) * (yc[1] - yc[0])) / A0[1],
                  -0.5 * ((prim[1][c] - prim[0][c]) * (xv2 - xv1) + (fP[0][c] - fP[1][c]) * (xc[1] - xc[0])) / A0[1])
            E2 = (0.5 * ((prim[2][c] - prim[0][c]) * (yv3 - yv2) + (fP[2][c] - fP[3][c]) * (yc[2] - yc[0])) / A0[2],
                  -0.5 * ((prim[2][c] - prim[0][c]) * (xv3 - xv2) + (fP[2][c] - fP[3][c]) * (xc[2] - xc[0])) / A0[2])
            E3 = (0.5 * ((prim[3][c] - prim[0][c]) * (yv3 - yv2) + (fP[4][c] - fP[5][c]) * (yc[3] - yc[0])) / A0[3],
                  -0.5 * ((prim[3][c] - prim[0][c]) * (xv3 - xv2) + (fP[4][c] - fP[5][c]) * (xc[3] - xc[0])) / A0[3])
            C0 = tuple((A0[1] * E1[i] + A0[2] * E2[i] + A0[3] * E3[i]) / (A0[1] + A0[2] + A0[3])
                       for i in range(2))
            gr[c] = (E1, E2, E3, C0)
        data.append((cons[0], prim[0], gr))

    out = np.zeros((K, Np, 5))
    for e in range(K):
        ldx = [sum(drop[i][j] * X[e][j] for j in range(Np)) for i in range(Np)]
        ldy = [sum(drop[i][j] * Y[e][j] for j in range(Np)) for i in range(Np)]
        qdv = np.zeros((Np, 5))
        for c in range(5):
            G = [data[e][2][c][f] if ET[e][2 * f] < 0 else data[ET[e][2 * f]][2][c][3] for f in range(3)]
            g1, g2, g3 = (G[f][0] ** 2 + G[f][1] ** 2 for f in range(3))
            fac3e = g1 * g1 + g2 * g2 + g3 * g3 + 3.0 * 1.e-14
            w = ((g2 * g3 + 1.e-14) / fac3e, (g1 * g3 + 1.e-14) / fac3e, (g1 * g2 + 1.e-14) / fac3e)
            Lx = sum(w[f] * G[f][0] for f in range(3))
            Ly = sum(w[f] * G[f][1] for f in range(3))
            for n in range(Np):
                qdv[n][c] = ldx[n] * Lx + ldy[n] * Ly
        AC, AP = data[e][0], data[e][1]
        for n in range(Np):
            drho, du, dv, dw, dp = qdv[n]
            rho = AP[0] + drho
            if rho < mind:
                for it in range(4):
                    drho *= 0.5
                    rho = AP[0] + drho
                    if rho > mind:
                        break
                    elif it == 3:
                        rho = AC[0]
            mom = [AC[1 + i] + AP[0] * d + drho * AP[1 + i] for i, d in enumerate((du, dv, dw))]
            kin = AP[1] ** 2 + AP[2] ** 2 + AP[3] ** 2
            cross = AP[1] * du + AP[2] * dv + AP[3] * dw
            ener = AC[4] + dp / (g - 1.) + 0.5 * drho * kin + AP[0] * cross
            pres = (g - 1.) * (ener - 0.5 * sum(m * m for m in mom) / rho)
            if pres < minp:
                for it in range(4):
                    dp *= 0.5
                    ener = AC[4] + dp / (g - 1.) + 0.5 * drho * kin + AP[0] * cross
                    pres = (g - 1.) * (ener - 0.5 * sum(m * m for m in mom) / rho)
                    if pres > minp:
                        break
                    elif it == 3:
                        ener = AP[4] / (g - 1.) + 0.5 * sum(m * m for m in mom) / rho
            out[e][n] = [rho, mom[0], mom[1], mom[2], ener]
    return out


@pytest.mark.parametrize("N", [1, 2, 4])
def test_limiter_against_transliteration_ffs(N):
    case = ffs_case(N=N)
    q = to_oracle(case.params.eqn, ffs_state(case, "smooth", seed=10 + N))
    got = O.limit(case.omesh, case.ophys, q)
    want = _limit_transliterated(case.ops, case.omesh, case.params.bcs, case.params.limiter, q)
    err = np.abs(got - want).max() / np.abs(want).max()
    assert err <= 1e-13, err


def test_limiter_against_transliteration_floors_and_dirichlet():
    """Dirichlet walls on a small square, a strong density/pressure jump so
    both floors act, and a second variant with z momentum."""
    case = euler_case(n=4, N=3, m=5, periodic=False)
    bcs = [(E.BC_EULER_DIRICHLET, [0.125, 0.0, 0.0, 0.0, 1.4, 0.1])]
    lim = (1.4, 0.3, 0.3)
    ph = O.phys(O.ORC_EULER, bcs=bcs, limiter=lim)
    x, y = case.omesh.x, case.omesh.y
    rho = np.where(x < 0.3, 1.0, 0.125) + 0.05 * np.sin(y)
    p = np.where(x * x + y * y < 4.0, 1.0, 0.1)
    u, v, w = 0.2 * np.cos(y), 0.1 * np.sin(x), 0.05 * np.cos(x + y)
    q = np.stack([rho, rho * u, rho * v, rho * w, p / 0.4 + 0.5 * rho * (u * u + v * v + w * w)], -1)
    got = O.limit(case.omesh, ph, q)
    want = _limit_transliterated(case.ops, case.omesh, bcs, lim, q)
    assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()


def test_limiter_against_transliteration_periodic():
    case = euler_case(n=5, N=2, m=5)
    ph = O.phys(O.ORC_EULER, limiter=(1.4, 1e-2, 1e-2))
    q = case.state(amp=5e-2)
    got = O.limit(case.omesh, ph, q)
    want = _limit_transliterated(case.ops, case.omesh, [], (1.4, 1e-2, 1e-2), q)
    assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()


def test_ssp_with_limiter_conserves_mass_periodic():
    """Limiter (average-preserving away from the floors) inside Apollo's
    stepping on a periodic mesh: total mass is conserved to rounding."""
    case = euler_case(n=6, N=2, m=5)
    ph = O.phys(O.ORC_EULER, limiter=(1.4, 1e-2, 1e-2))
    q0 = case.state(amp=2e-2)
    _, ms = case.oracle_rhs(q0)
    s0 = O.suggested_dt(case.omesh, 0.3, ms)
    q1, dts = O.ssp_advance(case.omesh, ph, 0.0, 4.5 * s0, s0, 0.3, q0)
    ave = O.limiter_weights(case.ops)
    J = np.array([O.elem_geom(case.ops, case.omesh.x[k], case.omesh.y[k])[4] for k in range(case.omesh.K)])
    m0 = np.sum(J * _ave_of(ave, q0)[:, 0])
    m1 = np.sum(J * _ave_of(ave, q1)[:, 0])
    assert len(dts) >= 4 and abs(m1 - m0) <= 1e-13 * abs(m0)



@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_limiter_weights_flatten_next_to_a_zero_gradient(N):
    """The limited gradient is sum_i w_i grad_i with w_1 = (g_2 g_3 + eps) /
    (g_1^2 + g_2^2 + g_3^2 + 3 eps), g_i = |grad_i|^2 (Hesthaven & Warburton
    2008, the EulerLimiter2D weights; wxhesthavenwarburtoneulerlimiter.cc:
    317-360): when one of the three gradients is exactly zero, the other
    two weights are eps / (g_j^2 + g_k^2) ~ 1e-13, so the element is
    flattened to its cell average.  Set-up: zero-gradient (copy) walls and
    a density that varies in y only, with velocity 0 and pressure 1.  On an
    element whose face 0 lies on the wall y = 0, the ghost cell repeats the
    cell average and both face ends carry rho(0) = 1, so that face's edge
    gradient is exactly zero for every variable (face 0 is visited first,
    before any state pointer is moved to a neighbour).  Weights that do not
    multiply the other two gradients (equal weights, w_i ~ g_i, a swapped
    index) leave a nonzero slope there; interior elements keep a slope."""
    g = 1.4
    mesh = E.structured_mesh(12, 0.0, 1.0, 0.0, 1.0, 0.1, 7, False)
    ops = O.Ops(N)
    omesh = O.Mesh(ops, mesh.VX, mesh.VY, mesh.EToV, mesh.vert_id, mesh.face_tag)
    ph = O.phys(O.ORC_EULER, bcs=[(O.ORC_BC_COPY, [])], limiter=(g, 1e-6, 1e-6))
    x, y = omesh.x, omesh.y
    rho = 1.0 + 0.5 * y
    z = np.zeros_like(rho)
    q = np.stack([rho, z, z, z, np.full_like(rho, 1.0 / (g - 1))], axis=-1)
    ql = O.limit(omesh, ph, q)
    avg = np.einsum("n,knc->kc", O.limiter_weights(ops), q)
    et = omesh.etetf
    v = mesh.EToV
    bottom = (et[:, 1] < 0) & (mesh.VY[v[:, 0]] == 0.0) & (mesh.VY[v[:, 1]] == 0.0)
    assert bottom.sum() == 12
    flat = np.abs(ql[bottom] - avg[bottom][:, None, :]).max()
    print("N=%d: %d elements on y = 0, max |limited - cell average| %.2e" % (N, bottom.sum(), flat))
    assert flat <= 1e-13
    interior = ~np.any(et[:, 1::2] < 0, axis=1)
    spread = ql[interior, :, 0].max(axis=1) - ql[interior, :, 0].min(axis=1)
    assert spread.min() > 1e-3
