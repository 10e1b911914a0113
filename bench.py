#!/usr/bin/env python3
"""Benchmark: DG RHS DOF-updates/s on one MI355X.

Headline workload (tm_cavity_256_N5, the configuration the metric is quoted
on): 2D Maxwell TM, 256 x 256 quads split into 2 triangles (K = 131072),
vertices jittered +-0.1 h (seed 1), N = 5 (Np = 21), PEC walls with Apollo's
transverse-magnetic boundary state, upwind flux, Ez = sin(3 pi x) sin(4 pi y),
H = 0, LSERK4 with Apollo's dt formula at the deck CFL 0.5/(2N+1), fixed dt.

A "step" is one LSERK4 step = 5 RHS evaluations fused with the stage
updates, replayed as one hipGraph.  DOF-updates/s = K Np m x (RHS
evaluations) / time with m = 3 (Ez, Bx, By).  `value` is the FP64 engine
(the reference's precision); FP32-MFMA and split-BF16x3 contractions and
Apollo's over-integrated cubature form are timed the same way beside it,
each with its speed-up over FP64 and its relative L2 error against the
analytic cavity mode and against the FP64 run after 500 steps (a fixed-length
accuracy run, whatever --steps is).  State lives in HBM before the timed
region starts.

Legs (each timed with HIP events on the engine's stream):
  vortex_64_N3       Euler vortex, K = 8192, N = 3, 100 steps, 3 precisions
  vortex_512_N4      Euler vortex, K = 524288, N = 4, 200 steps, 3 precisions,
                     density error against the FP64 run
  tm_fourier_384_N8  Maxwell TM, K = 294912, N = 8, 300 steps, 3 precisions
  adjoint            reverse LSERK4 on adjoint_vortex_256_N3 (200 steps,
                     checkpoint 20) and dJ/dq0 against central differences
  limiter            apl_limit at N = 4 and N = 1 on K = 524288
cpu_baseline: the oracle (Apollo's loop restated in C, OpenMP over elements)
on the host cores, for the headline workload and each leg's workload.

    python bench.py [--gpus 1] [--steps 500] [--warmup 20] [--workload NAME]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DG RHS DOF-updates/s: FP64 plain vs FP32-MFMA and adjoint, at stated L2 error"
PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "bf16": 2.5e15, "hbm": 8.0e12}
TRAFFIC = os.path.join(ROOT, "profiles", "hbm_traffic_maxwell_tm_N5_K131072.json")
HEAD_WORKLOAD = "tm_cavity_256_N5"


def flop_alg(K, Np, Nc, Ng, m):
    """Contraction flops of one RHS of Apollo's cubature form:
    K 2 m (3 Nc Np + 6 Ng Np) (1 FMA = 2 flops)."""
    return K * 2 * m * (3 * Nc * Np + 6 * Ng * Np)


def flop_exec(linear, K, N, Np, Nc, Ng, m):
    """Contraction flops the engine's kernel actually performs per RHS (no
    padding): the cubature form does V_c (Nc Np), A_r|A_s (2 Nc Np), the face
    traces (2 x 3 Ng (N+1)) and L (3 Ng Np) per variable; the nodal form of a
    linear flux does Dt_r|Dt_s (2 Np^2) and Lt (3 (N+1) Np)."""
    Nfp = N + 1
    if linear:
        per = 2 * Np * Np + 3 * Nfp * Np
    else:
        per = 3 * Nc * Np + 6 * Ng * Nfp + 3 * Ng * Np
    return K * 2 * m * per


def bytes_alg_stage(K, Np, m):
    """HBM bytes of one fused RHS + LSERK4 stage: read q and res, write q and
    res (FP64), plus 136 B of geometry and connectivity per element."""
    return K * (4 * Np * m * 8 + 136)


VARIANTS = {
    # name: (precision, algorithm)
    "fp64": (0, 0), "fp32": (1, 0), "bf16x3": (2, 0),
    "fp64_cub": (0, 1), "fp32_cub": (1, 1), "bf16x3_cub": (2, 1),
    # the cubature form on a named kernel: MFMA at every order, or the
    # vector-ALU kernel (N <= 3)
    "fp64_mfma": (0, 2), "fp32_mfma": (1, 2), "fp64_valu": (0, 3), "fp32_valu": (1, 3),
    # the nodal form of a linear flux on a named kernel
    "fp64_nmfma": (0, 4), "fp32_nmfma": (1, 4), "fp64_nvalu": (0, 5), "fp32_nvalu": (1, 5),
}
NODAL_ALGOS = (0, 4, 5)


def engine_on_stream(w, precision, algorithm, torch):
    from apollo_amd import engine as E
    eng = E.Engine(w.mesh(), w.params(precision, algorithm))
    stream = torch.cuda.Stream()
    E.lib().apl_set_stream(eng._h, ctypes.c_void_p(stream.cuda_stream))
    return eng, stream


def run_gpu(w, variant, steps, warmup, torch):
    """`steps` LSERK4 steps of one variant, timed with HIP events on the
    engine's stream; returns the timing and the final state."""
    precision, algorithm = VARIANTS[variant]
    eng, stream = engine_on_stream(w, precision, algorithm, torch)
    g, x, y = eng.geometry()
    q0 = np.ascontiguousarray(w.initial_state(x, y))
    dt = w.fixed_dt(eng, q0)
    with torch.cuda.stream(stream):
        qd = torch.from_numpy(q0.copy()).cuda()
        # warm-up (graph capture happens here)
        eng.lserk4(0.0, dt, max(warmup, 1), qd)
        qd.copy_(torch.from_numpy(q0))
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        ev0.record(stream)
        eng.lserk4(0.0, dt, steps, qd)
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_time(ev1)
    qT = qd.cpu().numpy()
    info = dict(K=eng.K, N=eng.N, Np=eng.Np, Nc=eng.Nc, Ng=eng.Ng, m=eng.m, dt=dt)
    eng.close()
    return dict(wall=wall, gpu_ms=gpu_ms, qT=qT, x=x, y=y, info=info, dt=dt, q0=q0)


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def variant_figures(w, variant, r, steps):
    """DOF-updates/s, per-launch time and roofline figures of one run."""
    info = r["info"]
    K, N, Np, Nc, Ng, m = (info[k] for k in ("K", "N", "Np", "Nc", "Ng", "m"))
    lin = w.linear and VARIANTS[variant][1] in NODAL_ALGOS
    kern_s = r["gpu_ms"] * 1e-3 / (5 * steps)
    d = dict(dof_updates_per_s=K * Np * m / kern_s, ms_per_step=r["gpu_ms"] / steps,
             kernel_us=kern_s * 1e6, form="nodal" if lin else "cubature",
             tflops_alg=flop_alg(K, Np, Nc, Ng, m) / kern_s / 1e12,
             tflops_exec=flop_exec(lin, K, N, Np, Nc, Ng, m) / kern_s / 1e12,
             gbs_alg=bytes_alg_stage(K, Np, m) / kern_s / 1e9)
    d["hbm_frac_alg"] = d["gbs_alg"] * 1e9 / PEAK["hbm"]
    pk = PEAK["fp64" if variant.startswith("fp64") else ("fp32" if variant.startswith("fp32") else "bf16")]
    d["mfma_frac_exec"] = d["tflops_exec"] * 1e12 / pk
    return d


def run_leg(name, variants, torch, steps=None):
    """A workload timed like the headline one over its own step count: each
    variant's figures, its speed-up over FP64 and its relative L2 error
    against the FP64 run (density only for Euler) and the analytic solution
    where there is one."""
    from apollo_amd import workloads as W
    w = W.get(name)
    steps = w.steps if steps is None else steps
    res = {}
    for v in variants:
        print("bench: %s %s steps=%d" % (name, v, steps), file=sys.stderr, flush=True)
        res[v] = run_gpu(w, v, steps, 3, torch)
    sel = (lambda q: q[..., 0]) if w.m == 4 else (lambda q: q)
    ref = res["fp64"]["qT"]
    out = {"workload": w.desc, "steps": steps, "dt": res["fp64"]["dt"], "variants": {}}
    for v in variants:
        r = res[v]
        d = variant_figures(w, v, r, steps)
        if v != "fp64":
            d["l2_vs_fp64"] = rel_l2(sel(r["qT"]), sel(ref))
            d["speedup_vs_fp64"] = res["fp64"]["gpu_ms"] / r["gpu_ms"]
        ex = w.exact_state(r["x"], r["y"], steps * r["dt"])
        if ex is not None:
            d["l2_vs_analytic"] = rel_l2(sel(r["qT"]), sel(ex))
        out["variants"][v] = d
    if w.m == 4:
        out["error_field"] = "density"
    return out


def run_adjoint(torch):
    """Reverse LSERK4 on adjoint_vortex_256_N3: timing, and dJ/dq0 against
    central finite differences along 8 Gaussian directions of two kinds,
    smooth random fields and white noise, at eps = 1e-6 (white noise also at
    1e-5 and 1e-7, to show how its error scales with eps), with the rounding
    noise of J that limits the white-noise differences."""
    from apollo_amd import workloads as W
    w = W.get("adjoint_vortex_256_N3")
    steps, ckpt = w.steps, w.ckpt
    eng, stream = engine_on_stream(w, 0, 0, torch)
    g, x, y = eng.geometry()
    q0 = np.ascontiguousarray(w.initial_state(x, y))
    dt = w.fixed_dt(eng, q0)
    disc = (0, 0.0, 0.0, 1.0)              # density, centre (0, 0), radius 1
    with torch.cuda.stream(stream):
        q0d = torch.from_numpy(q0).cuda()

        def objective(qinit):
            q = qinit.clone()
            eng.lserk4(0.0, dt, steps, q)
            return eng.objective_disc_sq(q, *disc)

        qT = torch.empty_like(q0d)
        J = objective(q0d)
        eng.lserk4(0.0, dt, steps, qT.copy_(q0d))
        qbarT = torch.zeros_like(q0d)
        eng.objective_disc_sq(qT, *disc, dJdq=qbarT)
        grad = torch.empty_like(q0d)
        eng.lserk4_adjoint(0.0, dt, steps, ckpt, q0d, qbarT, grad)      # warm-up
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        ev[0].record(stream)
        eng.lserk4_adjoint(0.0, dt, steps, ckpt, q0d, qbarT, grad)
        ev[1].record(stream)
        # the adjoint RHS alone, 100 evaluations on device pointers
        lam = torch.randn(q0d.shape, dtype=torch.float64, device=q0d.device)
        out = torch.empty_like(q0d)
        eng.adjoint_rhs(0.0, q0d, lam, out)
        ev[2].record(stream)
        for _ in range(100):
            eng.adjoint_rhs(0.0, q0d, lam, out)
        ev[3].record(stream)
        torch.cuda.synchronize()
        t_adj = ev[0].elapsed_time(ev[1]) * 1e-3
        t_arhs = ev[2].elapsed_time(ev[3]) * 1e-3 / 100

        def fd_err(d, eps):
            fd = (objective(q0d + eps * d) - objective(q0d - eps * d)) / (2 * eps)
            ad = float((grad * d).sum())
            return abs(fd - ad) / abs(ad)

        # rounding noise of J: its change when q0 moves by -1, 0 or +1 ulp per entry
        noise = 0.0
        for sd in (1, 2, 3):
            r = np.random.default_rng(sd).integers(-1, 2, size=q0.shape).astype(np.float64)
            noise = max(noise, abs(objective(q0d * (1.0 + np.finfo(np.float64).eps *
                                                    torch.from_numpy(r).cuda())) - J))
        gen = torch.Generator(device="cpu").manual_seed(8)
        white = [torch.randn(q0d.shape, generator=gen, dtype=torch.float64).cuda() for _ in range(8)]
        gd_white = [abs(float((grad * d).sum())) for d in white]
        errs5 = [fd_err(d, 1e-5) for d in white]
        errs6 = [fd_err(d, 1e-6) for d in white]
        errs7 = [fd_err(d, 1e-7) for d in white]
        # smooth Gaussian directions: 16 random Fourier modes of the periodic
        # domain per variable, N(0,1) amplitudes
        rng = np.random.default_rng(8)
        errs_smooth = []
        for _ in range(8):
            d = np.zeros(q0.shape)
            for v in range(q0.shape[-1]):
                for _ in range(16):
                    kx, ky = rng.integers(-3, 4, size=2)
                    ph = rng.uniform(0, 2 * np.pi)
                    d[..., v] += rng.standard_normal() * np.cos(2 * np.pi * (kx * x + ky * y) / w.L + ph)
            errs_smooth.append(fd_err(torch.from_numpy(d).cuda(), 1e-6))
    dofs = eng.K * eng.Np * w.m
    res = dict(workload=w.desc, steps=steps, ckpt_every=ckpt, dt=dt, J=J,
               seconds=t_adj,
               adjoint_dof_updates_per_s=dofs * 5 * steps / t_adj,
               adjoint_rhs_us=t_arhs * 1e6,
               adjoint_rhs_dof_updates_per_s=dofs / t_arhs,
               fd_directions=8, fd_eps=1e-6,
               fd_max_rel_err_smooth=max(errs_smooth),
               fd_max_rel_err_white_noise=max(errs6),
               fd_max_rel_err_white_noise_eps1e5=max(errs5),
               fd_max_rel_err_white_noise_eps1e7=max(errs7),
               J_rounding_noise=noise,
               fd_white_noise_predicted_rounding_err_eps1e6=max(noise / (1e-6 * g) for g in gd_white),
               fd_note="smooth: Gaussian random fields (16 Fourier modes per variable, N(0,1) "
                       "amplitudes); white noise: N(0,1) per degree of freedom.  Along white noise "
                       "g.d sums 5.2M terms of random sign and is small next to J, so the central "
                       "difference is limited by the rounding noise of J: its error is about "
                       "noise(J) / (eps |g.d|) (the predicted figure) and grows as eps shrinks "
                       "(compare eps 1e-5, 1e-6, 1e-7)",
               note="adjoint DOF-updates/s counts the 5 x steps adjoint RHS evaluations of the "
                    "reverse sweep over the whole reverse LSERK4 call (forward sweep and the "
                    "segment recompute, which keeps each step's stage inputs, included)")
    eng.close()
    return res


def bytes_alg_limit(K, Np, m):
    """HBM bytes of one limiter application: read and write the state
    (FP64) plus 140 B of geometry and connectivity per element (vertices,
    J, three normals and half lengths, three face codes)."""
    return K * (2 * Np * m * 8 + 140)


def run_limiter(torch, N, reps=50):
    """apl_limit in place on Apollo's 5-variable Euler, [-10,10]^2 with walls,
    512x512x2 jittered 0.15h seed 5, K=524288, isentropic-vortex state, at
    order N: applications/s and the algorithmic-byte rate."""
    from apollo_amd import engine as E
    mesh = E.structured_mesh(512, -10.0, 10.0, -10.0, 10.0, 0.15, 5, False)
    params = E.Params(N, E.EQN_EULER5, flux=E.FLUX_HLL, bcs=[(E.BC_EULER_WALL, [])],
                      limiter=(1.4, 1e-2, 1e-2))
    eng = E.Engine(mesh, params)
    stream = torch.cuda.Stream()
    E.lib().apl_set_stream(eng._h, ctypes.c_void_p(stream.cuda_stream))
    g, x, y = eng.geometry()
    q0 = np.ascontiguousarray(E.vortex_state(x, y, 0.0, m=5))
    with torch.cuda.stream(stream):
        qd = torch.from_numpy(q0.copy()).cuda()
        for _ in range(3):
            eng.limit(qd, qd)
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(reps):
            eng.limit(qd, qd)
        ev1.record(stream)
        torch.cuda.synchronize()
    sec = ev0.elapsed_time(ev1) * 1e-3 / reps
    K, Np, m = eng.K, eng.Np, eng.m
    eng.close()
    ach = bytes_alg_limit(K, Np, m) / sec
    return {"workload": "Hesthaven-Warburton Euler limiter (apl_limit, in place): Apollo's "
                        "5-variable Euler, [-10,10]^2 with walls, 512x512x2 jittered 0.15h seed 5, "
                        "K=524288, N=%d, isentropic-vortex state" % N,
            "reps": reps, "us_per_application": sec * 1e6, "dof_per_s": K * Np * m / sec,
            "roofline": {"bound": "hbm", "achieved": ach / 1e9, "peak": PEAK["hbm"] / 1e9,
                         "unit": "GB/s", "frac": ach / PEAK["hbm"],
                         "bytes_alg_per_element": bytes_alg_limit(1, Np, m)},
            "note": "time of apl_limit includes its host-side launch of three kernels and one "
                    "flag check per call"}


def _cpu_host():
    model = ""
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cores": usable}


def _oracle_case(w):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import cases
    case = cases.Case(w.mesh(), w.params())
    q = cases.to_oracle(w.eqn, w.initial_state(case.omesh.x, case.omesh.y))
    return case, q


def cpu_rhs_rate(name, nthreads, budget_s, max_reps=20, min_reps=1):
    """Oracle RHS evaluations of the full mesh (Apollo's 8-variable Maxwell /
    5-variable Euler layout), after one warm-up, timed for about budget_s."""
    from apollo_amd import workloads as W
    from oracle import oracle as O
    w = W.get(name)
    case, q = _oracle_case(w)
    t0 = time.perf_counter()
    O.rhs(case.omesh, case.ophys, 0.0, q, nthreads=nthreads)          # warm-up
    t1 = time.perf_counter() - t0
    reps = int(max(min_reps, min(max_reps, budget_s / max(t1, 1e-6))))
    t0 = time.perf_counter()
    for _ in range(reps):
        O.rhs(case.omesh, case.ophys, 0.0, q, nthreads=nthreads)
    el = time.perf_counter() - t0
    dofs = case.mesh.K * case.ops.Np * w.m
    return dict(value=dofs * reps / el, unit="DOF-updates/s", cores=nthreads, kind="port",
                sample="%d RHS evaluation(s) of the full mesh (K=%d, Apollo's %d-variable layout) "
                       "after 1 warm-up, %d OpenMP thread(s) over elements"
                       % (reps, case.mesh.K, q.shape[-1], nthreads),
                rhs_per_s=reps / el)


def cpu_lserk4_run(name, nthreads, steps):
    """The oracle's LSERK4 on the full mesh over `steps` steps at the
    workload's fixed dt (for vortex_64_N3 the whole 100-step run)."""
    from apollo_amd import workloads as W
    from oracle import oracle as O
    w = W.get(name)
    case, q = _oracle_case(w)
    _, ms = O.rhs(case.omesh, case.ophys, 0.0, q, nthreads=nthreads)
    dt = O.suggested_dt(case.omesh, w.cfl, ms)
    t0 = time.perf_counter()
    O.lserk4(case.omesh, case.ophys, 0.0, dt, steps, q, nthreads=nthreads)
    el = time.perf_counter() - t0
    dofs = case.mesh.K * case.ops.Np * w.m
    return dict(value=dofs * 5 * steps / el, unit="DOF-updates/s", cores=nthreads, kind="port",
                sample="%d LSERK4 steps (%d RHS evaluations) of the full mesh (K=%d, Apollo's "
                       "%d-variable layout), %d OpenMP thread(s)"
                       % (steps, 5 * steps, case.mesh.K, q.shape[-1], nthreads),
                seconds=el, steps=steps)


def cpu_baselines(nthreads):
    """The headline cpu_baseline and one per leg workload, each bounded to a
    few seconds of host time."""
    print("bench: cpu baseline, %d threads" % nthreads, file=sys.stderr, flush=True)
    head = cpu_rhs_rate(HEAD_WORKLOAD, nthreads, budget_s=12.0, min_reps=3)
    head["single_core"] = cpu_rhs_rate(HEAD_WORKLOAD, 1, budget_s=6.0, max_reps=3)
    head["host"] = _cpu_host()
    legs = {
        "vortex_64_N3": cpu_lserk4_run("vortex_64_N3", nthreads, 100),
        "vortex_64_N3_single_core": cpu_lserk4_run("vortex_64_N3", 1, 10),
        "vortex_512_N4": cpu_rhs_rate("vortex_512_N4", nthreads, budget_s=4.0, max_reps=4),
        "adjoint_vortex_256_N3": cpu_rhs_rate("adjoint_vortex_256_N3", nthreads, budget_s=3.0,
                                              max_reps=8),
        "tm_fourier_384_N8": cpu_rhs_rate("tm_fourier_384_N8", nthreads, budget_s=5.0, max_reps=3),
    }
    return head, legs


def measured_traffic(key):
    """HBM bytes per launch of a kernel from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor)."""
    try:
        with open(TRAFFIC) as f:
            return json.load(f)["kernels"][key]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def calibrate_copy(mib, reps=10):
    """A plain device-to-device copy of `mib` MiB, `reps` times: the reference
    kernel whose FETCH_SIZE / WRITE_SIZE counters scale the engine's."""
    import torch
    n = mib * 2 ** 20 // 8
    src = torch.arange(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(reps):
        dst.copy_(src)
    ev1.record()
    torch.cuda.synchronize()
    s = ev0.elapsed_time(ev1) * 1e-3 / reps
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "reps": reps,
                      "us_per_copy": s * 1e6, "gbs": 2 * n * 8 / s / 1e9}), flush=True)


def kernel_name(w, variant):
    """The engine's kernel for a variant: the nodal kernel for linear fluxes
    (algorithm 0), the vector-ALU cubature kernel at N <= 3 in FP64/FP32
    (algorithms 0, 1, 3; eight waves per group below 512 groups), else the
    MFMA cubature kernel."""
    N = w.order
    prec, algo = VARIANTS[variant]
    ty = {"fp64": "double", "fp32": "float", "bf16x3": "apl::Bf16x3"}[variant.split("_")[0]]
    if w.linear and algo == 5:
        return "rhs_valu_linear_kernel<%d,%d,%d,%s>" % (N, w.eqn, w.flux, ty)
    if w.linear and algo in NODAL_ALGOS:
        return "rhs_linear_kernel<%d,%d,%d,%s>" % (N, w.eqn, w.flux, ty)
    if N <= 3 and prec != 2 and algo != 2:
        nw = 8 if (w.K + 63) // 64 < 512 else 4
        return "rhs_valu_kernel<%d,%d,%d,%s,%d>" % (N, w.eqn, w.flux, ty, nw)
    return "rhs_kernel<%d,%d,%d,%s>" % (N, w.eqn, w.flux, ty)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default=HEAD_WORKLOAD)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-legs", action="store_true",
                    help="skip the vortex, N=8 Maxwell and limiter legs")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the fixed 500-step accuracy run of the headline workload")
    ap.add_argument("--only", default="",
                    help="time these variants only, comma-separated (fp64, fp32, bf16x3, "
                         "fp64_cub, fp32_cub, bf16x3_cub, fp64_mfma, fp32_mfma, fp64_valu, "
                         "fp32_valu)")
    ap.add_argument("--limiter-only", action="store_true", help="time the limiter legs only")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="run a plain MIB-MiB device copy instead (counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print("bench.py: the engine runs on one GPU only (the multi-GPU design is "
              "documented, not built); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    from apollo_amd import workloads as W
    assert torch.cuda.is_available()
    if args.limiter_only:
        print(json.dumps({"limiter_N4": run_limiter(torch, 4), "limiter_N1": run_limiter(torch, 1)}),
              flush=True)
        return
    w = W.get(args.workload)
    names = ["fp64", "fp32", "bf16x3"] + (["fp64_cub", "fp32_cub", "bf16x3_cub"] if w.linear else [])
    if args.only:
        names = args.only.split(",")
    res = {}
    for v in names:
        print("bench: %s %s steps=%d" % (args.workload, v, args.steps), file=sys.stderr, flush=True)
        res[v] = run_gpu(w, v, args.steps, args.warmup, torch)
    out = {v: variant_figures(w, v, r, args.steps) for v, r in res.items()}
    for v, r in res.items():
        if v != "fp64" and "fp64" in res:
            out[v]["speedup_vs_fp64"] = res["fp64"]["gpu_ms"] / r["gpu_ms"]
            out[v]["l2_vs_fp64_at_timed_steps"] = rel_l2(r["qT"], res["fp64"]["qT"])
    info = next(iter(res.values()))["info"]
    K, N, Np, Nc, Ng, m = (info[k] for k in ("K", "N", "Np", "Nc", "Ng", "m"))
    head = "fp64" if "fp64" in out else names[0]
    h = out[head]
    line = {
        "metric": METRIC,
        "value": h["dof_updates_per_s"],
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": h["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32", "bf16x3": "bf16x3"}[head.split("_")[0]],
        "data": "synthetic",
        "config": {"workload": w.desc, "name": w.name, "K": K, "N": N, "Np": Np, "Nc": Nc,
                   "Ng": Ng, "m": m, "dt": info["dt"], "rhs_evals_per_step": 5,
                   "precision": "FP64 everywhere (the reference's precision); nodal form of the "
                                "linear flux, identical to Apollo's cubature form to rounding"
                   if head == "fp64" else head},
        "variants": out,
    }
    # Roofline of the dominant kernel (the head variant's fused RHS + stage).
    # The nodal form executes ~2.5 flop per algorithmic byte, far below the
    # FP64 and FP32 ridges (9.8 and 19.7 flop/B), so its bound is HBM; the
    # cubature form is reported against its MFMA peak beside it.
    bpl = bytes_alg_stage(K, Np, m)
    lin = w.linear and VARIANTS[head][1] in NODAL_ALGOS
    kname = kernel_name(w, head)
    if lin:
        ach = bpl / (h["kernel_us"] * 1e-6)
        line["roofline"] = {"bound": "hbm", "achieved": ach / 1e9, "peak": PEAK["hbm"] / 1e9,
                            "unit": "GB/s", "frac": ach / PEAK["hbm"],
                            "traffic": measured_traffic("tm_cavity_256_N5/%s" % head),
                            "kernel": kname + " (fused RHS + LSERK4 stage), avg over %d launches"
                                      % (5 * args.steps),
                            "bytes_per_launch": bpl}
    else:
        pk = PEAK["fp64" if head.startswith("fp64") else "fp32"]
        ach = h["tflops_alg"] * 1e12
        line["roofline"] = {"bound": "mfma", "achieved": ach / 1e12, "peak": pk / 1e12,
                            "unit": "TFLOP/s", "frac": ach / pk, "traffic": None,
                            "kernel": kname + " (fused RHS + LSERK4 stage), avg over %d launches"
                                      % (5 * args.steps),
                            "flop_per_launch": flop_alg(K, Np, Nc, Ng, m)}
    line["roofline_variants"] = {}
    for v in out:
        if v == head:
            continue
        lv = w.linear and VARIANTS[v][1] in NODAL_ALGOS
        if lv:
            line["roofline_variants"][v] = {
                "bound": "hbm", "achieved": out[v]["gbs_alg"], "peak": PEAK["hbm"] / 1e9,
                "unit": "GB/s", "frac": out[v]["hbm_frac_alg"],
                "traffic": measured_traffic("tm_cavity_256_N5/%s" % v), "kernel": kernel_name(w, v)}
        else:
            pk = PEAK["fp64" if v.startswith("fp64") else ("fp32" if v.startswith("fp32") else "bf16")]
            line["roofline_variants"][v] = {
                "bound": "mfma", "achieved_exec": out[v]["tflops_exec"], "peak": pk / 1e12,
                "unit": "TFLOP/s", "frac": out[v]["tflops_exec"] * 1e12 / pk,
                "traffic": measured_traffic("tm_cavity_256_N5/%s" % v), "kernel": kernel_name(w, v)}
    if "fp32" in out and "fp64" in out:
        line["speedup_fp32_vs_fp64"] = out["fp32"]["speedup_vs_fp64"]
    if not args.no_accuracy and not args.only:
        # the headline workload's full 500-step run, whatever --steps is:
        # errors against the analytic cavity mode and the FP64 run
        acc = {}
        runs = {}
        for v in ("fp64", "fp32", "bf16x3", "fp64_cub"):
            if v not in VARIANTS or (v.endswith("_cub") and not w.linear):
                continue
            print("bench: accuracy %s %s steps=%d" % (w.name, v, w.steps), file=sys.stderr, flush=True)
            runs[v] = run_gpu(w, v, w.steps, 1, torch)
        for v, r in runs.items():
            T = w.steps * r["dt"]
            ex = w.exact_state(r["x"], r["y"], T)
            acc[v] = {"l2_vs_analytic": rel_l2(r["qT"], ex) if ex is not None else None,
                      "l2_vs_fp64": rel_l2(r["qT"], runs["fp64"]["qT"]) if v != "fp64" else 0.0}
        line["accuracy"] = {"steps": w.steps, "T": w.steps * runs["fp64"]["dt"], "variants": acc,
                            "note": "relative L2 over all variables after the workload's full "
                                    "step count, independent of --steps"}
    if not args.no_adjoint and not args.only:
        print("bench: adjoint", file=sys.stderr, flush=True)
        line["adjoint"] = run_adjoint(torch)
    if not args.no_legs and not args.only:
        three = ["fp64", "fp32", "bf16x3"]
        line["legs"] = {
            "vortex_64_N3": run_leg("vortex_64_N3", three, torch),
            "vortex_512_N4": run_leg("vortex_512_N4", three, torch),
            "tm_fourier_384_N8": run_leg("tm_fourier_384_N8", three, torch),
        }
        print("bench: limiter", file=sys.stderr, flush=True)
        line["legs"]["limiter_N4"] = run_limiter(torch, 4)
        line["legs"]["limiter_N1"] = run_limiter(torch, 1)
    if not args.no_cpu_baseline:
        nth = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        head_cpu, legs_cpu = cpu_baselines(nth)
        line["cpu_baseline"] = head_cpu
        line["cpu_baseline_legs"] = legs_cpu
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
