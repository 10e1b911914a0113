#!/usr/bin/env python3
"""Benchmark: DG RHS DOF-updates/s on one MI355X.

Default workload (Maxwell TM cavity, the configuration the headline metric
is quoted on): 256 x 256 quads split into 2 triangles (K = 131072), vertices
jittered +-0.1 h (seed 1), N = 5 (Np = 21), PEC walls with Apollo's
transverse-magnetic boundary state, upwind flux, Ez = sin(3 pi x) sin(4 pi y),
H = 0, LSERK4 with Apollo's dt formula at the deck CFL 0.5/(2N+1), fixed dt,
500 steps.

A "step" is one LSERK4 step = 5 RHS evaluations fused with the stage
updates, replayed as one hipGraph.  DOF-updates/s = K Np m x (RHS
evaluations) / time with m = 3 (Ez, Bx, By).  `value` is the FP32-MFMA
engine; the FP64 engine and Apollo's over-integrated cubature form are timed
the same way and reported beside it, each with its L2 error against the
analytic cavity mode (and FP32 against FP64).  State lives in HBM before the
timed region starts.

The adjoint leg (Euler isentropic vortex, [-5,5]^2 periodic, 256x256x2
jittered 0.1h seed 3, N = 3, 200 steps, checkpoint every 20) times the
reverse LSERK4 and checks dJ/dq0 against central finite differences along 8
Gaussian directions (eps = 1e-6; smooth random fields, and white noise
reported beside them).  Two more legs time FP64, FP32-MFMA and
split-BF16x3 on the Euler vortex at N = 4 (512x512x2, K = 524288, 200 steps,
density error against the FP64 run) and on Maxwell TM at N = 8 (384x384x2,
K = 294912, random 16-mode Ez, 300 steps, MFMA utilisation).

    python bench.py [--gpus 1] [--steps 500] [--warmup 20] [--workload c2]
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


def workload(name):
    """Mesh + parameters + initial state of a benchmark workload."""
    from apollo_amd import engine as E
    if name == "c2":
        n, N = 256, 5
        mesh = E.structured_mesh(n, 0.0, 1.0, 0.0, 1.0, 0.1, 1, False)
        params = dict(order=N, eqn=E.EQN_MAXWELL_TM3, flux=E.FLUX_UPWIND,
                      bcs=[(E.BC_MAXWELL_TM, [])])
        cfl = 0.5 / (2 * N + 1)
        desc = ("2D Maxwell TM (Ez,Bx,By), unit square, PEC walls (Apollo TM boundary state), "
                "256x256x2 jittered +-0.1h seed 1, K=131072, N=5, upwind flux, LSERK4")
        return dict(name=name, mesh=mesh, params=params, cfl=cfl, desc=desc, m=3, analytic="cavity")
    if name == "c5":
        n, N = 384, 8
        mesh = E.structured_mesh(n, 0.0, 1.0, 0.0, 1.0, 0.1, 4, False)
        params = dict(order=N, eqn=E.EQN_MAXWELL_TM3, flux=E.FLUX_UPWIND,
                      bcs=[(E.BC_MAXWELL_TM, [])])
        cfl = 0.5 / (2 * N + 1)
        desc = ("2D Maxwell TM, 384x384x2 jittered seed 4, K=294912, N=8, upwind flux, "
                "random 16-mode Ez")
        return dict(name=name, mesh=mesh, params=params, cfl=cfl, desc=desc, m=3, analytic="fourier")
    if name == "c3":
        n, N = 512, 4
        mesh = E.structured_mesh(n, -10.0, 10.0, -10.0, 10.0, 0.15, 2, True)
        params = dict(order=N, eqn=E.EQN_EULER4, bcs=[])
        desc = ("2D Euler isentropic vortex, [-10,10]^2 periodic, 512x512x2 jittered 0.15h seed 2, "
                "K=524288, N=4, LF flux")
        return dict(name=name, mesh=mesh, params=params, cfl=0.5, desc=desc, m=4, analytic="vortex",
                    L=20.0)
    if name == "c4":
        n, N = 256, 3
        mesh = E.structured_mesh(n, -5.0, 5.0, -5.0, 5.0, 0.1, 3, True)
        params = dict(order=N, eqn=E.EQN_EULER4, bcs=[])
        desc = ("adjoint: 2D Euler isentropic vortex, [-5,5]^2 periodic, 256x256x2 jittered 0.1h "
                "seed 3, K=131072, N=3, LF flux, J = int_{r<1} rho^2 at step 200, checkpoint 20")
        return dict(name=name, mesh=mesh, params=params, cfl=0.5, desc=desc, m=4, analytic="vortex",
                    L=10.0)
    if name in ("c6", "c6n1"):
        # c6n1: the same at N = 1, the order of Apollo's shock decks
        n, N = 512, (4 if name == "c6" else 1)
        mesh = E.structured_mesh(n, -10.0, 10.0, -10.0, 10.0, 0.15, 5, False)
        params = dict(order=N, eqn=E.EQN_EULER5, flux=E.FLUX_HLL, bcs=[(E.BC_EULER_WALL, [])],
                      limiter=(1.4, 1e-2, 1e-2))
        desc = ("Hesthaven-Warburton Euler limiter (apl_limit, in place): Apollo's 5-variable "
                "Euler, [-10,10]^2 with walls, 512x512x2 jittered 0.15h seed 5, K=524288, N=%d, "
                "isentropic-vortex state" % N)
        return dict(name=name, mesh=mesh, params=params, cfl=0.5, desc=desc, m=5, analytic="vortex5",
                    L=20.0)
    raise SystemExit("unknown workload %s" % name)


def initial_state(w, x, y):
    from apollo_amd import engine as E
    if w["analytic"] == "cavity":
        return E.tm_cavity_mode(x, y, 0.0, m=3)
    if w["analytic"] == "fourier":
        return E.random_fourier_ez(x, y, 16, 4, m=3)
    return E.periodic_vortex_state(x, y, 0.0, w["L"], m=4)


def exact_state(w, x, y, t):
    from apollo_amd import engine as E
    if w["analytic"] == "cavity":
        return E.tm_cavity_mode(x, y, t, m=3)
    if w["analytic"] == "vortex":
        return E.periodic_vortex_state(x, y, t, w["L"], m=4)
    return None


def flop_alg(K, N, Np, Nc, Ng, m):
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


def engine_on_stream(w, precision, algorithm, torch):
    from apollo_amd import engine as E
    eng = E.Engine(w["mesh"], E.Params(precision=precision, algorithm=algorithm, **w["params"]))
    stream = torch.cuda.Stream()
    E.lib().apl_set_stream(eng._h, ctypes.c_void_p(stream.cuda_stream))
    return eng, stream


def fixed_dt(w, eng, q0):
    """Apollo's dt formula, evaluated once from q0 and then held fixed."""
    if w["analytic"] in ("cavity", "fourier"):
        return eng.suggested_dt(w["cfl"], 1.0)      # c0 = 1
    tmp = np.zeros_like(q0)
    return eng.rhs(0.0, q0, tmp, cfl=w["cfl"])


def run_gpu(w, precision, steps, warmup, torch, algorithm=0):
    eng, stream = engine_on_stream(w, precision, algorithm, torch)
    g, x, y = eng.geometry()
    q0 = np.ascontiguousarray(initial_state(w, x, y))
    dt = fixed_dt(w, eng, q0)
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
    return dict(wall=wall, gpu_ms=gpu_ms, qT=qT, x=x, y=y, info=info, dt=dt)


def run_adjoint(torch, steps=200, ckpt=20, ndir=8, eps=1e-6):
    """Reverse LSERK4 on the adjoint workload: timing and the finite-difference
    check of dJ/dq0 along `ndir` Gaussian directions."""
    from apollo_amd import engine as E
    w = workload("c4")
    eng, stream = engine_on_stream(w, E.PREC_FP64, E.ALGO_AUTO, torch)
    g, x, y = eng.geometry()
    q0 = np.ascontiguousarray(initial_state(w, x, y))
    dt = fixed_dt(w, eng, q0)
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
        gen = torch.Generator(device="cpu").manual_seed(8)

        def fd_err(d):
            fd = (objective(q0d + eps * d) - objective(q0d - eps * d)) / (2 * eps)
            ad = float((grad * d).sum())
            return abs(fd - ad) / abs(ad)

        # white-noise directions: N(0,1) per degree of freedom
        errs = [fd_err(torch.randn(q0d.shape, generator=gen, dtype=torch.float64).cuda())
                for _ in range(ndir)]
        # smooth Gaussian directions: 16 random Fourier modes of the periodic
        # domain per variable, N(0,1) amplitudes
        rng = np.random.default_rng(8)
        L = w["L"]
        errs_smooth = []
        for _ in range(ndir):
            d = np.zeros(q0.shape)
            for v in range(q0.shape[-1]):
                for _ in range(16):
                    kx, ky = rng.integers(-3, 4, size=2)
                    ph = rng.uniform(0, 2 * np.pi)
                    d[..., v] += rng.standard_normal() * np.cos(2 * np.pi * (kx * x + ky * y) / L + ph)
            errs_smooth.append(fd_err(torch.from_numpy(d).cuda()))
    dofs = eng.K * eng.Np * w["m"]
    res = dict(workload=w["desc"], steps=steps, ckpt_every=ckpt, dt=dt, J=J,
               seconds=t_adj,
               adjoint_dof_updates_per_s=dofs * 5 * steps / t_adj,
               adjoint_rhs_us=t_arhs * 1e6,
               adjoint_rhs_dof_updates_per_s=dofs / t_arhs,
               fd_directions=ndir, fd_eps=eps, fd_max_rel_err=max(errs_smooth),
               fd_max_rel_err_white_noise=max(errs),
               fd_note="fd_max_rel_err: smooth Gaussian random-field directions (16 Fourier modes "
                       "per variable, N(0,1) amplitudes); white-noise directions (N(0,1) per degree "
                       "of freedom) move the two traces of a face independently by ~eps, which "
                       "straddles the Lax-Friedrichs max(c-, c+) at part of the face points, so "
                       "central differences there see the average of the two branches",
               note="adjoint DOF-updates/s counts the 5 x steps adjoint RHS evaluations of the "
                    "reverse sweep over the whole reverse LSERK4 call (forward sweep, segment "
                    "recompute and stage rebuild included)")
    eng.close()
    return res


def run_leg(name, steps, variants, torch):
    """A secondary workload timed like the headline one: DOF-updates/s and
    the L2 error of each variant against the FP64 run after `steps` steps
    (density only for Euler)."""
    w = workload(name)
    res = {}
    for vname, prec, algo in variants:
        print("bench: %s %s steps=%d" % (name, vname, steps), file=sys.stderr, flush=True)
        res[vname] = run_gpu(w, prec, steps, 5, torch, algo)
    info = res["fp64"]["info"]
    K, N, Np, Nc, Ng, m = (info[k] for k in ("K", "N", "Np", "Nc", "Ng", "m"))
    ref = res["fp64"]["qT"]
    sel = (lambda q: q[..., 0]) if m == 4 else (lambda q: q)
    out = {"workload": w["desc"], "steps": steps, "dt": info["dt"], "variants": {}}
    for vname, prec, algo in variants:
        r = res[vname]
        kern_s = r["gpu_ms"] * 1e-3 / (5 * steps)
        lin = m == 3 and algo == 0
        d = dict(dof_updates_per_s=K * Np * m / kern_s, kernel_us=kern_s * 1e6,
                 form="nodal" if lin else "cubature",
                 tflops_alg=flop_alg(K, N, Np, Nc, Ng, m) / kern_s / 1e12,
                 tflops_exec=flop_exec(lin, K, N, Np, Nc, Ng, m) / kern_s / 1e12,
                 gbs_alg=bytes_alg_stage(K, Np, m) / kern_s / 1e9)
        d["hbm_frac_alg"] = d["gbs_alg"] * 1e9 / PEAK["hbm"]
        if vname != "fp64":
            d["l2_vs_fp64"] = float(np.linalg.norm((sel(r["qT"]) - sel(ref)).ravel())
                                    / np.linalg.norm(sel(ref).ravel()))
        ex = exact_state(w, r["x"], r["y"], steps * r["dt"])
        if ex is not None:
            d["l2_vs_analytic"] = float(np.linalg.norm((sel(r["qT"]) - sel(ex)).ravel())
                                        / np.linalg.norm(sel(ex).ravel()))
        pk = PEAK["fp64" if vname == "fp64" else ("fp32" if vname == "fp32" else "bf16")]
        d["mfma_frac_exec"] = d["tflops_exec"] * 1e12 / pk
        out["variants"][vname] = d
    return out


def bytes_alg_limit(K, Np, m):
    """HBM bytes of one limiter application: read and write the state
    (FP64) plus 140 B of geometry and connectivity per element (vertices,
    J, three normals and half lengths, three face codes)."""
    return K * (2 * Np * m * 8 + 140)


def run_limiter(torch, reps=50, name="c6"):
    """Time apl_limit in place on a limiter workload (c6: N = 4, c6n1:
    N = 1) with HIP events on the engine's stream; report applications/s,
    DOF/s and the algorithmic-byte rate against the HBM peak."""
    from apollo_amd import engine as E
    w = workload(name)
    eng, stream = engine_on_stream(w, E.PREC_FP64, E.ALGO_AUTO, torch)
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
    return {"workload": w["desc"], "reps": reps, "us_per_application": sec * 1e6,
            "dof_per_s": K * Np * m / sec,
            "roofline": {"bound": "hbm", "achieved": ach / 1e9, "peak": PEAK["hbm"] / 1e9,
                         "unit": "GB/s", "frac": ach / PEAK["hbm"],
                         "bytes_alg_per_element": bytes_alg_limit(1, Np, m)},
            "note": "time of apl_limit includes its host-side launch of three kernels and one "
                    "flag check per call"}


def cpu_baseline(w, nthreads, budget_s=15.0):
    """Oracle RHS (Apollo's loop restated in C, Apollo's 8-variable Maxwell /
    5-variable Euler layout) on the host cores, on the full mesh."""
    from oracle import oracle as O
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import cases
    from apollo_amd import engine as E
    params = E.Params(**w["params"])
    case = cases.Case(w["mesh"], params)
    q = cases.to_oracle(params.eqn, initial_state(w, case.omesh.x, case.omesh.y))
    O.rhs(case.omesh, case.ophys, 0.0, q, nthreads=nthreads)       # warm-up
    t0 = time.perf_counter()
    O.rhs(case.omesh, case.ophys, 0.0, q, nthreads=nthreads)
    t1 = time.perf_counter() - t0
    reps = int(max(3, min(20, budget_s / max(t1, 1e-6))))
    for _ in range(2):
        O.rhs(case.omesh, case.ophys, 0.0, q, nthreads=nthreads)
    t0 = time.perf_counter()
    for _ in range(reps):
        O.rhs(case.omesh, case.ophys, 0.0, q, nthreads=nthreads)
    el = time.perf_counter() - t0
    dofs = case.mesh.K * case.ops.Np * w["m"]
    # one core: the same RHS, fewer repetitions
    O.rhs(case.omesh, case.ophys, 0.0, q, nthreads=1)
    reps1 = 3
    t0 = time.perf_counter()
    for _ in range(reps1):
        O.rhs(case.omesh, case.ophys, 0.0, q, nthreads=1)
    el1 = time.perf_counter() - t0
    model = ""
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return dict(value=dofs * reps / el, unit="DOF-updates/s", cores=nthreads, kind="port",
                sample="%d RHS evaluations of the full mesh (K=%d, Apollo's %d-variable layout) "
                       "after 3 warm-ups, OpenMP over elements" % (reps, case.mesh.K, q.shape[-1]),
                rhs_per_s=reps / el,
                single_core={"value": dofs * reps1 / el1, "rhs_per_s": reps1 / el1,
                             "sample": "%d RHS evaluations, 1 thread" % reps1},
                host={"cpu_model": model, "nproc": os.cpu_count(), "usable_cores": usable})


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-legs", action="store_true",
                    help="skip the Euler N=4 and Maxwell N=8 workloads")
    ap.add_argument("--only", default="",
                    help="time one variant only (fp64, fp32, bf16x3, fp64_cub, fp32_cub, bf16x3_cub)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="run a plain MIB-MiB device copy instead (counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the multi-GPU design is "
              "documented, not built); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    from apollo_amd import engine as E
    assert torch.cuda.is_available()
    if args.workload == "c6":
        print(json.dumps({"limiter_c6": run_limiter(torch),
                          "limiter_c6_n1": run_limiter(torch, name="c6n1")}), flush=True)
        return
    w = workload(args.workload)
    linear_ok = w["m"] == 3
    res = {}
    variants = [("fp64", E.PREC_FP64, E.ALGO_AUTO), ("fp32", E.PREC_FP32_MFMA, E.ALGO_AUTO),
                ("bf16x3", E.PREC_BF16X3, E.ALGO_AUTO)]
    if linear_ok:
        # Maxwell: the default is the nodal form of the linear flux; also time
        # Apollo's over-integrated cubature form
        variants += [("fp64_cub", E.PREC_FP64, E.ALGO_CUBATURE),
                     ("fp32_cub", E.PREC_FP32_MFMA, E.ALGO_CUBATURE),
                     ("bf16x3_cub", E.PREC_BF16X3, E.ALGO_CUBATURE)]
    for name, prec, algo in variants:
        if args.only and args.only != name:
            continue
        print("bench: %s %s steps=%d" % (args.workload, name, args.steps), file=sys.stderr, flush=True)
        r = run_gpu(w, prec, args.steps, args.warmup, torch, algo)
        r["linear"] = linear_ok and algo == E.ALGO_AUTO
        res[name] = r
    info = next(iter(res.values()))["info"]
    K, N, Np, Nc, Ng, m = (info[k] for k in ("K", "N", "Np", "Nc", "Ng", "m"))
    rhs_evals = 5 * args.steps
    out = {}
    for name, r in res.items():
        ms = r["gpu_ms"]
        kern_s = ms * 1e-3 / rhs_evals
        dofs = K * Np * m * rhs_evals / (ms * 1e-3)
        T = args.steps * r["dt"]
        ex = exact_state(w, r["x"], r["y"], T)
        err_exact = None
        if ex is not None:
            err_exact = float(np.linalg.norm((r["qT"] - ex).ravel()) / np.linalg.norm(ex.ravel()))
        out[name] = dict(dof_updates_per_s=dofs, ms_per_step=ms / args.steps,
                         kernel_us=kern_s * 1e6, l2_vs_analytic=err_exact, T=T,
                         form="nodal" if r["linear"] else "cubature",
                         tflops_alg=flop_alg(K, N, Np, Nc, Ng, m) / kern_s / 1e12,
                         tflops_exec=flop_exec(r["linear"], K, N, Np, Nc, Ng, m) / kern_s / 1e12,
                         gbs_alg=bytes_alg_stage(K, Np, m) / kern_s / 1e9)
        if name != "fp64" and "fp64" in res:
            a, b = r["qT"], res["fp64"]["qT"]
            out[name]["l2_vs_fp64"] = float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))
    head = "fp32" if "fp32" in out else next(iter(out))
    h = out[head]
    kname = "%s<%d,%d,%d,%s>" % ("rhs_linear_kernel" if res[head]["linear"] else "rhs_kernel", N,
                                 w["params"]["eqn"], w["params"].get("flux", 0),
                                 "double" if head.startswith("fp64") else "float")
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
        "dtype": "f64" if head.startswith("fp64") else "f32",
        "data": "synthetic",
        "config": {"workload": w["desc"], "K": K, "N": N, "Np": Np, "Nc": Nc, "Ng": Ng,
                   "m": m, "dt": info["dt"], "rhs_evals_per_step": 5,
                   "precision": "FP32-MFMA contractions, FP64 state/geometry/stage update"
                   if head == "fp32" else head},
        "variants": out,
        "speedup_fp32_vs_fp64": (out["fp32"]["dof_updates_per_s"] / out["fp64"]["dof_updates_per_s"])
        if ("fp32" in out and "fp64" in out) else None,
    }
    # Roofline of the dominant kernel (the head variant's fused RHS + stage):
    # the nodal form executes ~2.5 flop per algorithmic byte, far below the
    # MFMA ridge, so its bound is HBM; the cubature form is reported against
    # its MFMA peak beside it.
    bpl = bytes_alg_stage(K, Np, m)
    if res[head]["linear"]:
        ach = bpl / (h["kernel_us"] * 1e-6)
        line["roofline"] = {"bound": "hbm", "achieved": ach / 1e9, "peak": PEAK["hbm"] / 1e9,
                            "unit": "GB/s", "frac": ach / PEAK["hbm"],
                            "traffic": measured_traffic("%s/%s" % (w["name"], head)),
                            "kernel": kname + " (fused RHS + LSERK4 stage), avg over %d launches"
                                      % rhs_evals,
                            "bytes_per_launch": bpl}
    else:
        pk = PEAK["fp64" if head.startswith("fp64") else "fp32"]
        ach = h["tflops_alg"] * 1e12
        line["roofline"] = {"bound": "mfma", "achieved": ach / 1e12, "peak": pk / 1e12,
                            "unit": "TFLOP/s", "frac": ach / pk,
                            "traffic": measured_traffic("%s/%s" % (w["name"], head)),
                            "kernel": kname + " (fused RHS + LSERK4 stage), avg over %d launches"
                                      % rhs_evals,
                            "flop_per_launch": flop_alg(K, N, Np, Nc, Ng, m)}
    for cub in ("fp64_cub", "fp32_cub", "bf16x3_cub"):
        if cub in out:
            pk = PEAK["fp64" if cub == "fp64_cub" else ("fp32" if cub == "fp32_cub" else "bf16")]
            ach = out[cub]["tflops_exec"] * 1e12
            line.setdefault("roofline_cubature", {})[cub] = {
                "bound": "mfma", "achieved_exec": ach / 1e12, "peak": pk / 1e12, "unit": "TFLOP/s",
                "frac": ach / pk, "traffic": measured_traffic("%s/%s" % (w["name"], cub))}
    if args.workload == "c2" and not args.no_adjoint and not args.only:
        print("bench: adjoint (c4)", file=sys.stderr, flush=True)
        line["adjoint"] = run_adjoint(torch)
    if args.workload == "c2" and not args.no_legs and not args.only:
        three = [("fp64", E.PREC_FP64, E.ALGO_AUTO), ("fp32", E.PREC_FP32_MFMA, E.ALGO_AUTO),
                 ("bf16x3", E.PREC_BF16X3, E.ALGO_AUTO)]
        line["euler_c3"] = run_leg("c3", 200, three, torch)
        line["maxwell_c5"] = run_leg("c5", 300, three, torch)
        print("bench: limiter (c6)", file=sys.stderr, flush=True)
        line["limiter_c6"] = run_limiter(torch)
        line["limiter_c6_n1"] = run_limiter(torch, name="c6n1")
    if not args.no_cpu_baseline:
        nth = min(16, os.cpu_count() or 1)
        print("bench: cpu baseline, %d threads" % nth, file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(w, nth)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
