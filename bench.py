"""Benchmark: FP64 zone-updates/s of the fused OpenMHD step on one MI355X.

Workload (the configuration the metric is quoted on): 2D Kelvin-Helmholtz
layer, 4096 x 4096 cells, FP64, x periodic, y walls, minmod + HLLD + GLM +
TVD-RK2, CFL 0.4.  A "step" is one full TVD-RK2 step of the whole grid (two
fused launches + the CFL reduction); the state is resident in HBM before the
timed region starts.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line.  Progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
BYTES_PER_ZONE_UPDATE = 360  # 9 doubles: read U^n, write U*, read U^n + U*, write U^{n+1}
BYTES_STAGE = {1: 144, 2: 216}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse_traffic():
    """HBM bytes per stage-2 launch from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration)."""
    f = REPO / "profiles" / "pmc_sweep2d_kh4096.txt"
    if not f.exists():
        return None
    m = re.search(r"^traffic_bytes_per_launch_stage2:\s*([0-9.eE+]+)", f.read_text(), re.M)
    return float(m.group(1)) if m else None


def stream_copy_gbs(nbytes: int = 1 << 30, reps: int = 20) -> float:
    """What HBM delivers to a plain device copy on this GPU: bytes read plus
    bytes written per second, over `reps` copies of `nbytes` timed with HIP
    events (after two untimed copies)."""
    import torch

    src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    gbs = 2 * nbytes * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del src, dst
    torch.cuda.empty_cache()
    return gbs


def cpu_model() -> str:
    """The host CPU's model name and logical CPU count, from /proc/cpuinfo."""
    try:
        names = [ln.split(":", 1)[1].strip() for ln in open("/proc/cpuinfo") if ln.startswith("model name")]
        return f"{names[0]} ({len(names)} logical CPUs on the host)" if names else "unknown"
    except OSError:
        return "unknown"


def cpu_baseline(n: int, sample_steps: int):
    """The reference CPU build (oracle/_ref, the project's Fortran driver over the
    reference's own routines, OpenMP) timed on this host on a bounded sample of
    the same workload; falls back to the C restatement if _ref is absent."""
    import resource

    threads = min(16, os.cpu_count() or 1)
    drv = REPO / "oracle" / "_ref" / "main2d_ref_omp"
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))

    def big_stack():  # set_dt/glm_ss2 keep ix*jx automatic arrays on the stack
        resource.setrlimit(resource.RLIMIT_STACK, (resource.getrlimit(resource.RLIMIT_STACK)[1],) * 2)

    def ref_rate(nthreads, steps):
        cmd = [str(drv), "--kh", str(n), str(steps + 1), "--time"]
        log(f"cpu baseline ({nthreads} threads):", " ".join(cmd))
        out = subprocess.run(cmd, env=dict(env, OMP_NUM_THREADS=str(nthreads)), capture_output=True, text=True,
                             timeout=900, preexec_fn=big_stack)
        m = re.search(r"zone_updates_per_s\s+([0-9.eE+]+)", out.stdout)
        if out.returncode == 0 and m:
            return float(m.group(1))
        log("reference driver failed:", out.stdout[-500:], out.stderr[-500:])
        return None

    if drv.exists():
        v = ref_rate(threads, sample_steps)
        if v is not None:
            v1 = ref_rate(1, 1)
            return {"value": v, "unit": "zone-updates/s", "cores": threads, "kind": "reference",
                    "single_thread": v1, "cpu": cpu_model(),
                    "sample": f"KH {n}x{n} minmod, steps 2..{sample_steps + 1} of a {sample_steps + 1}-step run "
                              f"(the Fortran driver's own KH layer init), the reference's routines compiled unmodified with "
                              f"flang -O2 -fopenmp, {threads} OpenMP threads; single_thread: step 2 of a 2-step "
                              f"run on 1 thread"}
    # C restatement (port)
    sys.path.insert(0, str(REPO / "tests"))
    from omhd_testlib import Oracle, make_problem

    from openmhd_amd.problems import kelvin_helmholtz

    U, dx = kelvin_helmholtz(n, n, seed=1)
    orc = Oracle()
    p = make_problem(n + 2, n + 2, dx=dx, cfl=0.4, lm_type=1, bc=("periodic", "wall", "periodic"))
    orc.bc_u(U, p)
    orc.run(p, U, 1)
    t0 = time.perf_counter()
    orc.run(p, U, sample_steps)
    dt = time.perf_counter() - t0
    return {"value": n * n * sample_steps / dt, "unit": "zone-updates/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(),
            "sample": f"KH {n}x{n} minmod, {sample_steps} steps after one untimed step, C restatement "
                      f"(gcc -O2 -fopenmp, {threads} threads)"}


def bench_ensemble(a, steps=None):
    """256 independent KH runs, sides log-uniform in [64, 1024],
    300 steps, each with its own dt.  Batched (one launch per RK stage over all
    members, autotuned tiles; host dt and device dt + hipGraph) against the same
    members run one after another with the single-grid step.  Returns the
    figures as a dict."""
    import torch

    from openmhd_amd import Ensemble, Solver, make_config, problems

    torch.cuda.set_device(0)
    params = problems.kh_ensemble(a.members)
    states, cfgs = [], []
    for p in params:
        U, dx = problems.kh_member(p)
        states.append(U)
        cfgs.append(make_config(p["nx"], p["ny"], dx, cfl=0.4, lm_type=1, bc=("periodic", "wall")))
    cells = sum(p["nx"] * p["ny"] for p in params)
    log(f"ensemble: {len(params)} members, {cells} cells")
    steps = a.steps if steps is None else steps
    seq = 0.0
    for U, c in zip(states, cfgs):
        s = Solver(c.nx, c.ny, c.dx, cfl=0.4, lm_type=1, bc=("periodic", "wall"))
        s.upload(U)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.step(steps)
        seq += time.perf_counter() - t0
        s.close()
    log(f"one after another: {seq:.3f} s")
    out = {"members": len(params), "cells": cells, "steps": steps, "sequential_s": seq}
    for mode in ("host_dt", "device_dt_graph"):
        e = Ensemble(cfgs)
        for i, U in enumerate(states):
            e.upload(i, U)
        if mode != "host_dt":
            e.set_mode(device_dt=True, graph=True)
        rows, grid = e.autotune(3)
        e.step(2)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e.step(steps)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        e.close()
        log(f"batched {mode}: {el:.3f} s (rows {rows}, grid {grid})")
        out[mode] = {"seconds": el, "zone_updates_per_s": cells * steps / el, "speedup_vs_sequential": seq / el,
                     "tile_rows": rows, "grid": grid}
    out["sequential_zone_updates_per_s"] = cells * steps / seq
    return {"workload": f"KH ensemble, {len(params)} members, sides 64-1024, {steps} steps", **out}


def bench_3d(a, steps=None, warmup=None, grad_steps=100, every=10, fd=None):
    """3D resistive Harris sheet 512x256x256 (Bx = tanh(y), rho =
    0.2 + cosh^-2(y)/0.2, eta = 1e-3 + 1e-2 exp(-(x^2+y^2)/4), 1e-3 velocity
    noise), HLLD + resistive flux + GLM + TVD-RK2, minmod, x/z periodic, y
    walls: forward zone-updates/s over `steps` steps; then (unless --no-grad)
    the gradient of the reconnected flux at the end of `grad_steps` (100) steps
    with respect to (eta1, w) by the adjoint with checkpoints every `every`
    (10) steps, checked (unless --no-fd) by central differences with the
    frozen dt sequence.  Returns the figures as a dict."""
    import numpy as np
    import torch

    from openmhd_amd import Adjoint, Solver, problems

    torch.cuda.set_device(0)
    steps = a.steps if steps is None else steps
    warmup = a.warmup if warmup is None else warmup
    fd_on = (not a.no_fd) if fd is None else fd
    nx, ny, nz, dx = a.nx3, a.ny3, a.nz3, 0.125
    eta0, eta1, w = 1e-3, 1e-2, 2.0
    U, origin = problems.harris_sheet_3d(nx, ny, nz, dx)
    s = Solver(nx, ny, dx, nz=nz, cfl=0.4, lm_type=1, bc=("periodic", "wall", "periodic"),
               resistivity=None if a.no_res else (eta0, eta1, w), origin=origin, rows_per_wave=a.rows)
    s.upload(U)
    W0 = s.get_state()
    del U
    s.step(warmup)
    s.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.step(steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ms, n = s.timing()
    cells = nx * ny * nz
    res = {"workload": f"3D {'ideal' if a.no_res else 'resistive'} Harris sheet {nx}x{ny}x{nz}", "steps": steps,
           "ms_per_step": 1e3 * el / steps, "zone_updates_per_s": cells * steps / el,
           "hbm_frac_step": cells * steps / el * BYTES_PER_ZONE_UPDATE / 1e9 / HBM_PEAK_GBS,
           "stage_ms": [ms[0] / max(1, n[0]), ms[1] / max(1, n[1])]}
    log(f"forward: {res['ms_per_step']:.3f} ms/step")
    if not a.no_grad and not a.no_res:
        nsteps = grad_steps
        from openmhd_amd.diagnostics import reconnected_flux

        j0 = ny // 2  # rows j0-1 and j0 straddle the sheet centre y = 0

        s.set_state(W0)
        adj = Adjoint(s, nsteps, every)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dts, chs = adj.forward()
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t0
        WT = s.get_state()
        phi, g_by = reconnected_flux(WT, dx)
        lamT = np.zeros_like(WT)
        lamT[6, :, j0 - 1, :] = g_by
        lamT[6, :, j0, :] = g_by
        del WT
        # seed and result resident in HBM, as the forward's state is: the
        # timed backward moves nothing over PCIe
        lamT_d = torch.from_numpy(lamT).cuda()
        lam0_d = torch.empty_like(lamT_d)
        del lamT
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        adj.backward(lamT_d, lam0_d)
        torch.cuda.synchronize()
        t_bwd = time.perf_counter() - t0
        del lamT_d, lam0_d
        g = adj.eta_grad()
        adj.close()
        log(f"adjoint: forward {t_fwd:.3f} s, backward {t_bwd:.3f} s, dphi/d(eta0, eta1, w) = {g}")
        # central differences of the flux with the sign pattern of the base run
        # (the same derivative at the base point; no |By| kink is crossed)
        def flux_lin(W):
            return float((g_by * (W[6, :, j0 - 1, :] + W[6, :, j0, :])).sum())

        fd = {}
        eps_list = (1e-2, 1e-3, 1e-4, 1e-5) if fd_on else ()
        for k, name in ((1, "eta1"), (2, "w")):
            rel = []
            for eps in eps_list:
                L = []
                for sgn in (1, -1):
                    prm = [eta0, eta1, w]
                    prm[k] *= 1 + sgn * eps
                    s.set_resistivity(*prm)
                    s.set_state(W0)
                    s.step_fixed(dts, chs)
                    L.append(flux_lin(s.get_state()))
                d = (L[0] - L[1]) / (2 * eps * [eta0, eta1, w][k])
                rel.append(abs(d - g[k]) / abs(g[k]))
                log(f"d phi / d {name}: adjoint {g[k]:.10e}, central difference (eps {eps:g}) {d:.10e}, rel {rel[-1]:.1e}")
            fd[name] = rel
        s.set_resistivity(eta0, eta1, w)
        res.update({"grad_steps": nsteps, "checkpoint_every": every, "reconnected_flux": phi,
                    "dphi_deta1": g[1], "dphi_dw": g[2], "fd_eps": list(eps_list), "fd_rel_err": fd,
                    "adjoint_forward_s": t_fwd, "adjoint_backward_s": t_bwd, "backward_over_forward": t_bwd / t_fwd})
    s.close()
    return res


def bench_adjoint(a, fd=True, nsteps=None):
    """Orszag-Tang 512^2 + 16 seeded Fourier velocity modes
    (|k| <= 8, amplitude 1e-2), 200 MC/HLLD/GLM/TVD-RK2 steps, checkpoints
    every 20; loss = ||By(T) - By_target(T)||^2 with the target run from a
    second seeded perturbation; gradient w.r.t. the 16 mode amplitudes via the
    adjoint (torch.autograd through Rollout), checked by central differences
    along 16 seeded directions with the frozen dt/ch sequence (unless fd is
    False).  Returns the figures as a dict."""
    import numpy as np
    import torch

    from openmhd_amd import Adjoint, Rollout, Solver, problems

    torch.cuda.set_device(0)
    n, every = a.n if a.n != 4096 else 512, 20
    nsteps = a.steps if nsteps is None else nsteps
    dx = 2 * np.pi / n
    s = Solver(n, n, dx, cfl=0.4, lm_type=2)
    modes = problems.fourier_modes(1)
    tmodes = problems.fourier_modes(2)
    tgt_coef = np.full(16, 1e-2) * np.random.default_rng(2).uniform(0.5, 1.5, 16)
    s.set_state(problems.orszag_tang_perturbed(n, tmodes, tgt_coef))
    s.step(nsteps)
    target = torch.tensor(s.get_state()[6], device="cuda")

    def loss_of(coef, roll, dt=None, ch=None):
        W0 = problems.orszag_tang_perturbed(n, modes, coef)
        WT = roll(W0, dt, ch)
        return ((WT[6] - target) ** 2).sum()

    roll = Rollout(s, nsteps, every)
    coef = torch.full((16,), 1e-2, dtype=torch.float64, device="cuda", requires_grad=True)
    L = loss_of(coef, roll)
    L.backward()
    grad = coef.grad.detach().cpu().numpy()
    dts, chs = roll.dt.copy(), roll.ch.copy()
    # central differences of the forward with the frozen sequence
    # (the map is piecewise smooth: a larger step crosses limiter kinks, a
    # smaller one sees the forward's rounding noise; the sweep shows both)
    rng = np.random.default_rng(16)
    eps_list = (1e-7, 1e-8, 1e-9, 1e-10)
    rel = []
    with torch.no_grad():
        for k in range(16 if fd else 0):
            d = rng.standard_normal(16)
            d /= np.linalg.norm(d)
            dv = torch.tensor(d, device="cuda")
            an = float(grad @ d)
            r = []
            for eps in eps_list:
                lp = float(loss_of(coef + eps * dv, roll, dts, chs))
                lm = float(loss_of(coef - eps * dv, roll, dts, chs))
                r.append(abs((lp - lm) / (2 * eps) - an) / abs(an))
            rel.append(r)
            log(f"direction {k}: adjoint {an:.12e} rel. diff. at eps {eps_list}: " + " ".join(f"{x:.1e}" for x in r))
    rel = np.array(rel).reshape(-1, len(eps_list))
    best = rel.min(axis=1) if len(rel) else np.full(1, np.nan)
    # timing: plain steps, the adjoint's forward pass (with checkpoints), the backward pass
    W0 = problems.orszag_tang_perturbed(n, modes, np.full(16, 1e-2))
    W0d = torch.tensor(W0, device="cuda")
    lamT = torch.randn_like(W0d)
    adj = Adjoint(s, nsteps, every)
    t = {}
    for rep in range(2):  # the first round warms up
        s.set_state(W0d)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.step_fixed(dts, chs)
        torch.cuda.synchronize()
        t["plain_fixed"] = time.perf_counter() - t0
        s.set_state(W0d)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.step(nsteps)
        torch.cuda.synchronize()
        t["plain_cfl"] = time.perf_counter() - t0
        s.set_state(W0d)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        adj.forward(dts, chs)
        torch.cuda.synchronize()
        t["forward"] = time.perf_counter() - t0
        out = torch.empty_like(W0d)
        t0 = time.perf_counter()
        adj.backward(lamT, out)
        torch.cuda.synchronize()
        t["backward"] = time.perf_counter() - t0
    adj.close()
    res = {"workload": f"adjoint: Orszag-Tang {n}^2 + 16 modes, {nsteps} steps, checkpoints every {every}",
           "loss": float(L), "grad_norm": float(np.linalg.norm(grad)), "seconds": t,
           "backward_over_forward": t["backward"] / t["forward"],
           "backward_over_plain_step": t["backward"] / t["plain_cfl"],
           "zone_updates_per_s_forward": n * n * nsteps / t["forward"]}
    if fd:
        res.update({"fd_eps": list(eps_list), "fd_rel_err_median_per_eps": [float(x) for x in np.median(rel, axis=0)],
                    "fd_rel_err_best_max": float(best.max()), "fd_rel_err_best_median": float(np.median(best))})
    return res


def secondary(a):
    """The other configurations, bounded so that the default run stays within
    a few minutes: the 256-member ensemble (300 steps, both dt modes), the
    Orszag-Tang 512^2 adjoint (200 steps, checkpoints every 20; the
    finite-difference check is in --workload adjoint and the GPU tests), and
    the 3D resistive Harris sheet at 512x256x256 (forward over 10 steps after
    2, adjoint over 10 steps with checkpoints every 10)."""
    import torch

    out = {}
    for name, fn in (("ensemble", lambda: bench_ensemble(a, steps=300)),
                     ("adjoint_2d", lambda: bench_adjoint(a, fd=False, nsteps=200)),
                     ("resistive_3d", lambda: bench_3d(a, steps=10, warmup=2, grad_steps=10, every=10, fd=False))):
        t0 = time.perf_counter()
        try:
            out[name] = fn()
        except Exception as e:  # a secondary figure never hides the headline
            log(f"secondary {name} failed:", repr(e))
            out[name] = {"error": repr(e)}
        out[name]["wall_s"] = time.perf_counter() - t0
        torch.cuda.empty_cache()
        log(f"secondary {name}: {out[name]['wall_s']:.1f} s")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: the configuration's own count: KH 500, ensemble 300, adjoint 200, "
                         "3D 100 including the warmup)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--rows", type=int, default=0, help="rows per wavefront chunk (0: library default)")
    ap.add_argument("--dt-mode", default="device", choices=("host", "device"),
                    help="kh: dt from the reduced CFL speed on the host each step (host) or on the GPU (device)")
    ap.add_argument("--cpu-sample-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", default="kh", choices=("kh", "ensemble", "3d", "adjoint"),
                    help="kh: KH 4096^2 (the headline line); ensemble: 256 KH members; adjoint: OT 512^2 gradient; "
                         "3d: Harris sheet 512x256x256 forward + resistivity gradient (own JSON lines)")
    ap.add_argument("--members", type=int, default=256)
    ap.add_argument("--nx3", type=int, default=512)
    ap.add_argument("--ny3", type=int, default=256)
    ap.add_argument("--nz3", type=int, default=256)
    ap.add_argument("--no-grad", action="store_true", help="3d: forward timing only")
    ap.add_argument("--no-fd", action="store_true", help="3d: skip the finite-difference check of the gradient")
    ap.add_argument("--no-res", action="store_true", help="3d: ideal MHD (no resistive flux), forward only")
    ap.add_argument("--no-secondary", action="store_true",
                    help="kh: skip the secondary configurations (ensemble, 2D adjoint, 3D resistive)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy (rocprofv3 FETCH/WRITE_SIZE calibration)")
    a = ap.parse_args()
    if a.steps is None:
        a.steps = {"kh": 500, "ensemble": 300, "adjoint": 200, "3d": max(1, 100 - a.warmup)}[a.workload]
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        sys.exit("bench.py runs on one GPU: the grid is one domain on one MI355X and sharding it "
                 "across GPUs is not built -- run with --gpus 1")

    if a.workload != "kh":
        sys.path.insert(0, str(REPO))
        res = {"ensemble": bench_ensemble, "3d": bench_3d, "adjoint": bench_adjoint}[a.workload](a)
        print(json.dumps(res), flush=True)
        return

    # the CPU baseline runs first, as a child process started before this
    # process initialises the GPU
    cpu = None
    if not a.no_cpu and not a.calibrate_copy:
        try:
            cpu = cpu_baseline(a.n, a.cpu_sample_steps)
        except Exception as e:  # the baseline is reported, never the measured value
            log("cpu baseline failed:", repr(e))

    import torch

    sys.path.insert(0, str(REPO))
    from openmhd_amd import Solver2D
    from openmhd_amd.problems import kelvin_helmholtz

    torch.cuda.set_device(0)
    if a.calibrate_copy:
        src = torch.empty(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        src.fill_(1.0)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 8, "copy_bytes_written": dst.numel() * 8, "launches": 5}))
        return
    n = a.n
    log(f"building KH {n}x{n} initial state on the host")
    U, dx = kelvin_helmholtz(n, n, seed=1)
    s = Solver2D(n, n, dx, cfl=0.4, lm_type=1, time_type=0, bc=("periodic", "wall"), rows_per_wave=a.rows)
    s.upload(U)
    del U
    if a.dt_mode == "device":
        s.set_mode(device_dt=True)
    log(f"warmup {a.warmup} steps")
    s.step(a.warmup)
    s.set_timing(True)
    torch.cuda.synchronize()
    log(f"timing {a.steps} steps")
    t0 = time.perf_counter()
    s.step(a.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms, launches = s.timing()
    elapsed = t1 - t0
    cells = n * n
    zups = cells * a.steps / elapsed
    ms_step = 1e3 * elapsed / a.steps
    st = [ms[i] / max(1, launches[i]) for i in range(2)]  # average launch duration, ms
    dom = 1 if st[1] >= st[0] else 0
    achieved = BYTES_STAGE[dom + 1] * cells / (st[dom] * 1e-3) / 1e9
    traffic = parse_traffic()
    log(f"{ms_step:.3f} ms/step, {zups / 1e9:.3f} G zone-updates/s; stage1 {st[0]:.3f} ms, stage2 {st[1]:.3f} ms")
    copy_gbs = stream_copy_gbs()
    log(f"device copy: {copy_gbs:.0f} GB/s (read + write)")
    out = {
        "metric": "FP64 zone-updates/s & %HBM peak; batched and adjoint time vs plain GPU step",
        "value": zups,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded KH initial state generated on the host)",
        "config": {"workload": f"2D Kelvin-Helmholtz {n}x{n} FP64, x periodic, y walls, minmod + HLLD + GLM + "
                               f"TVD-RK2, CFL 0.4", "nx": n, "ny": n, "limiter": "minmod", "flux": "HLLD",
                   "time": "TVD-RK2", "dt": a.dt_mode, "parallelism": "single GPU, one domain"},
        "hbm_frac_step": zups * BYTES_PER_ZONE_UPDATE / 1e9 / HBM_PEAK_GBS,
        "roofline": {"bound": "hbm", "kernel": f"k_march stage {dom + 1} (k_march<LIM,{dom + 1},TT,0,0>)", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "avg_launch_ms": st[dom], "algorithmic_bytes_per_launch": BYTES_STAGE[dom + 1] * cells,
                     "stream_copy": copy_gbs, "frac_of_copy": achieved / copy_gbs},
        "cpu_baseline": cpu,
    }
    s.close()
    del s
    if not a.no_secondary:
        out["secondary"] = secondary(a)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
