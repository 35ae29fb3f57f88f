"""Generate the golden fixtures in tests/golden/ from the reference routines.

The reference Fortran (2D_basic, 2D_KH/bc.f90, 3D_basic) is compiled unmodified
by `make -C oracle ref` into oracle/_ref/*.so; this script calls those
subroutines on crafted inputs and stores inputs + outputs as .npz (plain
arrays, no pickles).  Branch labels for the HLLD vectors come from the oracle
and are used only to choose inputs that cover every branch.

    python tests/golden/gen_golden.py [3d | recon | recon_run | ot_divb | snapshots | solvers]   (needs oracle/_ref, built here)
"""
from __future__ import annotations

import hashlib
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
from omhd_testlib import (KH_SIGN, Oracle, Ref, farr, orszag_tang, prim_to_cons, random_prim,  # noqa: E402
                          ref_run_2d, smooth_state)

OUT = HERE


def save(name, **arrs):
    np.savez_compressed(OUT / f"{name}.npz", **{k: np.asarray(v) for k, v in arrs.items()})
    print("wrote", name, {k: np.asarray(v).shape for k, v in arrs.items()})


def hlld_pool(rng, n):
    """Interface state pairs spread over the HLLD branches: sub/supersonic
    flows, weak and strong normal fields, near-HLLC-G switching states."""
    L = random_prim(rng, (n,))
    R = random_prim(rng, (n,))
    kind = rng.integers(0, 6, n)
    # strong normal flow -> FL / FR
    s = kind == 0
    L[s, 0:3] *= 6.0
    R[s, 0:3] = L[s, 0:3] * rng.uniform(0.8, 1.2, (s.sum(), 3))
    # tiny normal field -> HLLC-G
    s = kind == 1
    L[s, 5:8] *= 1e-4
    R[s, 5:8] *= 1e-4
    # strong field, weak pressure -> Alfven close to fast -> HLLC-G
    s = kind == 2
    L[s, 5:8] *= 4.0
    R[s, 5:8] = L[s, 5:8]
    L[s, 3] *= 0.01
    R[s, 3] *= 0.01
    # moderate drift -> L*, R*
    s = kind == 3
    off = rng.normal(0, 2.0, (s.sum(), 1))
    L[s, 0:3] += off
    R[s, 0:3] += off
    return L, R


def main():
    orc = Oracle()
    ref2 = Ref(2)
    ref3 = Ref(3)
    rng = np.random.default_rng(20260511)

    # ---- u2v --------------------------------------------------------------
    ix, jx = 12, 9
    U = np.asfortranarray(prim_to_cons(random_prim(rng, (ix, jx))))
    V = farr(ix, jx, 4)
    ref2.u2v(U, V)
    save("u2v_2d", U=U, V=V)

    # ---- limiter (2D dir 1,2; 3D dir 3) ----------------------------------
    out = {}
    for typ in range(5):
        for d in (1, 2):
            wk = np.asfortranarray(rng.normal(size=(11, 10)))
            wk[3, 3], wk[4, 3], wk[5, 3] = -0.0, 0.0, -0.0  # signed-zero gradients
            wk[6, 5] = wk[7, 5]  # gB == 0
            wk[2, 6] = wk[2, 7] = wk[2, 8]
            wL = np.full((11, 10), 7.0, order="F")
            wR = np.full((11, 10), -7.0, order="F")
            ref2.limiter(wk, wL, wR, d, typ)
            out[f"wk_{typ}_{d}"] = wk
            out[f"wL_{typ}_{d}"] = wL
            out[f"wR_{typ}_{d}"] = wR
        wk = np.asfortranarray(rng.normal(size=(6, 5, 9)))
        wk[2, 2, 3] = -0.0
        wL = np.full((6, 5, 9), 7.0, order="F")
        wR = np.full((6, 5, 9), -7.0, order="F")
        ref3.limiter(wk, wL, wR, 3, typ)
        out[f"wk_{typ}_3"] = wk
        out[f"wL_{typ}_3"] = wL
        out[f"wR_{typ}_3"] = wR
    save("limiter", **out)

    # ---- HLLD: every branch x dir ----------------------------------------
    out = {}
    for d in (1, 2, 3):
        L, R = hlld_pool(rng, 20000)
        br = np.array([orc.hlld(L[i], R[i], 1.3, d - 1)[1] for i in range(len(L))])
        pick = []
        for b in range(1, 9):
            idx = np.flatnonzero(br == b)
            assert len(idx) >= 4, f"branch {b} dir {d} not covered ({len(idx)})"
            pick.extend(idx[:8].tolist())
        pick = np.array(pick)
        n = len(pick)
        if d < 3:  # lay the pairs out along the interface direction of a 2D grid
            shp = (n + 2, 3) if d == 1 else (3, n + 2)
            VL = farr(*shp, 9)
            VR = farr(*shp, 9)
            VL[...] = 1.0
            VR[...] = 1.0
            if d == 1:
                VL[1:n + 1, 1, :] = L[pick]
                VR[1:n + 1, 1, :] = R[pick]
            else:
                VL[1, 1:n + 1, :] = L[pick]
                VR[1, 1:n + 1, :] = R[pick]
            F = farr(*shp, 9)
            ref2.flux_solver(F, VL, VR, 1.3, d)
        else:
            shp = (3, 3, n + 2)
            VL = farr(*shp, 9)
            VR = farr(*shp, 9)
            VL[...] = 1.0
            VR[...] = 1.0
            VL[1, 1, 1:n + 1, :] = L[pick]
            VR[1, 1, 1:n + 1, :] = R[pick]
            F = farr(*shp, 9)
            ref3.flux_solver(F, VL, VR, 1.3, d)
        out[f"VL_{d}"], out[f"VR_{d}"], out[f"F_{d}"] = VL, VR, F
        out[f"branch_{d}"] = br[pick]
    out["ch"] = np.array(1.3)
    save("hlld", **out)

    # ---- glm_ss2, rk (2D/3D), set_dt (2D/3D), bc ---------------------------
    out = {}
    U = np.asfortranarray(prim_to_cons(random_prim(rng, (10, 8))))
    U2 = U.copy(order="F")
    ref2.glm_ss2(U2, 1.37, 0.0123)
    out["glm_in"], out["glm_out"] = U, U2
    F = np.asfortranarray(rng.normal(size=(10, 8, 9)))
    G = np.asfortranarray(rng.normal(size=(10, 8, 9)))
    for mode in range(4):
        A = U.copy(order="F")
        B = np.asfortranarray(rng.normal(size=(10, 8, 9)))
        out[f"rk2d_{mode}_in_wk"], out[f"rk2d_{mode}_in_wk1"] = A.copy(order="F"), B.copy(order="F")
        ref2.rk(mode, A, B, F, G, None, 0.0123, 0.0371)
        out[f"rk2d_{mode}_wk"], out[f"rk2d_{mode}_wk1"] = A, B
    out["rk2d_F"], out["rk2d_G"] = F, G
    U3 = np.asfortranarray(prim_to_cons(random_prim(rng, (6, 5, 7))))
    F3, G3, H3 = (np.asfortranarray(rng.normal(size=(6, 5, 7, 9))) for _ in range(3))
    for mode in range(4):
        A = U3.copy(order="F")
        B = np.asfortranarray(rng.normal(size=(6, 5, 7, 9)))
        out[f"rk3d_{mode}_in_wk"], out[f"rk3d_{mode}_in_wk1"] = A.copy(order="F"), B.copy(order="F")
        ref3.rk(mode, A, B, F3, G3, H3, 0.0123, 0.0371)
        out[f"rk3d_{mode}_wk"], out[f"rk3d_{mode}_wk1"] = A, B
    out["rk3d_F"], out["rk3d_G"], out["rk3d_H"] = F3, G3, H3
    V = farr(10, 8, 4)
    ref2.u2v(U, V)
    vmax, dt = ref2.set_dt(U, V, 0.0371, 0.4)
    out["setdt2d_U"], out["setdt2d_V"], out["setdt2d_out"] = U, V, np.array([vmax, dt])
    V3 = farr(6, 5, 7, 4)
    ref3.u2v(U3, V3)
    vmax, dt = ref3.set_dt(U3, V3, 0.0371, 0.4)
    out["setdt3d_U"], out["setdt3d_V"], out["setdt3d_out"] = U3, V3, np.array([vmax, dt])
    A = np.asfortranarray(rng.normal(size=(7, 6, 9)))
    out["bc_in"] = A.copy(order="F")
    ref2.bc(A)
    out["bc_out"] = A
    A = np.asfortranarray(rng.normal(size=(7, 6, 9)))
    out["bckh_in"] = A.copy(order="F")
    ref2.bc_for_U(A)
    out["bckh_out"] = A
    VL = np.asfortranarray(rng.normal(size=(7, 6, 9)))
    VR = np.asfortranarray(rng.normal(size=(7, 6, 9)))
    out["bcF_inL"], out["bcF_inR"] = VL.copy(order="F"), VR.copy(order="F")
    ref2.bc_for_F(VL, VR)
    out["bcF_L"], out["bcF_R"] = VL.copy(order="F"), VR.copy(order="F")
    ref2.bc_for_G(VL, VR)
    out["bcG_L"], out["bcG_R"] = VL, VR
    A = np.asfortranarray(rng.normal(size=(6, 5, 7, 9)))
    out["bc3d_in"] = A.copy(order="F")
    ref3.bc(A)
    out["bc3d_out"] = A
    save("small", **out)

    # ---- whole runs: OT 200^2 MC 1000 steps, KH walls 128^2 minmod 100 ------
    U, dx = orszag_tang(200)
    ref2.bc(U)
    U0sha = hashlib.sha256(U.tobytes(order="F")).hexdigest()
    dts, chs = ref_run_2d(ref2, U, 1000, dx, 0.4, 2)
    interior = np.ascontiguousarray(U[1:-1, 1:-1, :].transpose(2, 1, 0))
    sys.path.insert(0, str(HERE.parents[1]))
    from openmhd_amd.diagnostics import div_b_l2  # noqa: E402

    save("ot200_mc_1000", dt=dts, ch=chs, dx=np.array(dx), u0_sha256=np.array(U0sha),
         final_sha256=np.array(hashlib.sha256(interior.tobytes()).hexdigest()),
         field_sum=interior.sum(axis=(1, 2)), field_absmax=np.abs(interior).max(axis=(1, 2)),
         sample=interior[:, ::7, ::7], divb_l2=np.array(div_b_l2(U, dx)))

    sys.path.insert(0, str(HERE.parents[1]))
    from openmhd_amd.problems import kelvin_helmholtz  # noqa: E402

    U, dx = kelvin_helmholtz(128, 128, seed=7)
    ref2.bc_for_U(U)
    U0 = U.copy(order="F")
    dts, chs = ref_run_2d(ref2, U, 100, dx, 0.4, 1, bc="kh")
    save("kh128_minmod_100", U0=U0, dt=dts, ch=chs, dx=np.array(dx), final=U)

    # ---- small 2D periodic run with every limiter (10 steps) ---------------
    out = {}
    for lm in range(5):
        U = smooth_state(np.random.default_rng(lm), 20, 18)
        ref2.bc(U)
        out[f"U0_{lm}"] = U.copy(order="F")
        dts, chs = ref_run_2d(ref2, U, 10, 0.05, 0.4, lm)
        out[f"U_{lm}"], out[f"dt_{lm}"] = U, dts
    save("limiters_run", **out)

    # ---- small 3D periodic run (3D_basic/main.f90 order) --------------------
    from omhd_testlib import ref_run_3d  # noqa: E402

    U = smooth_state(np.random.default_rng(33), 14, 12, 10)
    ref3.bc(U)
    U0 = U.copy(order="F")
    dts, chs = ref_run_3d(ref3, U, 4, 0.1, 0.4, 2)
    save("run3d", U0=U0, U=U, dt=dts, ch=chs)


def only_3d():
    ref3 = Ref(3)
    from omhd_testlib import ref_run_3d  # noqa: E402

    U = smooth_state(np.random.default_rng(33), 14, 12, 10)
    ref3.bc(U)
    U0 = U.copy(order="F")
    dts, chs = ref_run_3d(ref3, U, 4, 0.1, 0.4, 2)
    save("run3d", U0=U0, U=U, dt=dts, ch=chs)


def recon():
    """2D_reconnection: per-face boundary tables (bc_for_U/F/G) and the 2D CPU
    form of the resistive flux (its own eta(x, y) profile)."""
    from omhd_testlib import RefRec  # noqa: E402

    rr = RefRec()
    rng = np.random.default_rng(44)
    U = farr(9, 7, 9)
    U[...] = rng.normal(size=U.shape)
    Uo = U.copy(order="F")
    rr.bc_for_U(Uo)
    VL, VR = farr(9, 7, 9), farr(9, 7, 9)
    VL[...] = rng.normal(size=VL.shape)
    VR[...] = rng.normal(size=VR.shape)
    FL, FR = VL.copy(order="F"), VR.copy(order="F")
    rr.bc_for_F(FL, FR)
    GL, GR = VL.copy(order="F"), VR.copy(order="F")
    rr.bc_for_G(GL, GR)
    # resistive flux on a smooth field near the resistivity peak (x = y = 0)
    ix, jx, dx = 12, 10, 0.3
    x = -0.5 * (ix - 2) * dx - 0.5 * dx + dx * np.arange(ix)
    y = -0.5 * (jx - 2) * dx - 0.5 * dx + dx * np.arange(jx)
    X, Y = np.meshgrid(x, y, indexing="ij")
    Ur = farr(ix, jx, 9)
    Ur[...] = 0.1 * rng.normal(size=Ur.shape)
    Ur[..., 5] = np.tanh(Y) + 0.2 * np.sin(X)
    Ur[..., 6] = 0.3 * np.sin(0.7 * X + 0.4 * Y)
    Ur[..., 7] = 0.5 * np.cos(0.5 * X - 0.9 * Y)
    out = {}
    for d in (1, 2):
        F = farr(ix, jx, 9)
        F[...] = rng.normal(size=F.shape)
        out[f"F0_{d}"] = F.copy(order="F")
        rr.flux_resistive(F, Ur, x, y, dx, d)
        out[f"F_{d}"] = F
    save("recon", bcU_in=U, bcU_out=Uo, V_L=VL, V_R=VR, bcF_L=FL, bcF_R=FR, bcG_L=GL, bcG_R=GR, res_U=Ur, res_x=x,
         res_y=y, res_dx=dx, **out)


def solvers():
    """Flux types 0 (LLF), 1 (HLL), 2 (HLLC), 4 (LHLLD) of the reference's
    flux_solver on the HLLD fixture's interface pairs, dir 1/2 (2D form) and 3
    (3D form); whole 2D runs (periodic and KH walls) and 3D periodic runs."""
    d0 = np.load(OUT / "hlld.npz", allow_pickle=False)
    ref2, ref3 = Ref(2), Ref(3)
    out = {}
    for typ in (0, 1, 2, 4):
        for d in (1, 2, 3):
            VL = np.asfortranarray(d0[f"VL_{d}"])
            VR = np.asfortranarray(d0[f"VR_{d}"])
            F = farr(*VL.shape)
            (ref2 if d < 3 else ref3).flux_solver(F, VL, VR, float(d0["ch"]), d, typ)
            out[f"F_{typ}_{d}"] = F
    # whole runs of each type: periodic (MC) and KH walls (minmod), 8 steps
    rng = np.random.default_rng(66)
    for typ in (0, 1, 2, 4):
        for bc, lm in (("periodic", 2), ("kh", 1)):
            U = smooth_state(rng, 22, 20)
            if bc == "kh":
                ref2.bc_for_U(U)
            else:
                ref2.bc(U)
            U0 = U.copy(order="F")
            dts, chs = ref_run_2d(ref2, U, 8, 0.05, 0.4, lm, bc=bc, flux_type=typ)
            out[f"run_{typ}_{bc}_U0"], out[f"run_{typ}_{bc}_U"] = U0, U
            out[f"run_{typ}_{bc}_dt"], out[f"run_{typ}_{bc}_ch"] = dts, chs
    # 3D periodic runs of each type (3D_basic routines), MC, 4 steps
    from omhd_testlib import ref_run_3d  # noqa: E402

    for typ in (0, 1, 2, 4):
        U = smooth_state(rng, 14, 12, 10)
        ref3.bc(U)
        U0 = U.copy(order="F")
        dts, chs = ref_run_3d(ref3, U, 4, 0.1, 0.4, 2, flux_type=typ)
        out[f"run3d_{typ}_U0"], out[f"run3d_{typ}_U"] = U0, U
        out[f"run3d_{typ}_dt"], out[f"run3d_{typ}_ch"] = dts, chs
    save("solvers", **out)


def snapshots():
    """Snapshot files written by the reference's own fileio_output (2D and 3D
    forms), committed byte for byte with the arrays they hold."""
    import ctypes as C

    from omhd_testlib import REF_DIR, ptr  # noqa: E402

    rng = np.random.default_rng(55)
    for dim, shp in ((2, (8, 6)), (3, (6, 5, 4))):
        L = C.CDLL(str(REF_DIR / f"libreffio{dim}d.so"))
        U = farr(*shp, 9)
        V = farr(*shp, 4)
        U[...] = rng.normal(size=U.shape)
        V[...] = rng.normal(size=V.shape)
        axes = [np.ascontiguousarray(rng.normal(size=n)) for n in shp]
        t = C.c_double(1.2345678901234567)
        name = OUT / f"snapshot_ref{dim}d.dat"
        fn = str(name).encode().ljust(256, b" ")  # character*256, blank-padded
        ext = [C.byref(C.c_int(n)) for n in shp]
        args = [C.c_char_p(fn)] + ext + [C.byref(t)] + [ptr(a) for a in axes] + [ptr(U), ptr(V), C.c_long(256)]
        L.fileio_output_(*args)
        save(f"snapshot_ref{dim}d", U=U, V=V, t=t.value, **{f"axis{i}": a for i, a in enumerate(axes)})
        print("wrote", name, name.stat().st_size, "bytes")


def ot_divb():
    """Adds the L2 norm of div B of the reference's OT 200^2 MC state after
    1000 steps (openmhd_amd.diagnostics.div_b_l2) to ot200_mc_1000.npz,
    after checking that the rerun reproduces the stored final-state hash."""
    sys.path.insert(0, str(HERE.parents[1]))
    from openmhd_amd.diagnostics import div_b_l2

    d = dict(np.load(OUT / "ot200_mc_1000.npz", allow_pickle=False))
    ref2 = Ref(2)
    U, dx = orszag_tang(200)
    ref2.bc(U)
    ref_run_2d(ref2, U, 1000, dx, 0.4, 2)
    interior = np.ascontiguousarray(U[1:-1, 1:-1, :].transpose(2, 1, 0))
    assert hashlib.sha256(interior.tobytes()).hexdigest() == str(d["final_sha256"])
    d["divb_l2"] = np.array(div_b_l2(U, dx))
    save("ot200_mc_1000", **d)


def recon_run():
    """2D_reconnection end to end: eta(x, y) and set_dt_resistive of
    flux_resistive.f90 at sample points, and two short runs of the project's
    driver (openmhd_amd/fortran/main2d_recon.f90) linked by flang against the
    reference's 2D_basic + 2D_reconnection routines (oracle/_ref/
    main2d_recon_ref): one CFL-limited, one limited by the resistive step."""
    import ctypes as C
    import subprocess
    import tempfile

    from omhd_testlib import REF_DIR, RefRec, recon_problem

    from openmhd_amd import problems, snapshot

    rr = RefRec()
    rr.L.eta_.restype = C.c_double
    rng = np.random.default_rng(45)
    pts = np.concatenate([rng.uniform(-30, 30, (300, 2)), rng.uniform(-1, 1, (200, 2)),
                          np.array([[0.0, 0.0], [-0.0, 0.0], [25.0, 0.0], [0.0, -25.0], [17.7, 17.7], [1e-8, 0.0]])])
    eta = np.array([rr.L.eta_(C.byref(C.c_double(a)), C.byref(C.c_double(b))) for a, b in pts])
    dtc = np.array([[1.0, 0.1, 0.4], [1e-4, 0.1, 0.4], [0.3, 0.01, 0.35], [2e-3, 0.0125, 0.8]])
    dto = []
    for dt, dx, cfl in dtc:
        v = C.c_double(dt)
        rr.L.set_dt_resistive_(C.byref(v), C.byref(C.c_double(dx)), C.byref(C.c_double(cfl)))
        dto.append(v.value)
    out = dict(eta_pts=pts, eta=eta, sdr_in=dtc, sdr_out=np.array(dto))
    exe = REF_DIR / "main2d_recon_ref"
    for name, lx, lm, tt, flux in (("a", 6.4, 1, 0, 3), ("b", 0.32, 2, 1, 3)):
        nx, ny, steps = 32, 12, 8
        U, _, x, y = problems.harris_sheet_2d(nx, ny, lx=lx)
        p, tab = recon_problem(x, y, lm_type=lm, flux_type=flux, time_type=tt)
        Oracle().bc_u(U, p)
        with tempfile.TemporaryDirectory() as d:
            fin, fo, tr = f"{d}/in.dat", f"{d}/out.dat", f"{d}/tr.bin"
            snapshot.write(fin, 0.0, x, y, U)
            subprocess.run([str(exe), str(nx), str(ny), str(steps), "--input", fin, "--output", fo, "--trace", tr,
                            "--lm", str(lm), "--tt", str(tt), "--flux", str(flux)], check=True)
            _, _, _, Uf, _ = snapshot.read(fo)
            t = np.fromfile(tr)
        out.update({f"{name}_U0": U, f"{name}_x": x, f"{name}_y": y, f"{name}_dt": t[:steps], f"{name}_ch": t[steps:],
                    f"{name}_U": Uf, f"{name}_case": np.array([lm, tt, flux, steps])})
    save("recon_run", **out)


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "3d":
        only_3d()
    elif len(sys.argv) > 1 and sys.argv[1] == "recon":
        recon()
    elif len(sys.argv) > 1 and sys.argv[1] == "recon_run":
        recon_run()
    elif len(sys.argv) > 1 and sys.argv[1] == "ot_divb":
        ot_divb()
    elif len(sys.argv) > 1 and sys.argv[1] == "snapshots":
        snapshots()
    elif len(sys.argv) > 1 and sys.argv[1] == "solvers":
        solvers()
    else:
        main()
