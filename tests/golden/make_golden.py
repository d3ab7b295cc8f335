"""Generate the golden fixtures of tests/golden/ from the reference itself.

Runs the reference Fortran (GDF_Balsara_P2_0316, built by
oracle/build_ref.sh into oracle/_ref/) on seeded inputs and stores inputs and
reference outputs as .npz.  Only this script runs the reference; the tests
read the fixtures.  State arrays are stored as their 52 active (mode,
variable) planes (the other 28 slots of uh are always zero,
GDF_Balsara_P2_0316/Lh.f90:59) to keep the files small.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))

# active (mode d, variable n), 0-based, in device plane order
ACTIVE = [(d, n) for n in range(8) for d in range(6)] + [(6, 5), (8, 5), (7, 6), (9, 6)]


def pack(uh):
    return np.stack([uh[:, :, d, n] for (d, n) in ACTIVE], axis=-1)


def unpack(a):
    uh = np.zeros(a.shape[:2] + (10, 8), order="F")
    for k, (d, n) in enumerate(ACTIVE):
        uh[:, :, d, n] = a[:, :, k]
    return uh


def zeros(Nx, Ny):
    return (np.zeros((Nx + 2, Ny + 2, 10, 8), order="F"), np.zeros((Nx + 1, Ny, 3), order="F"),
            np.zeros((Nx, Ny + 1, 3), order="F"))


def periodic_halo(uh):
    uh = uh.copy(order="F")
    uh[-1, :] = uh[1, :]
    uh[0, :] = uh[-2, :]
    uh[:, -1] = uh[:, 1]
    uh[:, 0] = uh[:, -2]
    return uh


def ot_init(variant, Nx, Ny=None):
    Ny = Ny or Nx
    r = O.run_ref(variant, "init3", *zeros(Nx, Ny))
    return r["uh"], r["Bx"], r["By"]


def stress_state(uh, Bx, By, seed, amp=0.08):
    """OT state with w, Bz and every higher mode perturbed (seeded)."""
    rng = np.random.default_rng(seed)
    uh = uh.copy(order="F")
    Nx, Ny = uh.shape[0] - 2, uh.shape[1] - 2
    I = (slice(1, Nx + 1), slice(1, Ny + 1))
    uh[I + (0, 3)] += 0.3 * rng.standard_normal((Nx, Ny))        # rho*w
    uh[I + (0, 7)] += 0.2 * rng.standard_normal((Nx, Ny))        # Bz
    scale = np.array([np.abs(uh[I + (0, n)]).mean() + 0.2 for n in range(8)])
    for (d, n) in ACTIVE:
        if d == 0:
            continue
        uh[I + (d, n)] += amp * scale[n] * rng.standard_normal((Nx, Ny))
    # keep pressure comfortably positive at the averages
    uh[I + (0, 4)] += 1.0
    Bx = Bx + 0.05 * rng.standard_normal(Bx.shape)
    By = By + 0.05 * rng.standard_normal(By.shape)
    return periodic_halo(uh), np.asfortranarray(Bx), np.asfortranarray(By)


def pp_stress(uh, seed):
    """cells whose density or pressure goes negative at GLL points"""
    rng = np.random.default_rng(seed)
    uh = uh.copy(order="F")
    Nx, Ny = uh.shape[0] - 2, uh.shape[1] - 2
    for i in range(1, Nx + 1):
        for j in range(1, Ny + 1):
            r = rng.random()
            if r < 0.15:     # density dips below zero at xi = -1
                uh[i, j, 1, 0] = 1.3 * uh[i, j, 0, 0]
            elif r < 0.30:   # density overshoot above 2*rho_bar
                uh[i, j, 2, 0] = -1.2 * uh[i, j, 0, 0]
            elif r < 0.55:   # pressure dips below zero
                uh[i, j, 1, 4] = -0.9 * uh[i, j, 0, 4]
                uh[i, j, 3, 4] = 0.3 * uh[i, j, 0, 4]
    return periodic_halo(uh)


def save(name, **kw):
    path = os.path.join(OUT, name + ".npz")
    np.savez_compressed(path, **kw)
    print(f"wrote {path} ({os.path.getsize(path) / 1024:.0f} KiB)")


def main():
    two_pi = 2 * np.pi
    dom_ot = np.array([0.0, two_pi, 0.0, two_pi])

    # --- init_data (init_data.f90 with init3/init1/init4) ------------------
    for mode, tag in (("init3", "ot"), ("init1", "vortex"), ("init4", "rotor")):
        r = O.run_ref("r16_f4", mode, *zeros(16, 16))
        save(f"init_{tag}_16", uh=pack(r["uh"]), Bx=r["Bx"], By=r["By"],
             domain=np.array([r["xa"], r["xb"], r["ya"], r["yb"]]), M=r["M"], tend=r["tend"])

    uh0, Bx0, By0 = ot_init("r16_f4", 16)

    # --- Lh + LhEz for each flux type on a stress state ---------------------
    su, sx, sy = stress_state(uh0, Bx0, By0, seed=11)
    for ft in (1, 2, 3, 4):
        r = O.run_ref(f"r16_f{ft}", "lh", su, sx, sy, domain=dom_ot)
        save(f"lh_f{ft}_16", uh=pack(su), Bx=sx, By=sy, domain=dom_ot, bc=1,
             du=pack(r["du"]), dEz1=r["dEz1"], dEz2=r["dEz2"], EzVertex=r["EzVertex"])

    # non-square grid with outflow boundaries
    u8, b8x, b8y = ot_init("r12x8_f4", 12, 8)
    s8, s8x, s8y = stress_state(u8, b8x, b8y, seed=12)
    r = O.run_ref("r12x8_f4", "lh", s8, s8x, s8y, domain=dom_ot, bc=2)
    save("lh_f4_12x8_outflow", uh=pack(s8), Bx=s8x, By=s8y, domain=dom_ot, bc=2,
         du=pack(r["du"]), dEz1=r["dEz1"], dEz2=r["dEz2"], EzVertex=r["EzVertex"], uh_bc=pack(r["uh"]))

    # --- TVB with stale halos (halo values from another state) -------------
    st, stx, sty = stress_state(uh0, Bx0, By0, seed=13, amp=0.25)
    stale = periodic_halo(stress_state(uh0, Bx0, By0, seed=14, amp=0.25)[0])
    st = st.copy(order="F")
    for sl in ((0, slice(None)), (-1, slice(None)), (slice(None), 0), (slice(None), -1)):
        st[sl] = stale[sl]
    for M in (0.0, 10.0):
        r = O.run_ref("r16_f4", "tvb", st, stx, sty, domain=dom_ot, M=M)
        changed = int(np.sum(np.any(np.abs(r["uh"][1:-1, 1:-1] - st[1:-1, 1:-1]) > 0, axis=(2, 3))))
        print(f"  tvb M={M}: {changed} of 256 cells changed")
        save(f"tvb_M{int(M)}_16", uh=pack(st), Bx=stx, By=sty, domain=dom_ot, M=M, out=pack(r["uh"]),
             cells_changed=changed)

    # --- positivity limiter -------------------------------------------------
    sp = pp_stress(stress_state(uh0, Bx0, By0, seed=15)[0], seed=16)
    r = O.run_ref("r16_f4", "pp", sp, sx, sy, domain=dom_ot)
    changed = int(np.sum(np.any(np.abs(r["uh"][1:-1, 1:-1] - sp[1:-1, 1:-1]) > 0, axis=(2, 3))))
    print(f"  pp: {changed} of 256 cells changed")
    save("pp_16", uh=pack(sp), Bx=sx, By=sy, domain=dom_ot, out=pack(r["uh"]), cells_changed=changed)

    # --- div_free_Balsara, dt and diagnostics --------------------------------
    r = O.run_ref("r16_f4", "divfree", su, sx, sy, domain=dom_ot)
    save("divfree_16", uh=pack(su), Bx=sx, By=sy, domain=dom_ot, out=pack(r["uh"]))
    rd = O.run_ref("r16_f4", "dt", su, sx, sy, domain=dom_ot)
    rg = O.run_ref("r16_f4", "diag", su, sx, sy, domain=dom_ot)
    save("dt_diag_16", uh=pack(su), Bx=sx, By=sy, domain=dom_ot,
         dt=rd["dt"], alphax=rd["alphax"], alphay=rd["alphay"],
         umax=rg["umax"], totaldiv=rg["totaldiv"], pmin=rg["pmin"])

    # --- KXRCF_Detector (P6: uG declared, flags reset, Ck = 1) -------------
    r = O.run_ref("r16_f4", "kxrcf", su, sx, sy, domain=dom_ot)
    print(f"  kxrcf stress 16: {int(r['trouble'].sum())} of 256 flagged")
    save("kxrcf_16", uh=pack(su), Bx=sx, By=sy, domain=dom_ot, Ck=1.0, trouble=r["trouble"])
    r0 = O.run_ref("r32_f4_pp_kx", "init4", *zeros(32, 32))
    dom_r = np.array([r0["xa"], r0["xb"], r0["ya"], r0["yb"]])
    r = O.run_ref("r32_f4_pp_kx", "kxrcf", periodic_halo(r0["uh"]), r0["Bx"], r0["By"], domain=dom_r)
    print(f"  kxrcf rotor 32: {int(r['trouble'].sum())} of 1024 flagged")
    save("kxrcf_rotor_32", uh=pack(periodic_halo(r0["uh"])), Bx=r0["Bx"], By=r0["By"], domain=dom_r, Ck=1.0,
         trouble=r["trouble"])

    # --- whole RK3 runs (prologue + steps to tend) ---------------------------
    runs = [
        ("run_f4_16", "r16_f4", (16, 16), 1, 50.0, 0.05),
        ("run_f2_16", "r16_f2", (16, 16), 1, 10.0, 0.05),
        ("run_f1_16", "r16_f1", (16, 16), 1, 10.0, 0.05),
        ("run_f3_16", "r16_f3", (16, 16), 1, 10.0, 0.05),
        ("run_f4_pp_16_outflow", "r16_f4_pp", (16, 16), 2, 0.0, 0.05),
        ("run_f4_pp_32", "r32_f4_pp", (32, 32), 1, 50.0, 0.05),
        ("run_f2_32", "r32_f2", (32, 32), 1, 10.0, 0.08),
        # TVB gated by KXRCF (TVB_Limiter_P2.f90:12,22,116 enabled, Ck = 1)
        ("run_f2_kx_32", "r32_f2_kx", (32, 32), 1, 50.0, 0.3),
    ]
    for name, variant, (Nx, Ny), bc, M, tend in runs:
        ui, bxi, byi = ot_init(variant, Nx, Ny)
        r = O.run_ref(variant, "run", ui, bxi, byi, domain=dom_ot, bc=bc, M=M, tend=tend)
        print(f"  {name}: {int(r['count'])} steps, t={r['t']}")
        save(name, uh0=pack(ui), Bx0=bxi, By0=byi, domain=dom_ot, bc=bc, M=M, tend=tend,
             uh=pack(r["uh"]), Bx=r["Bx"], By=r["By"], t=r["t"], count=r["count"],
             umax=r["umax"], totaldiv=r["totaldiv"], pmin=r["pmin"])

    # rotor (init4.txt on [0,1]^2) with the positivity limiter
    r0 = O.run_ref("r32_f4_pp", "init4", *zeros(32, 32))
    dom_r = np.array([r0["xa"], r0["xb"], r0["ya"], r0["yb"]])
    r = O.run_ref("r32_f4_pp", "run", r0["uh"], r0["Bx"], r0["By"], domain=dom_r, M=10.0, tend=0.01)
    print(f"  rotor: {int(r['count'])} steps")
    save("run_rotor_f4_pp_32", uh0=pack(r0["uh"]), Bx0=r0["Bx"], By0=r0["By"], domain=dom_r, bc=1, M=10.0,
         tend=0.01, uh=pack(r["uh"]), Bx=r["Bx"], By=r["By"], t=r["t"], count=r["count"],
         umax=r["umax"], totaldiv=r["totaldiv"], pmin=r["pmin"])
    r = O.run_ref("r32_f4_pp_kx", "run", r0["uh"], r0["Bx"], r0["By"], domain=dom_r, M=10.0, tend=0.01)
    print(f"  rotor, KXRCF-gated TVB: {int(r['count'])} steps, {int(r['trouble'].sum())} flagged at the end")
    save("run_rotor_f4_pp_kx_32", uh0=pack(r0["uh"]), Bx0=r0["Bx"], By0=r0["By"], domain=dom_r, bc=1, M=10.0,
         tend=0.01, Ck=1.0, uh=pack(r["uh"]), Bx=r["Bx"], By=r["By"], t=r["t"], count=r["count"],
         umax=r["umax"], totaldiv=r["totaldiv"], pmin=r["pmin"], trouble=r["trouble"])


# initN.txt families beyond init1/3/4 (tag, driver mode, ic id); the run
# fixtures use the defaults of the init file with the positivity limiter
IC_LIBRARY = [("blast", "init5", 5), ("cloud", "init6", 6), ("cloud2", "init6b", 61), ("loop", "init7", 7),
              ("briowu", "init8", 8), ("briowu_y", "init8b", 81), ("ryujones", "init9", 9),
              ("ryujones_y", "init9b", 91)]


def main_ic():
    for tag, mode, ic in IC_LIBRARY:
        r = O.run_ref("r16_f4", mode, *zeros(16, 16))
        bc = 1 if ic == 7 else 2
        save(f"init_{tag}_16", uh=pack(r["uh"]), Bx=r["Bx"], By=r["By"], ic=ic, bc=bc,
             domain=np.array([r["xa"], r["xb"], r["ya"], r["yb"]]), M=r["M"], tend=r["tend"])
    # calculate_Az.f90 on the magnetic loop
    r0 = O.run_ref("r16_f4", "init7", *zeros(16, 16))
    dom = np.array([r0["xa"], r0["xb"], r0["ya"], r0["yb"]])
    r = O.run_ref("r16_f4", "az", r0["uh"], r0["Bx"], r0["By"], domain=dom)
    save("az_loop_16", uh=pack(r0["uh"]), Bx0=r0["Bx"], By0=r0["By"], domain=dom, Bx=r["Bx"], By=r["By"])
    # calculate_L2error.f90 on the projected vortex, translated by t20
    r0 = O.run_ref("r16_f4", "init1", *zeros(16, 16))
    dom = np.array([r0["xa"], r0["xb"], r0["ya"], r0["yb"]])
    t20s = np.array([0.0, 0.3, 1.7])
    L2 = np.stack([O.run_ref("r16_f4", "l2err", r0["uh"], r0["Bx"], r0["By"], domain=dom, t20=t)["L2"]
                   for t in t20s])
    save("l2err_vortex_16", uh=pack(r0["uh"]), Bx=r0["Bx"], By=r0["By"], domain=dom, t20=t20s, L2=L2)
    # whole RK3 runs of shock problems (P0 data, outflow boundaries, positivity)
    for tag, mode, tend in (("blast", "init5", 0.002), ("cloud", "init6", 0.02), ("briowu_y", "init8b", 0.02)):
        r0 = O.run_ref("r16_f4_pp", mode, *zeros(16, 16))
        dom = np.array([r0["xa"], r0["xb"], r0["ya"], r0["yb"]])
        r = O.run_ref("r16_f4_pp", "run", r0["uh"], r0["Bx"], r0["By"], domain=dom, bc=2, M=r0["M"], tend=tend)
        print(f"  run {tag}: {int(r['count'])} steps, t={r['t']}, pmin={r['pmin']:.3e}")
        save(f"run_{tag}_f4_pp_16", uh0=pack(r0["uh"]), Bx0=r0["Bx"], By0=r0["By"], domain=dom, bc=2, M=r0["M"],
             tend=tend, uh=pack(r["uh"]), Bx=r["Bx"], By=r["By"], t=r["t"], count=r["count"],
             umax=r["umax"], totaldiv=r["totaldiv"], pmin=r["pmin"])


if __name__ == "__main__":
    if sys.argv[1:] == ["ic"]:
        main_ic()
    else:
        main()
        main_ic()
