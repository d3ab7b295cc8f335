"""Precision study of the mixed form on the smooth MHD vortex.

    python freljord-cfd-arsenal_amd/precision_study.py [--n 256 512 1024 2048] [--tend 1.0]
        [--domain XA XB YA YB] [--out study.json]

The vortex of init1.txt (rho = 1, p = 1, u_inf = (1, 1), magnetic
perturbation of amplitude 1, gamma = 5/3, periodic; the domain of
init1.txt:41-44 is [-10, 10]^2, --domain overrides it) is advanced with
HLLD, TVB on every cell with M = 1e6 (init1.txt:55, which leaves this smooth
flow unlimited) and CFL 0.15 (calculate_dt.f90:8) to tend at each
resolution, once in plain FP64 and once in the mixed form.  Reported per
resolution:
- the L2 error of each form against the exact solution (calculate_L2error,
  all eight fields) and the convergence order between resolutions;
- the relative L2 difference of the mixed form against FP64 for every field.
  For the orthogonal Legendre basis, the L2 norm of a cell's polynomial is
  the norm of its modal coefficients weighted by the diagonal mass matrix
  1, 1/3, 1/3, 4/45, 1/9, 4/45 of the six P2 modes, continued by
  4/175, 4/135, 4/135, 4/175 for the cubic modes that Bx and By carry
  (get_basis.f90:60-159), which is the weighting used here.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FIELDS = ("rho", "rhou", "rhov", "rhow", "E", "Bx", "By", "Bz")
MODE_NORM = np.array([1.0, 1 / 3, 1 / 3, 4 / 45, 1 / 9, 4 / 45, 4 / 175, 4 / 135, 4 / 135, 4 / 175])   # get_basis.f90:60-159


def _pkg():
    name = "freljord_cfd_arsenal_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(HERE, "__init__.py"),
                                                  submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def field_rel_l2(a, b):
    """relative L2 difference of every field between two modal arrays
    (interior cells, mass-matrix weighted)"""
    a, b = a[1:-1, 1:-1], b[1:-1, 1:-1]
    w = np.sqrt(MODE_NORM[:a.shape[2]])[None, None, :]
    out = {}
    for n, nm in enumerate(FIELDS):
        bn = b[:, :, :, n] * w
        den = np.linalg.norm(bn)
        num = np.linalg.norm(a[:, :, :, n] * w - bn)
        out[nm] = float(num / den) if den > 0 else float(num)
    return out


def run_vortex(N, precision, tend, chunk=0.1, log=None, domain=None):
    """advance the vortex to tend; returns (modal array, L2 errors, steps, seconds, div B)"""
    P = _pkg()
    d = P.ic_defaults(P.IC_VORTEX)
    s = P.Solver(N, N, flux_type=P.HLLD, bc=d["bc"], domain=tuple(domain or d["domain"]), M=d["M"], CFL=0.15,
                 pp=0, indicator=P.IND_ALL, precision=precision)
    try:
        s.init_data(P.IC_VORTEX)
        s.sync()
        t0 = time.perf_counter()
        s.rk3_prologue()
        steps, t, mark = 0, 0.0, chunk
        while t < tend:
            t = s.step(tend, sync=True)
            steps += 1
            if log and t >= mark:
                log(f"  N={N} {'mixed' if precision == P.MIXED else 'fp64'}: t={t:.3f} steps={steps}")
                mark += chunk
        sec = time.perf_counter() - t0
        err = s.calculate_L2error(t)
        uh = s.download()[0]
        div = s.diagnostics()["totaldiv"]
        return uh, err, steps, sec, div
    finally:
        s.close()


def study(Ns, tend, chunk=0.1, log=None, domain=None):
    P = _pkg()
    rows = []
    for N in Ns:
        u64, e64, n64, s64, d64 = run_vortex(N, P.FP64, tend, chunk, log, domain)
        umx, emx, nmx, smx, dmx = run_vortex(N, P.MIXED, tend, chunk, log, domain)
        rows.append({"N": N, "steps": n64, "steps_mixed": nmx, "sec_fp64": s64, "sec_mixed": smx,
                     "L2_fp64": [float(x) for x in e64], "L2_mixed": [float(x) for x in emx],
                     "mixed_vs_fp64": field_rel_l2(umx, u64), "divB_fp64": d64, "divB_mixed": dmx})
        del u64, umx
        if log:
            r = rows[-1]
            log(f"N={N}: L2(rho) fp64 {r['L2_fp64'][0]:.3e} mixed {r['L2_mixed'][0]:.3e}; "
                f"mixed vs fp64 rho {r['mixed_vs_fp64']['rho']:.2e}; {r['sec_fp64']:.1f} s / {r['sec_mixed']:.1f} s")
    for a, b in zip(rows, rows[1:]):
        r = b["N"] / a["N"]
        for k in ("fp64", "mixed"):
            b["order_" + k] = [float(np.log(x / y) / np.log(r)) if x > 0 and y > 0 else None
                               for x, y in zip(a["L2_" + k], b["L2_" + k])]
    return rows


def markdown(rows, tend, domain):
    out = [f"Smooth MHD vortex on {domain}, HLLD, CFL 0.15, t = {tend}: L2 error of rho against the exact solution, "
           "convergence order, and the mixed form against FP64 (relative L2, every field)", "",
           "| N | steps | L2 rho FP64 | order | L2 rho mixed | order | " + " | ".join(FIELDS) + " | div B mixed | s FP64 | s mixed |",
           "|---|---|---|---|---|---|" + "---|" * len(FIELDS) + "---|---|---|"]
    for r in rows:
        o64 = r.get("order_fp64", [None])[0]
        omx = r.get("order_mixed", [None])[0]
        f = lambda x: "-" if x is None else f"{x:.2f}"
        out.append(f"| {r['N']}² | {r['steps']} | {r['L2_fp64'][0]:.3e} | {f(o64)} | {r['L2_mixed'][0]:.3e} | {f(omx)} | "
                   + " | ".join(f"{r['mixed_vs_fp64'][k]:.1e}" for k in FIELDS)
                   + f" | {r['divB_mixed']:.1e} | {r['sec_fp64']:.1f} | {r['sec_mixed']:.1f} |")
    return "\n".join(out)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--n", type=int, nargs="+", default=[256, 512, 1024, 2048])
    ap.add_argument("--tend", type=float, default=1.0)
    ap.add_argument("--chunk", type=float, default=0.1, help="time between progress lines")
    ap.add_argument("--domain", type=float, nargs=4, metavar=("XA", "XB", "YA", "YB"),
                    help="periodic domain (default: init1.txt's [-10, 10]^2)")
    ap.add_argument("--out", help="write the rows as JSON")
    a = ap.parse_args()
    dom = tuple(a.domain) if a.domain else _pkg().ic_defaults(_pkg().IC_VORTEX)["domain"]
    rows = study(a.n, a.tend, a.chunk, log=lambda m: print(m, flush=True), domain=dom)
    print(markdown(rows, a.tend, dom), flush=True)
    if a.out:
        with open(a.out, "w") as f:
            json.dump({"tend": a.tend, "domain": dom, "rows": rows}, f, indent=1)


if __name__ == "__main__":
    main()
