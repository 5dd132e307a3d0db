"""Precision of the binary32 builds against a binary64 run of the same scheme.

    python cubism-jr-gpu-mpcf_amd/precision.py [--n 512] [--steps 200] [--bubbles 50] [--json out.json] [--md out.md]

Workload: the bubble-cloud collapse (mpcf_ic_bubble_cloud: 50 air bubbles,
lognormal radii mu = ln 0.02, sigma = 0.3, centres in a ball of radius 0.3,
water at p = 100 and bubbles at p = 1), absorbing boundaries, CFL 0.3.

The reference run is MPCF_FLAG_FP64: storage, WENO/HLLC arithmetic and the
speed-of-sound / dt reduction in binary64, starting from the binary32 initial
state widened exactly.  Each binary32 variant then advances the same initial
state with the reference's dt sequence (dt_max = dt64[k] with a CFL large
enough that the bound is what binds), so the error measured is that of the
scheme's arithmetic and storage, not a drift in time.  Separately, the dt each
variant's own reduction would have chosen is compared with dt64[k].

Variants:
  parity : the bit-exact build (the reference's binary32 arithmetic)
  fast   : FMA-contracted binary32 arithmetic
  p64    : fast + binary64 pressure recovery p = (E - KE - P)/G and binary64
           speed-of-sound / dt reduction (MPCF_FLAG_P64)
  p64e   : p64 with E stored as a binary32 pair, E = hi + lo (MPCF_FLAG_E_PAIR)
Reported per field: relative L1 = sum|u32 - u64| / sum|u64| and relative
Linf = max|u32 - u64| / max|u64| over the interior, plus the difference of
the conserved integrals (mass, momenta, energy) from the binary64 run's,
|I32 - I64| / (h^3 sum|u64|), all evaluated on the device in binary64
(mpcf_compare, mpcf_integrals).  The pressure p = (E - |rho u|^2/(2 rho) - P)/G
of each variant's final state (in binary64 on the host, E = hi + lo for
p64e) is compared with the binary64 run's the same way.  The storage floor is
the error that rounding the binary64 run's final state to binary32 would
give by itself: E rounded alone, and p recomputed from that rounded state.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

import pympcf

FIELDS = ("rho", "rho_u", "rho_v", "rho_w", "E", "G", "P")
VARIANTS = {
    "parity": dict(),
    "fast": dict(fast=True),
    "p64": dict(p64=True),
    "p64e": dict(epair=True),
}


def pressure(u):
    """p = (E - |rho u|^2 / (2 rho) - P) / G of a (7, nz, ny, nx) binary64 state"""
    ke = (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / (2.0 * u[0])
    return (u[4] - ke - u[6]) / u[5]


def rel_errs(got, ref):
    d = abs(got - ref)
    return float(d.sum() / abs(ref).sum()), float(d.max() / abs(ref).max())


def run(n=512, steps=200, nbubbles=50, seed=0, cfl=0.3, variants=("parity", "fast", "p64", "p64e"), log=None):
    log = log or (lambda *a: None)
    h = 1.0 / n
    t0 = time.perf_counter()
    fields, bubbles = pympcf.ic_bubble_cloud(n, n, n, nbubbles, seed)
    log(f"initial state {n}^3 with {nbubbles} bubbles in {time.perf_counter() - t0:.1f} s")
    out = {"workload": f"bubble-cloud collapse {n}^3, {nbubbles} bubbles, {steps} LSRK3 steps, CFL {cfl}",
           "n": n, "steps": steps, "variants": {}}
    with pympcf.Block(n, n, n, h, fp64=True) as ref:
        ref.data = fields
        i0 = ref.integrals()
        dts, sos64 = [], []
        t0 = time.perf_counter()
        for k in range(steps):
            dts.append(ref.step(cfl))
            sos64.append(ref.last_sos64())
        t64 = time.perf_counter() - t0
        i64 = ref.integrals()
        log(f"binary64 run: {steps} steps in {t64:.1f} s, t_end = {sum(dts):.6e}")
        out["fp64"] = {"seconds": t64, "ms_per_step": 1e3 * t64 / steps, "t_end": sum(dts),
                       "integrals_start": i0, "integrals_end": i64}
        u64 = ref.download_f64()
        p64 = pressure(u64)
        # the binary32 storage floor: the binary64 final state rounded to binary32
        e32 = u64[4].astype(np.float32).astype(np.float64)
        u32 = u64.astype(np.float32).astype(np.float64)
        out["storage_floor"] = {"E": rel_errs(e32, u64[4]), "p": rel_errs(pressure(u32), p64)}
        del e32, u32
        for name in variants:
            kw = VARIANTS[name]
            with pympcf.Block(n, n, n, h, **kw) as b:
                b.data = fields
                dt_rel = 0.0
                t0 = time.perf_counter()
                for k in range(steps):
                    dt = b.step(2 * cfl, dt_max=dts[k])
                    if dt != dts[k]:
                        raise RuntimeError(f"{name}: step {k} took dt {dt!r}, expected {dts[k]!r}")
                    # the dt this variant's own reduction gives at CFL cfl
                    s = b.last_sos64() if (kw.get("p64") or kw.get("epair")) else b.last_dt()[1]
                    own = cfl * h / s
                    dt_rel = max(dt_rel, abs(own - dts[k]) / dts[k])
                tv = time.perf_counter() - t0
                cmp = b.compare(ref)
                iv = b.integrals()
                pv = rel_errs(pressure(b.download_f64()), p64)
            res = {
                "seconds": tv,
                "rel_l1": {f: cmp[q]["rel_l1"] for q, f in enumerate(FIELDS)},
                "rel_linf": {f: cmp[q]["rel_linf"] for q, f in enumerate(FIELDS)},
                # |I32 - I64| / (h^3 sum|u64|): the integral's difference on the
                # scale of the field itself (net momentum is close to zero, so
                # a relative difference of the integral would say nothing)
                "integral_diff": {f: abs(iv[q] - i64[q]) / (cmp[q]["l1_ref"] * h ** 3) if cmp[q]["l1_ref"] else 0.0
                                  for q, f in enumerate(FIELDS)},
                "dt_own_rel_err_max": dt_rel,
                "p_rel_l1": pv[0],
                "p_rel_linf": pv[1],
            }
            out["variants"][name] = res
            log(f"{name}: {tv:.1f} s; rel L1 " + " ".join(f"{f}={res['rel_l1'][f]:.2e}" for f in FIELDS))
            log(f"{name}: rel Linf " + " ".join(f"{f}={res['rel_linf'][f]:.2e}" for f in FIELDS)
                + f"; own-dt rel err {dt_rel:.2e}; p rel L1 {pv[0]:.2e} Linf {pv[1]:.2e}")
    return out


def to_markdown(res, command):
    """The report as a short markdown summary."""
    F = FIELDS
    L = [f"# Precision: binary32 builds against a binary64 run of the same scheme, {res['workload'].split(',')[0]}\n",
         f"Command (repo root, MI355X):\n\n    {command}\n",
         f"Workload: {res['workload']}; absorbing boundaries.",
         f"binary64 run (MPCF_FLAG_FP64): {res['fp64']['ms_per_step']:.1f} ms/step, t_end = {res['fp64']['t_end']:.6e}.",
         "Every binary32 build replays the binary64 dt sequence, so the numbers below are arithmetic and storage "
         "error, not a time offset.  Errors are over the interior, in binary64, on the device (mpcf_compare).\n"]

    def table(title, key, fmt):
        L.append(title + "\n")
        L.append("| build | " + " | ".join(F) + " |")
        L.append("|---|" + "---|" * len(F))
        for name, v in res["variants"].items():
            L.append(f"| {name} | " + " | ".join(format(v[key][f], fmt) for f in F) + " |")
        L.append("")

    table("Relative L1 = sum|u32 - u64| / sum|u64|:", "rel_l1", ".2e")
    table("Relative Linf = max|u32 - u64| / max|u64|:", "rel_linf", ".2e")
    table("Conserved integrals after the run, |I32 - I64| / (h^3 sum|u64|):", "integral_diff", ".1e")
    fl = res["storage_floor"]
    L.append("Pressure p = (E - |rho u|^2/(2 rho) - P)/G of the final state (binary64 on the host), against the "
             "binary64 run's, beside E; the storage floor is the binary64 final state rounded to binary32:\n")
    L.append("| build | E rel L1 | E rel Linf | p rel L1 | p rel Linf |")
    L.append("|---|---|---|---|---|")
    for name, v in res["variants"].items():
        L.append(f"| {name} | {v['rel_l1']['E']:.2e} | {v['rel_linf']['E']:.2e} | {v['p_rel_l1']:.2e} | "
                 f"{v['p_rel_linf']:.2e} |")
    L.append(f"| binary32 storage floor | {fl['E'][0]:.2e} | {fl['E'][1]:.2e} | {fl['p'][0]:.2e} | {fl['p'][1]:.2e} |")
    L.append("")
    L.append("| build | own dt, max relative error over the run | seconds |")
    L.append("|---|---|---|")
    for name, v in res["variants"].items():
        L.append(f"| {name} | {v['dt_own_rel_err_max']:.2e} | {v['seconds']:.1f} |")
    return "\n".join(L) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--bubbles", type=int, default=50)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--json", default=None)
    ap.add_argument("--md", default=None, help="also write the report as markdown")
    a = ap.parse_args(argv)
    res = run(a.n, a.steps, a.bubbles, a.seed, log=lambda *m: print("[precision]", *m, flush=True))
    s = json.dumps(res, indent=1)
    if a.json:
        with open(a.json, "w") as f:
            f.write(s)
    else:
        print(s)
    if a.md:
        cmd = f"python3 cubism-jr-gpu-mpcf_amd/precision.py --n {a.n} --steps {a.steps} --bubbles {a.bubbles} --md <out.md>"
        with open(a.md, "w") as f:
            f.write(to_markdown(res, cmd))


if __name__ == "__main__":
    sys.exit(main())
