"""Run a Lua `simulation` configuration on the device, as the reference's
shoccs binary does (src/app/shoccs.cpp: input file, --check): the
statistics of every step go to a CSV file with the columns of the reference's
system.csv (src/systems/heat.cpp:46-48), and the configuration's io table
writes the U and Error fields, the cut-point geometry and the XDMF3 series
(src/io/field_io.cpp).

    python -m shoccs_amd config.lua [--check] [--csv logs/system.csv] [--fused] [--no-io]
"""
from __future__ import annotations

import argparse
import os
import sys
import time


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m shoccs_amd")
    ap.add_argument("input_file")
    ap.add_argument("--check", action="store_true", help="read the configuration and the problem, do not run")
    ap.add_argument("--csv", default=os.path.join("logs", "system.csv"), help="statistics file ('' for none)")
    ap.add_argument("--fused", action="store_true", help="fused RK4 form (integrator.form = 'fused')")
    ap.add_argument("--no-io", action="store_true", help="ignore the configuration's io table")
    args = ap.parse_args(argv)

    import shoccs_amd as S

    cfg = S.Config.from_file(args.input_file)
    if args.check:
        prob = cfg.problem(None)
        print(f"configuration ok: kind {cfg.kind}, mesh {cfg.n}, {len(cfg.spheres)} shape(s), "
              f"{prob.N} points, |R| = {prob.nR}, dt = {prob.dt(cfg.h_cfl, cfg.p_cfl):.6g}")
        prob.close()
        return 0
    ctx = S.Context(0)
    prob = cfg.problem(ctx)
    if cfg.eigenvalues:
        # hyperbolic_eigenvalues::stats and its log (Timestamp,MaxEigen)
        from . import spectral

        stat = spectral.hyperbolic_eigenvalue_stat(prob)
        if args.csv:
            os.makedirs(os.path.dirname(args.csv) or ".", exist_ok=True)
            with open(args.csv, "w") as f:
                f.write(f"Timestamp,MaxEigen\n{time.strftime('%Y-%m-%d %H:%M:%S')},{stat:.17g}\n")
        print(f"max eigen statistic {stat:.6e}")
        prob.close()
        ctx.close()
        return 0
    u = prob.field()
    kw = cfg.run_kwargs()
    if args.fused and kw["integrator"] == S.RK4:
        kw["integrator"] = S.RK4_FUSED
    io = S.FieldIO(prob, **cfg.io) if cfg.io and not args.no_io else None
    t0 = time.perf_counter()
    res = prob.run(u, io=io, **kw)
    wall = (time.perf_counter() - t0) * 1e3
    if args.csv:
        os.makedirs(os.path.dirname(args.csv) or ".", exist_ok=True)
        dt = prob.dt(cfg.h_cfl, cfg.p_cfl)
        with open(args.csv, "w") as f:
            f.write("Timestamp,Time,Step,Linf,Min,Max,Domain_Linf,Domain_ic,Rx_Linf,Rx_ic,Ry_Linf,Ry_ic,"
                    "Rz_Linf,Rz_ic,Wall_ms\n")
            stamp = time.strftime("%Y-%m-%d %H:%M:%S")
            for k, row in enumerate(res["stats"]):
                tk = res["time"] if k == len(res["stats"]) - 1 else k * dt
                vals = ",".join(f"{v:.17g}" for v in row)
                f.write(f"{stamp},{tk:.17g},{k},{vals},{wall if k == len(res['stats']) - 1 else 0:.3f}\n")
    print(f"steps {res['step']}, time {res['time']:.6g}, completed {res['completed']}, "
          f"linf {res['linf']:.6e}")
    if io is not None:
        io.close()
    prob.close()
    ctx.close()
    return 0 if res["completed"] else 1


if __name__ == "__main__":
    sys.exit(main())
