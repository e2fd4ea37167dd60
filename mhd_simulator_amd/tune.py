"""Autotuner of the projection's launch shapes and cycle schedule (no counterpart in the
reference, whose WGSL workgroup shapes and V-cycle are fixed).

    from mhd_simulator_amd import tune
    best = tune.tune_projection(pr, u_star, refine_tol=1e-8)   # times the candidates on the GPU
    tune.apply(best)                                            # process-wide launch shapes
    pr.project(u_star, precision="mixed", **best["opts"])       # cycle schedule
    tune.save(best, "tuning.json"); tune.apply(tune.load("tuning.json", key=best["key"]))

A candidate is a set of launch shapes (mhdp_set_launch_tuning: the z-marching chunk length and
block cap, the brick kernels' occupancy target) and a cycle schedule (the smoother - the
reference's weighted Jacobi or red-black Gauss-Seidel -, its weight, the coarse-correction
scale, level-1 visits per cycle, the mixed solver's inner tolerance).  Every candidate runs one
untimed projection, then `repeats` timed ones (wall clock around a synchronized call; inputs
resident on the GPU); it must converge with the cycle it asked for (a cycle that breaks down and
is solved again with the reference's V-cycle is a failed candidate), and the fastest median wins.
`tune_projection` searches in two stages by default: the schedules at the built-in launch shapes,
then the launch shapes at the winning schedule.  `tune_batch` does the same for one
mhdp_project_batch call over many problems.  The launch shapes are process-wide: tune while
nothing else projects, then apply the winner.
"""
from __future__ import annotations

import ctypes as C
import itertools
import json
import os
import statistics
import time

from . import _lib

# the candidate axes (the defaults are the library's built-in shapes: z 32, cap 1024, occupancy 2)
MARCH_Z = (16, 32, 64)
MARCH_CAP = (1024, 2048)
BRICK_OCC = (2, 3)
COARSE_VISITS = (0, 1, 2)  # 0: the library's automatic choice by grid size
INNER_TOL = (0.0,)  # the mixed solver's inner FP32 tolerance; 0: the library's (1e-4)
BUILTIN_LAUNCH = {"march_z": 32, "march_cap": 1024, "brick_occ": 2}
# the schedule stage's axes (0: the library's tuned default for the smoother: weight 0.9, scale
# 1.6 with Jacobi, 1.8 with red-black, 1.9 from 2^24 cells)
SMOOTHER = ("jacobi", "red-black")
SMOOTH_WEIGHT = (0.0, 0.8, 0.95)   # the library accepts weights in [0, 1)
COARSE_SCALE = (0.0, 1.4, 1.8, 1.95)  # and scales in [0, 2)


def current() -> dict:
    """The launch shapes in force (mhdp_get_launch_tuning)."""
    t = _lib.LaunchTuning()
    rc = _lib.load().mhdp_get_launch_tuning(C.byref(t))
    if rc != _lib.OK:
        raise RuntimeError(f"mhdp_get_launch_tuning failed ({rc})")
    return t.as_dict()


def set_launch(shapes: dict) -> None:
    """Replace the process-wide launch shapes; unknown keys are an error, missing keys keep their
    current value."""
    now = current()
    unknown = set(shapes) - set(now)
    if unknown:
        raise KeyError(f"unknown launch shape(s): {sorted(unknown)}")
    now.update(shapes)
    t = _lib.LaunchTuning(**now)
    rc = _lib.load().mhdp_set_launch_tuning(C.byref(t))
    if rc != _lib.OK:
        raise ValueError(f"invalid launch shapes {now}")


def candidates(march_z=MARCH_Z, march_cap=MARCH_CAP, brick_occ=BRICK_OCC, coarse_visits=COARSE_VISITS,
               inner_tol=INNER_TOL, smoother=(None,), smooth_weight=(0.0,), coarse_scale=(0.0,)):
    """Every combination of the axes: [{"launch": {...}, "opts": {...}}, ...].  A schedule axis
    at its default (None / 0) leaves the option out: the library's choice."""
    out = []
    for z, cap, occ, visits, tau, sm, w, sc in itertools.product(march_z, march_cap, brick_occ, coarse_visits,
                                                               inner_tol, smoother, smooth_weight, coarse_scale):
        opts = {"coarse_visits": int(visits)}
        if tau:
            opts["inner_tol"] = float(tau)
        if sm is not None:
            opts["smoother"] = str(sm)
        if w:
            opts["smooth_weight"] = float(w)
        if sc:
            opts["coarse_scale"] = float(sc)
        out.append({"launch": {"march_z": int(z), "march_cap": int(cap), "brick_occ": int(occ)}, "opts": opts})
    return out


def schedules(smoother=SMOOTHER, smooth_weight=SMOOTH_WEIGHT, coarse_scale=COARSE_SCALE,
              coarse_visits=COARSE_VISITS, inner_tol=INNER_TOL):
    """The schedule stage: every cycle schedule at the built-in launch shapes."""
    return candidates(march_z=(BUILTIN_LAUNCH["march_z"],), march_cap=(BUILTIN_LAUNCH["march_cap"],),
                      brick_occ=(BUILTIN_LAUNCH["brick_occ"],), coarse_visits=coarse_visits,
                      inner_tol=inner_tol, smoother=smoother, smooth_weight=smooth_weight,
                      coarse_scale=coarse_scale)


def launches(opts, march_z=MARCH_Z, march_cap=MARCH_CAP, brick_occ=BRICK_OCC):
    """The launch stage: every launch shape with one schedule."""
    return [{"launch": {"march_z": int(z), "march_cap": int(cap), "brick_occ": int(occ)}, "opts": dict(opts)}
            for z, cap, occ in itertools.product(march_z, march_cap, brick_occ)]


def search(run, cands, repeats=3, clock=time.perf_counter):
    """Time run(candidate) -> ok for every candidate (one untimed call first) and return
    (best, table).  run may return a dict of solve statistics instead of True: the last one is
    kept in the row.  A candidate whose run raises or returns False is kept in the table as failed
    and never wins; the launch shapes in force before the search are restored at the end."""
    if repeats < 1:
        raise ValueError("repeats must be >= 1")
    before = current()
    table = []
    try:
        for cand in cands:
            set_launch(cand["launch"])
            row = dict(cand, ms=None, ok=False)
            try:
                out = run(cand)
                ok = bool(out)
                times = []
                for _ in range(repeats):
                    t0 = clock()
                    out = run(cand) if ok else out
                    times.append((clock() - t0) * 1e3)
                    ok = ok and bool(out)
                if ok and isinstance(out, dict):
                    row["stats"] = out
                row["ok"] = ok
                row["ms"] = statistics.median(times) if ok else None
            except Exception as e:  # a failing candidate is recorded, not fatal
                row["error"] = str(e)
            table.append(row)
    finally:
        set_launch(before)
    good = [r for r in table if r["ok"]]
    if not good:
        raise RuntimeError("no candidate converged")
    best = min(good, key=lambda r: r["ms"])
    return best, table


def _key(dims_list, precision, refine_tol):
    dims = ";".join("x".join(str(int(v)) for v in d) for d in dims_list)
    return f"{dims}|{precision}|{refine_tol:g}"


def two_stage(run, repeats, schedule_cands=None, launch_axes=None, clock=time.perf_counter):
    """The schedule stage, then the launch stage at its winner: (best, table of both stages)."""
    best1, table1 = search(run, schedule_cands if schedule_cands is not None else schedules(), repeats, clock)
    best2, table2 = search(run, launches(best1["opts"], **(launch_axes or {})), repeats, clock)
    best = best2 if best2["ms"] <= best1["ms"] else best1
    return best, table1 + table2


def tune_projection(pr, u_star, *, precision="mixed", refine_tol=1e-8, repeats=3, cands=None, schedule=None,
                    launch_axes=None, **project_kw):
    """Tune one PressureProjection on a device-resident u_star (torch tensor on the GPU): the two
    stages (schedule: the first stage's candidates, default schedules(); launch_axes: the second
    stage's axes for launches()), or exactly the candidates `cands` when given."""
    import torch

    u_out = torch.empty_like(u_star)

    def run(cand):
        pr.project(u_star, precision=precision, refine_tol=refine_tol, u_out=u_out, **cand["opts"], **project_kw)
        torch.cuda.synchronize()
        if pr.stats.get("cycle_fallback"):
            raise RuntimeError("the cycle broke down (solved again with the reference's V-cycle)")
        return {k: pr.stats[k] for k in ("iterations", "outer_steps") if k in pr.stats} or True

    best, table = two_stage(run, repeats, schedule, launch_axes) if cands is None else search(run, cands, repeats)
    return {"key": _key([pr.grid.dims], precision, refine_tol), "launch": best["launch"], "opts": best["opts"],
            "ms": best["ms"], "table": table}


def tune_batch(projections, u_stars, *, refine_tol=1e-8, repeats=3, cands=None, schedule=None, launch_axes=None,
               **batch_kw):
    """Tune one mhdp_project_batch call over `projections` (device-resident u_stars)."""
    import torch

    from .fluid import project_batch

    def run(cand):
        project_batch(projections, u_stars, precision="mixed", refine_tol=refine_tol, **cand["opts"], **batch_kw)
        torch.cuda.synchronize()
        if any(p.stats.get("cycle_fallback") for p in projections):
            raise RuntimeError("a cycle broke down (solved again with the reference's V-cycle)")
        return True

    best, table = two_stage(run, repeats, schedule, launch_axes) if cands is None else search(run, cands, repeats)
    return {"key": _key([p.grid.dims for p in projections], "mixed", refine_tol), "launch": best["launch"],
            "opts": best["opts"], "ms": best["ms"], "table": table}


def apply(result: dict) -> dict:
    """Put a tuning result's launch shapes in force; returns the projection options to pass."""
    set_launch(result["launch"])
    return dict(result["opts"])


def save(result: dict, path: str) -> None:
    """Add (or replace) `result` under its key in the JSON file at `path`."""
    data = {}
    if os.path.exists(path):
        with open(path) as f:
            data = json.load(f)
    data[result["key"]] = {k: result[k] for k in ("key", "launch", "opts", "ms")}
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def load(path: str, key: str) -> dict:
    """The saved result for `key` (KeyError when the file holds none)."""
    with open(path) as f:
        data = json.load(f)
    return data[key]


def main(argv=None):
    """python -m mhd_simulator_amd.tune [--workload plate|batch64] [--size N] [--out tuning.json]:
    tune on a synthetic workload of the given shape (plate: the two-phase droplets over a plate,
    mixed to 1e-8; batch64: the 64 uneven droplet problems in one batch call) and print every
    candidate's time and the winner."""
    import argparse

    import torch

    from . import configs as Cf
    from .fluid import FluidConfig, PressureProjection
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--workload", choices=["plate", "batch64"], default="plate")
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--out", default="")
    ap.add_argument("--inner-tol", default="", help="comma list of inner tolerances to sweep (default: the library's)")
    ap.add_argument("--builtin-launch", action="store_true", help="sweep the schedule only, at the built-in shapes")
    ap.add_argument("--launch-only", action="store_true",
                    help="sweep the launch shapes and level-1 visits only, with the library's cycle (no two stages)")
    a = ap.parse_args(argv)
    axes = {}
    if a.inner_tol:
        axes["inner_tol"] = tuple(float(v) for v in a.inner_tol.split(","))
    if a.builtin_launch:
        cands = schedules(**axes)
    elif a.launch_only:
        cands = candidates(**axes)
    else:
        cands = None  # two stages: schedules, then launch shapes at the winner
    torch.cuda.set_device(0)
    if a.workload == "plate":
        p = Cf.config5(a.size, droplets=max(1, round(200 * (a.size / 512) ** 3)))
        cfg = FluidConfig(relative_tolerance=1e-7, absolute_divergence_tolerance=0.0, max_iterations=4000)
        prs = [PressureProjection(p.grid, p.fluid, p.alpha, cfg)]
        u = torch.from_numpy(p.ustar).cuda()
        res = tune_projection(prs[0], u, refine_tol=1e-8, repeats=a.repeats, cands=cands, precond="mg")
    else:
        ps = Cf.config4()
        prs = [PressureProjection(q.grid, q.fluid, q.alpha) for q in ps]
        us = [torch.from_numpy(q.ustar).cuda() for q in ps]
        res = tune_batch(prs, us, refine_tol=1e-8, repeats=a.repeats, cands=cands, precond="mg")
    # the library's automatic smoother: red-black on grids of at least 2^15 cells, else Jacobi
    auto_rb = a.workload == "batch64" or a.size ** 3 >= (1 << 15)
    builtin = ({"coarse_visits": 0}, {"coarse_visits": 0, "smoother": "red-black" if auto_rb else "jacobi"})
    default = next((r for r in res["table"] if r["launch"] == BUILTIN_LAUNCH and r["opts"] in builtin), None)
    for r in sorted(res["table"], key=lambda r: (r["ms"] is None, r["ms"] or 0.0)):
        ms = f"{r['ms']:.2f} ms" if r["ms"] is not None else f"failed ({r.get('error', 'no convergence')})"
        its = f" {json.dumps(r['stats'])}" if "stats" in r else ""
        print(f"{json.dumps(r['launch'])} {json.dumps(r['opts'])}: {ms}{its}", flush=True)
    print(json.dumps({k: res[k] for k in ("key", "launch", "opts", "ms")}), flush=True)
    if default is not None:
        print(f"built-in shapes and schedule: {default['ms']:.2f} ms", flush=True)
    if a.out:
        save(res, a.out)
    for pr in prs:
        pr.close()


if __name__ == "__main__":
    main()
