#!/usr/bin/env python3
"""Benchmark of the pfc-relax relaxation loop on one MI355X.

Workload: 3-component graphene/hBN heterostructure PFC on a 4096 x 4096
periodic grid (hetero.in parameters, Gaussian coupling width ~a/2, seeded
50/50 patchwork of ~256-point domains, noise +-0.01, conserved dynamics,
dt = 0.25), fields resident in HBM.  A "step" is one semi-implicit update of
the grid (the reference's step(), src/pfc-relax.c:1133-1247).  Each timed
window is K consecutive steps run as relax() runs them: it starts at a
multiple of 100, so it holds the t % 100 refresh steps, and it evaluates the
free energy before every step t with t % 100 == 0 (hetero.in prints every 100
steps).  The window is bracketed by stream synchronisation.

The headline `value` is the FP64 form, the reference's precision.  The mixed
FP32/FP64 form runs the same window beside it (`mixed`), with its distance
from FP64 (`mixed_vs_fp64_error` after the windows, `mixed_vs_fp64_run` over
the workload's 5000 steps).

Prints one JSON line.  `roofline` is for the dominant kernel of the FP64 step
(the longer of its row and column passes), timed with HIP events on the
engine's stream over a second window of plain (non-refresh) steps; it also
carries both passes' fractions and the whole step's fraction under the
three-sweep byte model.
`cpu_baseline` times the CPU oracle on a bounded sample of the same workload;
`batch` is the batch form on 96 systems of uneven sizes (2^k and 3*2^k sides
from 256 to 2048, one or three fields), beside the same systems stepped one
by one; `other_configs` times the 2048^2 polycrystal, the 8192^2
heterostructure and a 1000^2 polycrystal (a side only the generic path
takes) in both forms.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0   # MI355X HBM3E
TPRINT = 100            # hetero.in's print interval: one evaluate per 100 steps


def col_bytes(Nx, Ny, U, esz):
    """Algorithmic HBM bytes of one column-pass launch: per field it reads the
    row transforms J and writes the inverse-column transforms I ((Nx/2) x Ny
    complex each, DC and Nyquist packed in column 0) and reads and writes the
    state N ((Nx/2 + 1) x Ny complex)."""
    return 2 * U * (Nx // 2 + (Nx // 2 + 1)) * Ny * esz


def row_bytes(Nx, Ny, U, esz):
    """Row pass: per field reads I and writes J, (Nx/2) x Ny complex each."""
    return 2 * U * (Nx // 2) * Ny * esz


def survey_bytes(Nx, Ny, U, real_size):
    """Three-sweep model of the step: U (16 NxNy + 96 (Nx/2+1) Ny) B in FP64."""
    return U * (16 * Nx * Ny + 96 * (Nx // 2 + 1) * Ny) * real_size / 8


def progress(msg):
    """One line on stderr per finished phase of the benchmark."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def relax_window(g, t0, nsteps, tprint):
    """Steps t0 .. t0 + nsteps - 1 as relax() runs them (src/pfc-relax.c:
    1292-1312): an evaluate (the free energy and means that print_output
    logs) before every step t with t % tprint == 0, and each step t with
    t % 100 == 0 refreshing the spectral state.  hetero.in writes no energy
    density file, so each print is pfc_step_evaluate (the evaluate's sums
    taken by the first step after it), as pfc-relax runs it.  Returns the
    number of evaluates."""
    t, end, nev = t0, t0 + nsteps, 0
    while t < end:
        n = min(end, (t // tprint + 1) * tprint) - t
        if t % tprint == 0:
            g.step_evaluate(t, n)
            nev += 1
        else:
            g.step(t, n)
        t += n
    return nev


def run_gpu(params, fields, Nx, Ny, dx, dy, precision, steps, warmup, tprint=TPRINT):
    """Warmup, then a window of `steps` steps starting at a multiple of 100, so
    that it holds a refresh step and an evaluate (one per tprint steps, as
    relax() runs them), timed by the wall clock alone (the figure reported).
    Then a second window of plain steps (no refresh, an even number, at most
    98) records HIP events around each pass on the engine's stream: the mean
    launch time of each pass's plain kernel, for the roofline."""
    from pfc_amd.engine import Engine
    g = Engine(params, Nx, Ny, dx, dy, precision)
    g.set_fields(fields)
    g.step(0, warmup)
    s0 = -(-warmup // 100) * 100
    g.step(warmup, s0 - warmup)
    g.nan_step()                       # synchronises the stream
    t0 = time.perf_counter()
    nev = relax_window(g, s0, steps, tprint)
    g.nan_step()
    wall = time.perf_counter() - t0
    s1 = -(-(s0 + steps) // 100) * 100
    g.step(s0 + steps, s1 - s0 + 1 - steps)     # through the refresh step s1
    g.nan_step()
    g.timing(True)
    g.step(s1 + 1, max(2, min(steps, 98) // 2 * 2))
    ms, n = g.timing_read()
    g.timing(False)
    assert g.nan_step() == -1, "NaN in the benchmark run"
    return g, wall, ms, n, nev, s0


def accuracy_study(params, fields, Nx, Ny, dx, dy, nsteps, checkpoints, atoms=True):
    """Mixed form against the FP64 form from the same start over the
    workload's full length: relative L2 of the densities and relative error
    of the free energy per area at each checkpoint.  Beside it two FP64
    controls against the FP64 form: (1) started from the FP32-rounded initial
    fields (one FP32 rounding, ~3e-8 relative), and (2) perturbed after every
    step by relative noise of the size of one mixed step's error (measured
    here: one mixed step and one FP64 step from the same FP64 state), the
    round-off the mixed form injects per step, in FP64 arithmetic.  If (2)
    strays as far as the mixed form, the distance is the relaxation's
    sensitivity to any per-step error of that size, not a defect of the FP32
    arithmetic."""
    from pfc_amd import FP64, MIXED
    from pfc_amd.engine import Engine
    gm = Engine(params, Nx, Ny, dx, dy, MIXED)
    gd = Engine(params, Nx, Ny, dx, dy, FP64)
    gp = Engine(params, Nx, Ny, dx, dy, FP64)
    gq = Engine(params, Nx, Ny, dx, dy, FP64)
    # one mixed step against one FP64 step from the same state
    gm.set_fields(fields)
    gd.set_fields(fields)
    gm.step(1, 1)
    gd.step(1, 1)
    x1 = gd.get_fields()
    one_step = float(np.linalg.norm(gm.get_fields() - x1) / np.linalg.norm(x1))
    del x1
    gm.set_fields(fields)
    gd.set_fields(fields)
    gp.set_fields(fields.astype(np.float32).astype(np.float64))
    gq.set_fields(fields)
    t = 0
    rows = []
    last = [c for c in checkpoints if c <= nsteps][-1]
    for c in checkpoints:
        if c > nsteps:
            break
        for s0 in range(t, c, 1000):
            n = min(1000, c - s0)
            for g in (gm, gd, gp):
                g.step(s0, n)
            gp.nan_step()
            progress(f"accuracy {Nx}x{Ny}: mixed, FP64 and rounded-start runs at step {s0 + n}")
        for s in range(t, c):
            gq.step(s, 1)
            gq.perturb(one_step, 1000003 * s + 17)
            if (s + 1) % 1000 == 0:
                gq.nan_step()
                progress(f"accuracy {Nx}x{Ny}: perturbed FP64 run at step {s + 1}")
        t = c
        fm, _ = gm.evaluate()
        fd, _ = gd.evaluate()
        fp, _ = gp.evaluate()
        fq, _ = gq.evaluate()
        xd = gd.get_fields()
        nd = np.linalg.norm(xd)
        xm = gm.get_fields()
        em = float(np.linalg.norm(xm - xd) / nd)
        xp = gp.get_fields()
        ep = float(np.linalg.norm(xp - xd) / nd)
        xq = gq.get_fields()
        eq = float(np.linalg.norm(xq - xd) / nd)
        row = {"step": c, "energy_per_area_rel": abs(fm - fd) / abs(fd),
               "density_rel_l2": em, "fave_fp64": fd,
               "fp64_from_fp32_rounded_start": {"energy_per_area_rel": abs(fp - fd) / abs(fd),
                                                "density_rel_l2": ep},
               "fp64_perturbed_every_step": {"energy_per_area_rel": abs(fq - fd) / abs(fd),
                                             "density_rel_l2": eq}}
        if c == last and atoms:
            row["atoms"] = atom_positions(xd, xm, xp, dx, dy, xq)
        progress(f"accuracy {Nx}x{Ny}: checkpoint {c} done")
        del xm, xp, xd, xq
        rows.append(row)
    for g in (gm, gd, gp, gq):
        g.close()
    return {"one_step_mixed_vs_fp64_density_rel_l2": one_step,
            "perturbation_per_step_rel": one_step, "checkpoints": rows}


def atom_positions(xd, xm, xp, dx, dy, xq):
    """Atoms of the summed density (the coordinator's conversion, lattice
    constant 1) of the FP64 run, and the displacement of each mixed-run and
    control-run atom from its nearest FP64 atom, in lattice constants."""
    from pfc_amd import LATTICE_A
    from pfc_amd.coordinator import compare_atoms, coordinate
    if not all(np.isfinite(x).all() for x in (xd, xm, xp, xq)):
        return None
    ad, box, nr, fl = coordinate(xd.sum(axis=0), dx, dy, LATTICE_A, 1.0)
    out = {"field": "sum of the 3 densities", "unit": "lattice constants", "rings_fp64": nr,
           "coordinator_flags_fp64": fl}
    for name, x in (("mixed", xm), ("fp64_from_fp32_rounded_start", xp), ("fp64_perturbed_every_step", xq)):
        a, _, _, _ = coordinate(x.sum(axis=0), dx, dy, LATTICE_A, 1.0)
        out[name] = compare_atoms(ad, a, box)
        progress(f"atoms of the {name} run compared")
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(params, fields, Nx, Ny, dx, dy, budget_s):
    """Times the FP64 CPU oracle (OpenMP) on a bounded number of steps, on
    the host's CPU share: OMP_NUM_THREADS where the environment sets it (the
    GPU box allots 16 host cores to its one GPU), else every core."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    o = Oracle(params, Nx, Ny, dx, dy)
    o.set_fields(fields)
    o.step(1)   # warm (page in)
    n = 0
    t0 = time.perf_counter()
    while True:
        o.step(2 + n)
        n += 1
        el = time.perf_counter() - t0
        if el > budget_s or n >= 50:
            break
    o.close()
    return {"value": Nx * Ny * n / el, "unit": "grid-point-steps/s", "cores": threads,
            "kind": "port", "cpu_model": cpu_model(), "host_cpus_visible": os.cpu_count(),
            "sample": f"{n} FP64 steps of the same {Nx}x{Ny} 3-field workload on the CPU oracle "
                      f"(oracle/pfc_oracle.c: the reference's step() restated in C, own "
                      f"radix-4/2/3 FFT, OpenMP, {threads} threads), {el:.1f} s"}


BATCH_SIDES = (256, 384, 512, 768, 1024, 1536, 2048)


def batch_bench(nsys, steps, warmup, seed=96):
    """The batch form on the batched-relaxation workload: nsys square systems,
    sides drawn from BATCH_SIDES (2^k and 3*2^k FFT lengths), 70 % one-field
    graphene / 30 % three-field graphene/hBN patchworks, random lattice
    orientations and noise, mixed precision, autotuned; beside it the same
    systems stepped one by one through pfc_step."""
    from pfc_amd import MIXED, LATTICE_A, graphene_params, hetero_params
    from pfc_amd.engine import BATCH_AUTOTUNE, Batch, Engine
    from pfc_amd.inputs import crystal, hetero_patchwork
    rng = np.random.default_rng(seed)
    specs = [(int(rng.choice(BATCH_SIDES)), 3 if rng.random() < 0.3 else 1) for _ in range(nsys)]
    t0 = time.perf_counter()
    engines = []
    for k, (n, U) in enumerate(specs):
        if U == 1:
            p = graphene_params()
            dx = LATTICE_A / 8.0
            f = crystal(n, n, dx, dx, 0.2, -0.2, LATTICE_A, sym=3,
                        theta=float(rng.uniform(0.0, np.pi / 3)), noise=0.01, seed=k)[None]
        else:
            p = hetero_params(dt=0.25, c=1.0, sigma=2.0 / 7.2552)
            dx = 0.7
            f = hetero_patchwork(n, n, dx, dx, domain=128, seed=k, noise=0.01)
        e = Engine(p, n, n, dx, dx, MIXED)
        e.set_fields(f)
        engines.append(e)
    gen_s = time.perf_counter() - t0
    gp = sum(n * n for n, _ in specs)
    t0 = time.perf_counter()
    b = Batch(engines, BATCH_AUTOTUNE)
    tune_s = time.perf_counter() - t0
    b.step(0, warmup)
    t0 = time.perf_counter()
    b.step(warmup, steps)
    wall = time.perf_counter() - t0
    shapes = {f"{n}x{n} U={U}": {"row_points": b.shape(0, n, U), "column_points": b.shape(1, n, U)}
              for n, U in sorted(set(specs))}
    b.close()
    t = warmup + steps
    for e in engines:
        e.step(t, 2)
    for e in engines:
        e.nan_step()
    t += 2
    t0 = time.perf_counter()
    for e in engines:
        e.step(t, steps)
    for e in engines:
        e.nan_step()
    wall1 = time.perf_counter() - t0
    for e in engines:
        e.close()
    # algorithmic bytes of one step over all systems: the row and column
    # passes' bytes of each (mixed form: FP32 complex elements, 8 B)
    step_bytes = sum(row_bytes(n, n, U, 8) + col_bytes(n, n, U, 8) for n, U in specs)
    gbs = step_bytes / (wall / steps) / 1e9
    gbs1 = step_bytes / (wall1 / steps) / 1e9
    return {
        "workload": f"{nsys} systems, sides drawn from {list(BATCH_SIDES)} (seed {seed}), "
                    f"{sum(1 for _, u in specs if u == 1)} one-field / "
                    f"{sum(1 for _, u in specs if u == 3)} three-field, mixed precision",
        "grid_points": gp,
        "value": gp * steps / wall,
        "unit": "grid-point-steps/s",
        "ms_per_step": 1e3 * wall / steps,
        "steps": steps,
        "one_by_one": {"value": gp * steps / wall1, "ms_per_step": 1e3 * wall1 / steps},
        "speedup_vs_one_by_one": wall1 / wall,
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": gbs / PEAK_HBM_GBS, "one_by_one_frac": gbs1 / PEAK_HBM_GBS,
                     "algorithmic_bytes_per_step": step_bytes,
                     "note": "every system's row- and column-pass bytes per step (each pass "
                             "reads and writes its fields' spectral buffers once) over the "
                             "wall-clock step time of the whole batch"},
        "autotune_s": tune_s,
        "shapes": shapes,
        "input_generation_s": gen_s,
    }


def other_configs(steps, warmup=3):
    """The other measured configurations at full size, both forms: the 2048^2
    graphene polycrystal (64 seeded grains, conserved, dt = 1) and the 8192^2
    three-field heterostructure (dt = 0.25), with the inputs of their GPU
    tests, and a 1000^2 polycrystal on the generic path, timed like the
    headline (windows starting at a refresh step, one evaluate per 100
    steps).  ms per step and grid-point-steps/s."""
    from pfc_amd import FP64, MIXED, LATTICE_A, graphene_params, hetero_params
    from pfc_amd.inputs import hetero_patchwork, polycrystal
    out = {}
    dx = LATTICE_A / 8.0
    cases = [
        ("graphene polycrystal 2048x2048, 1 field, dt=1",
         lambda: (graphene_params(c=1.0, dt=1.0),
                  polycrystal(2048, 2048, dx, dx, 64, 0.2, -0.2, LATTICE_A, sym=3, noise=0.01,
                              boundary=2, seed=64)[None], 2048, dx, 1)),
        ("hetero 3-field 8192x8192, dt=0.25",
         lambda: (hetero_params(dt=0.25, c=1.0, sigma=2.0 / 7.2552),
                  hetero_patchwork(8192, 8192, 0.7, 0.7, n_domains=128, seed=8192, noise=0.01),
                  8192, 0.7, 3)),
        # a side the tuned kernels do not cover: the generic path (runtime
        # mixed-radix transforms, one launch sequence per step, timed whole)
        ("graphene polycrystal 1000x1000, 1 field, dt=1 (generic path)",
         lambda: (graphene_params(c=1.0, dt=1.0),
                  polycrystal(1000, 1000, dx, dx, 16, 0.2, -0.2, LATTICE_A, sym=3, noise=0.01,
                              boundary=2, seed=1000)[None], 1000, dx, 1)),
    ]
    for name, make in cases:
        params, fields, N, h, U = make()
        generic = "generic path" in name
        row = {}
        for pname, prec in (("fp64", FP64), ("mixed", MIXED)):
            g, wall, ms, n, nev, _ = run_gpu(params, fields, N, N, h, h, prec, steps, warmup)
            esz = 16 if prec == FP64 else 8
            rms, cms = ms[0] / max(n, 1), ms[1] / max(n, 1)
            r = {"ms_per_step": 1e3 * wall / steps,
                 "grid_point_steps_per_s": N * N * steps / wall,
                 "step_model_frac": survey_bytes(N, N, U, esz // 2) * steps / wall / 1e9
                 / PEAK_HBM_GBS}
            if not generic:
                r.update({"row_pass_ms": rms, "column_pass_ms": cms,
                          "row_pass_frac": row_bytes(N, N, U, esz) / (rms * 1e-3) / 1e9
                          / PEAK_HBM_GBS,
                          "column_pass_frac": col_bytes(N, N, U, esz) / (cms * 1e-3) / 1e9
                          / PEAK_HBM_GBS})
            row[pname] = r
            g.close()
        out[name] = row
        del fields
    return out


def traffic_from_profile(Nx, Ny, key):
    tf = os.path.join(ROOT, "profiles", f"traffic_step_{Nx}x{Ny}_u3.txt")
    if os.path.exists(tf):
        for line in open(tf):
            if line.startswith(f"{key}:"):
                return float(line.split(":")[1])
    return None


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--cpu-budget", type=float, default=15.0,
                    help="seconds of CPU-oracle stepping for the baseline (0: skip)")
    ap.add_argument("--accuracy-steps", type=int, default=5000,
                    help="length of the mixed-vs-FP64 comparison run (0: skip)")
    ap.add_argument("--batch-systems", type=int, default=96,
                    help="systems of the batched-relaxation leg (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=100)
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the batched-relaxation leg and print its JSON object")
    ap.add_argument("--other-steps", type=int, default=100,
                    help="steps per form of the 2048^2 and 8192^2 configurations (0: skip)")
    ap.add_argument("--c4-accuracy-steps", type=int, default=0,
                    help="run only the 8192^2 mixed-vs-FP64 comparison over this many steps "
                         "(checkpoints 1000, 5000, 20000) and print its JSON object")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported: the whole "
              "grid lives on one MI355X and this build has no multi-GPU path", file=sys.stderr)
        return 2

    if a.batch_only:
        print(json.dumps(batch_bench(a.batch_systems, a.batch_steps, 10)))
        return 0
    if a.c4_accuracy_steps > 0:
        from pfc_amd import hetero_params
        from pfc_amd.inputs import hetero_patchwork
        p = hetero_params(dt=0.25, c=1.0, sigma=2.0 / 7.2552)
        f = hetero_patchwork(8192, 8192, 0.7, 0.7, n_domains=128, seed=8192, noise=0.01)
        r = accuracy_study(p, f, 8192, 8192, 0.7, 0.7, a.c4_accuracy_steps, (1000, 5000, 20000),
                           atoms=False)
        print(json.dumps({"workload": "hetero 3-field 8192x8192, 128 domains, dt=0.25", **r}))
        return 0

    from pfc_amd import FP64, MIXED, hetero_params
    from pfc_amd.inputs import hetero_patchwork

    Nx = Ny = a.size
    U = 3
    dx = dy = 0.7
    sigma = 2.0 / 7.2552          # real-space Gaussian width ~ a/2
    params = hetero_params(dt=0.25, c=1.0, sigma=sigma)
    t0 = time.perf_counter()
    fields = hetero_patchwork(Nx, Ny, dx, dy, domain=256, seed=2024, noise=0.01)
    gen_s = time.perf_counter() - t0

    res = {}
    for name, prec in (("fp64", FP64), ("mixed", MIXED)):
        g, wall, ms, n, nev, s0 = run_gpu(params, fields, Nx, Ny, dx, dy, prec, a.steps, a.warmup)
        progress(f"{name} windows done: {1e3 * wall / a.steps:.4f} ms per step")
        esz = 16 if prec == FP64 else 8
        col_ms = ms[1] / max(n, 1)
        row_ms = ms[0] / max(n, 1)
        res[name] = dict(
            wall=wall, ms_step=1e3 * wall / a.steps, gps=Nx * Ny * a.steps / wall,
            col_ms=col_ms, row_ms=row_ms, evaluates=nev, window_start=s0,
            col_gbs=col_bytes(Nx, Ny, U, esz) / (col_ms * 1e-3) / 1e9,
            row_gbs=row_bytes(Nx, Ny, U, esz) / (row_ms * 1e-3) / 1e9,
            step_model_gbs=survey_bytes(Nx, Ny, U, esz // 2) * a.steps / wall / 1e9,
        )
        fe, ne = g.evaluate()
        res[name]["fave"] = fe
        res[name]["fields"] = g.get_fields()
        g.close()

    m, d = res["mixed"], res["fp64"]
    err_l2 = float(np.linalg.norm(m["fields"] - d["fields"]) / np.linalg.norm(d["fields"]))
    err_f = abs(m["fave"] - d["fave"]) / abs(d["fave"])
    after = d["window_start"] + a.steps
    del m["fields"], d["fields"]

    acc = None
    if a.accuracy_steps > 0:
        acc = accuracy_study(params, fields, Nx, Ny, dx, dy, a.accuracy_steps,
                             (100, 500, 1000, 2000, 5000, 10000, 20000))

    batch = None
    if a.batch_systems > 0:
        batch = batch_bench(a.batch_systems, a.batch_steps, 10)
        progress("batch done")

    others = other_configs(a.other_steps) if a.other_steps > 0 else None
    if others:
        progress("other configurations done")

    cpu = None
    if a.cpu_budget > 0:
        cpu = cpu_baseline(params, fields, Nx, Ny, dx, dy, a.cpu_budget)
        progress("CPU baseline done")

    def kernel_roof(r, pname, esz):
        dom = "column" if r["col_ms"] >= r["row_ms"] else "row"
        ach = r["col_gbs"] if dom == "column" else r["row_gbs"]
        # the engine's choice of column pass for plain steps (pfc_create):
        # per field for three fields at Ny <= 1024 and Ny >= 4096 (FP64) or
        # 8192 (mixed)
        split = os.environ.get("PFC_COLSPLIT", "") != "0" and (
            os.environ.get("PFC_COLSPLIT", "") == "1" or Ny <= 1024
            or Ny >= (4096 if pname == "fp64" else 8192))
        col_kernel = "k_col_step_f (column pass per field, " if split else "k_col_step (column pass, "
        return {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": ach / PEAK_HBM_GBS,
                "traffic": traffic_from_profile(Nx, Ny, f"{pname}_{dom}_bytes_per_launch"),
                "kernel": (col_kernel if dom == "column" else
                           ("k_row_step_p (row pass, persistent, " if pname == "fp64" else
                            "k_row_step (row pass, ")) + pname + ")",
                "algorithmic_bytes_per_launch": (col_bytes if dom == "column" else row_bytes)(
                    Nx, Ny, U, esz),
                "avg_launch_ms": r["col_ms"] if dom == "column" else r["row_ms"],
                "row_pass_frac": r["row_gbs"] / PEAK_HBM_GBS,
                "column_pass_frac": r["col_gbs"] / PEAK_HBM_GBS,
                "step_model_frac": r["step_model_gbs"] / PEAK_HBM_GBS}

    roof = kernel_roof(d, "fp64", 16)
    roof["step_model_note"] = ("step_model_frac: the three-sweep model's bytes per step, "
                               "U (16 NxNy + 96 (Nx/2+1) Ny) B in FP64, times the steps, over "
                               "the whole timed wall clock (evaluates and refreshes included)")
    out = {
        "metric": "grid-point-steps/s at 4096² 3-field, mixed vs FP64 on 1 GPU; HBM GB/s",
        "value": d["gps"],
        "unit": "grid-point-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": d["ms_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded graphene/hBN Voronoi patchwork (hetero.in parameters)",
        "config": {"workload": f"hetero 3-field PFC {Nx}x{Ny} conserved dt=0.25, FP64 (the "
                               f"reference's precision); the mixed FP32/FP64 form beside it",
                   "grid": [Nx, Ny], "fields": U, "parallelism": "1 GPU",
                   "timed_window": f"steps {d['window_start']}..{d['window_start'] + a.steps - 1} "
                                   f"with {d['evaluates']} evaluate(s) (one before each step t "
                                   f"with t % {TPRINT} == 0, as relax() prints) and the t % 100 "
                                   f"refresh steps"
                                   + (f"; a window of {a.steps} < {TPRINT} steps still holds one "
                                      f"of each, {TPRINT / a.steps:.0f}x relax()'s rate"
                                      if a.steps < TPRINT else "")},
        "fp64": {"value": d["gps"], "ms_per_step": d["ms_step"],
                 "column_pass_GBps": d["col_gbs"], "row_pass_GBps": d["row_gbs"],
                 "column_pass_ms": d["col_ms"], "row_pass_ms": d["row_ms"],
                 "step_model_GBps": d["step_model_gbs"]},
        "mixed": {"value": m["gps"], "ms_per_step": m["ms_step"],
                  "column_pass_GBps": m["col_gbs"], "row_pass_GBps": m["row_gbs"],
                  "column_pass_ms": m["col_ms"], "row_pass_ms": m["row_ms"],
                  "step_model_GBps": m["step_model_gbs"],
                  "roofline": kernel_roof(m, "mixed", 8),
                  "accuracy_note": "energy per area against FP64 is within 1e-6 over the first "
                                   "~100 steps of this random start only; mixed_vs_fp64_run sets "
                                   "it beside an FP64 run perturbed every step by the mixed "
                                   "form's one-step error"},
        "speedup_mixed_vs_fp64": m["gps"] / d["gps"],
        "mixed_vs_fp64_error": {"density_rel_l2": err_l2, "energy_per_area_rel": err_f,
                                "after_steps": after},
        "mixed_vs_fp64_run": acc,
        "roofline": roof,
        "cpu_baseline": cpu,
        "batch": batch,
        "other_configs": others,
        "input_generation_s": gen_s,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
