"""Benchmark of the DEM time step on one MI355X.

Workload (the configuration the metric is quoted on): 50,000 convex
sphero-polyhedra -- each the convex hull of 8-14 random points on an
ellipsoid 1 : 0.8 : 0.6, size uniform in [4, 6] mm, Minkowski radius 5 % of
the size -- poured into a box 40 x 40 x 200 mean sizes under gravity,
linkCells neighbour list, ForceLaw Avalanche (kn = 1e7, kt = 0.8 kn, en = 0.3,
mu = 0.6), dt = 2e-7 s, FP64.  The pour is synthetic: bodies packed row by
row from the floor (compact, principal axes along the box axes up to 3 deg),
random velocities in [-0.5, 0.5] m/s, seed 2.

A "step" is one iteration of Rockable's integrate() loop: velocity-Verlet
predictor, contact kinematics + force law, ordered per-body resultants,
accelerations, kick, and the neighbour-list rebuild every interVerlet = 100
steps.  value = bodies x steps / wall seconds of the timed steps (state
resident in HBM).  The roofline object is the force phase (kinematics + law +
resultants), the hot kernels of the step.  cpu_baseline times the CPU oracle
(the restatement of Rockable's OpenMP DefaultParallelMode path) on a bounded
sample of the same workload.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def make_deck(n_bodies, seed, dirpath):
    from rockable_amd import decks as D
    shapes, deck = D.compact_pour_deck(n_bodies=n_bodies, seed=seed, tilt_deg=3.0, gap_frac=0.0, vmax=0.5,
                                       law="Avalanche", interverlet_steps=100)
    return D.write_deck(dirpath, shapes, deck), deck


def force_phase_bytes(nb, nl, law):
    """Algorithmic HBM bytes of one force phase: each body's kinematics read
    once (pos 24 + Q 32 + vel 24 + vrot 24 + group 4 = 108 B) and its force
    and moment written (48 B); per listed interaction its key (12 B), damping
    (8 B) and history read and written (ft, mom: 96 B; Avalanche adds fn and
    dn: 32 B)."""
    per_inter = {"Default": 116, "Avalanche": 148, "StickedLinks": 184}.get(law, 116)
    return 156 * nb + per_inter * nl


def cpu_baseline(deck_path, target_seconds, threads):
    from oracle.rk_oracle import Oracle
    o = Oracle(deck_path, threads=threads)
    o.update_nl()
    o.run_steps(20)
    steps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < target_seconds:
        o.run_steps(20)
        steps += 20
    dt = time.perf_counter() - t0
    nb = o.nbodies
    return {"value": nb * steps / dt, "unit": "body-steps/s", "cores": threads, "kind": "port",
            "sample": "CPU oracle (OpenMP restatement of Rockable's DefaultParallelMode step), same deck, "
                      "%d steps after 20 warm-up steps, %.1f s" % (steps, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=3000)
    ap.add_argument("--bodies", type=int, default=50000)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "16")))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--deck-dir", default=None)
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported (the path would shard by "
              "spatial slabs or by packing, see DESIGN.md)" % a.gpus, file=sys.stderr)
        sys.exit(2)

    from rockable_amd.engine import Engine
    d = a.deck_dir or tempfile.mkdtemp(prefix="rk_bench_")
    t0 = time.perf_counter()
    path, deck = make_deck(a.bodies, a.seed, d)
    print("deck: %d bodies written in %.1f s" % (len(deck.bodies), time.perf_counter() - t0), file=sys.stderr,
          flush=True)

    e = Engine(path, device=0)
    e.set_option("record_geometry", 0)
    e.update_nl()
    t0 = time.perf_counter()
    e.step(a.warmup)
    e.synchronize()
    print("warm-up: %d steps in %.2f s" % (a.warmup, time.perf_counter() - t0), file=sys.stderr, flush=True)

    # timed region: K steps, device state resident, profile events on the stream
    e.set_option("profile", 1)
    e.reset_kernel_time()
    c0 = e.counts()
    e.synchronize()
    t0 = time.perf_counter()
    e.step(a.steps)
    e.synchronize()
    wall = time.perf_counter() - t0
    c1 = e.counts()
    nb = c1["nbodies"]
    kt = {k: e.kernel_time(k) for k in ("predict", "interact", "gather", "breakage", "nl")}
    value = nb * a.steps / wall

    # roofline of the force phase (interact + gather), per launch
    n_int, ms_int = kt["interact"]
    n_gat, ms_gat = kt["gather"]
    nl_listed = 0.5 * (c0["nlisted"] + c1["nlisted"])
    bytes_per = force_phase_bytes(nb, nl_listed, "Avalanche")
    t_force = (ms_int / max(n_int, 1) + ms_gat / max(n_gat, 1)) * 1e-3
    achieved = bytes_per / t_force / 1e9

    out = {
        "metric": "body-steps/s", "value": value, "unit": "body-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": 1e3 * wall / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "plain GPU form: %d convex sphero-polyhedra (hulls of 8-14 points), compact pour, "
                               "linkCells, ForceLaw Avalanche, dt 2e-7 s" % nb,
                   "bodies": nb, "listed_per_body": c1["nlisted"] / nb, "active_per_body": c1["nactive"] / nb,
                   "nl_rebuilds_in_timed_region": c1["nl_updates"] - c0["nl_updates"],
                   "kernel_ms_per_step": {k: (v[1] / max(v[0], 1)) for k, v in kt.items() if v[0]},
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": None,
                     "kernel": "force phase (k_interact + k_gather)", "algorithmic_bytes_per_launch": bytes_per},
    }
    if not a.no_cpu:
        e.close()
        out["cpu_baseline"] = cpu_baseline(path, a.cpu_seconds, a.cpu_threads)
        try:
            out["cpu_baseline"]["cpu_model"] = subprocess.run(
                "lscpu | grep 'Model name' | head -1", shell=True, capture_output=True, text=True).stdout.split(
                ":")[-1].strip()
        except Exception:
            pass
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
