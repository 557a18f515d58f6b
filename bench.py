"""Benchmark of the DEM time step on one MI355X.

Workload (the configuration the metric is quoted on): 50,000 convex
sphero-polyhedra -- each the convex hull of 8-14 random points on an
ellipsoid 1 : 0.8 : 0.6 (slivers under a quarter of the ellipsoid's volume
drawn again), size uniform in [4, 6] mm, Minkowski radius 5 % of the size --
poured into a box 40 x 40 x 200 mean sizes under gravity, linkCells
neighbour list, ForceLaw Avalanche (kn = 1e7, kt = 0.8 kn, en = 0.3, mu =
0.6), dt = 2e-7 s, FP64.

The pour is synthetic and settled by the engine itself before anything is
timed: bodies packed row by row from the floor (principal axes along the box
axes up to 3 deg, velocities in [-0.5, 0.5] m/s, seed 2), settled for
`--settle` steps under 400 g with the same contact law, dt and stiffness
(interVerlet 20 steps), then gravity is lowered to 100, 25, 6.25, 1.5625 and
1 g, `--relax` steps each, so that the elastic spring-back of each stage
damps out instead of throwing the packing up; bodies that left the box are
dropped; the timed run uses interVerlet = 100 steps.  (20,000 steps at
dt = 2e-7 s are only 4 ms: a pour that starts loose stays a dilute gas,
hence the settling.)

A "step" is one iteration of Rockable's integrate() loop: velocity-Verlet
predictor, contact kinematics + force law, ordered per-body resultants,
accelerations, kick, and the neighbour-list rebuild every interVerlet = 100
steps.  value = bodies x steps / wall seconds of the timed steps (state
resident in HBM).  The roofline object is the force phase (kinematics + law +
resultants), the hot kernels of the step; kernel durations come from HIP
events recorded on the engine stream inside the timed region.  cpu_baseline
times the CPU oracle (the restatement of Rockable's OpenMP
DefaultParallelMode path) on a bounded sample of the same settled packing.

Beside the plain form, the JSON line carries "forms": "batched" (256
independent sphere packings in one engine, autotuned launch shapes and
graph-captured steps, with a sample of the same packings run one engine
after another) and "avalanche" (a 1,000,000-body polyhedral avalanche with
the displacement-check rebuild: the settled packing tiled and released on a
slope; the plain GPU form, fixed batches and the batched + learned-rebuild
form, timed on the same steps of the same trajectory) and
"bonds" (200,000 sphero-cubes in 2,000 bonded clusters hitting a plane, bonds
breaking on impact).  --forms "" leaves them out.
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
# HBM bytes per force-phase launch on the 50 005-body settled pour:
# rocprofv3 FETCH_SIZE and WRITE_SIZE, each scaled by what it reads for a
# plain 256 MiB device copy (factors 2.0 and 1.0).  With the far-entry skip
# (k_interact_seg + k_gather + k_heavy_finish): 137.0 MB read + 27.9 MB
# written, profiles/c2_force_phase_traffic.md; without it (k_interact +
# k_gather + k_heavy_finish, measured before the skip): 171.3 + 36.3 MB.
TRAFFIC_BYTES_PER_LAUNCH = {1: 164.9e6, 0: 207.6e6}


def make_settled_deck(n_bodies, seed, dirpath, settle_steps, relax_steps, log):
    """Loose compact pour -> settled under 400 g by the engine -> gravity
    lowered to 1 g in steps of 4x (each stage long enough for the spring-back
    of the contacts to damp out) -> bodies that left the box dropped -> conf
    with the workload's own parameters (see the module docstring)."""
    import numpy as np
    from rockable_amd import decks as D
    from rockable_amd.engine import Engine
    dt = 2e-7
    g = 9.81
    shapes, deck = D.compact_pour_deck(n_bodies=n_bodies, seed=seed, tilt_deg=3.0, gap_frac=0.0, vmax=0.5,
                                       law="Avalanche", dt=dt, interverlet_steps=20)
    deck.gravity = (0.0, 0.0, -400.0 * g)
    loose = D.write_deck(dirpath, shapes, deck, name="loose.conf")
    e = Engine(loose, device=0)
    e.set_option("record_geometry", 0)
    e.update_nl()
    t0 = time.perf_counter()
    stages = [(400.0, settle_steps)] + [(m, relax_steps) for m in (100.0, 25.0, 6.25, 1.5625, 1.0)]
    for mult, steps in stages:
        e.set_option("gravity_z", -mult * g)
        done = 0
        while done < steps:
            n = min(5000, steps - done)
            e.step(n)
            e.synchronize()
            done += n
        c = e.counts()
        v = np.linalg.norm(e.bodies()["vel"][deck.nDriven:], axis=1)
        log("settle at %g g: %d steps, %.1f s, listed/body %.2f, active/body %.2f, speed mean %.3g max %.3g m/s" % (
            mult, steps, time.perf_counter() - t0, c["nlisted"] / c["nbodies"], c["nactive"] / c["nbodies"],
            v.mean(), v.max()))
    raw = os.path.join(dirpath, "settled_raw.conf")
    e.save_conf(raw)
    b = e.bodies()
    e.close()
    # keep the driven walls and every body whose centre is inside the box
    lo = np.zeros(3)
    hi = np.array([40.0, 40.0, 200.0]) * 5e-3
    inside = np.all((b["pos"] >= lo) & (b["pos"] <= hi), axis=1)
    inside[:deck.nDriven] = True
    kept = D.conf_drop_bodies(raw, os.path.join(dirpath, "settled_kept.conf"), inside)
    path = D.conf_with(kept, os.path.join(dirpath, "settled.conf"), gravity=(0.0, 0.0, -g), t=0.0, tmax=1.0,
                       interConf=1.0, interVerlet=100 * dt)
    return path, int((~inside).sum())


# options of every timed engine: no geometry records (the step writes the
# contact geometry only when an output is due), and the far-entry skip when
# asked for (--skip-far; otherwise the engine's default)
SKIP_FAR = None


def timed_options(e):
    e.set_option("record_geometry", 0)
    if SKIP_FAR is not None:
        e.set_option("skip_far", SKIP_FAR)


def form_autotune(e, a):
    """The forms' launch shapes: autotuned, or with --no-autotune the
    engine's defaults (no captured graphs), as a profiled run takes them."""
    if a.no_autotune:
        return 0, {k: int(e.get_option(k)) for k in e.TUNED}
    return e.autotune(steps_per_trial=a.tune_trial)


def interverlet_steps(e):
    """Steps between two scheduled rebuilds of the engine's deck
    (interVerlet / dt, the loop's own test at Rockable.cpp:2693)."""
    import math
    c = e.counts()
    iv = e.conf_scalar("interVerlet")
    return max(1, int(math.ceil(iv / c["dt"] - 0.5)))


def form_roofline(e, steps, law):
    """Force-phase roofline of a form, on the engine in its timed state:
    steps more steps with HIP events bracketing the force phase on the
    engine stream, the algorithmic bytes of each launch over its average
    duration."""
    c0 = e.counts()
    e.set_option("profile", 1)
    e.reset_kernel_time()
    e.step(steps)
    e.synchronize()
    n, ms = e.kernel_time("force")
    e.set_option("profile", 0)
    c1 = e.counts()
    nl_listed = 0.5 * (c0["nlisted"] + c1["nlisted"])
    b = force_phase_bytes(c1["nbodies"], nl_listed, law)
    t = ms / max(n, 1) * 1e-3
    ach = b / t / 1e9 if t > 0 else 0.0
    out = {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": ach / PEAK_HBM_GBS,
           "traffic": None, "kernel": "force phase (%s law)" % law, "algorithmic_bytes_per_launch": b,
           "force_phase_ms": 1e3 * t, "launches": n, "listed_per_body": nl_listed / c1["nbodies"],
           "note": "HIP events over %d steps right after the timed region, same state" % steps}
    sk = skip_accounting(e, min(steps, 200), c1["nbodies"], nl_listed, law, t)
    if sk is not None:
        out["far_entry_skip"] = sk
    return out


def skip_accounting(e, steps, nb, nl, law, t_force):
    """With the far-entry skip on, the share of listed entries the force
    phase leaves alone (counted over `steps` further steps, outside every
    timed region) and the bytes a launch then touches: 17 B per entry tested
    (activity byte, key, pair), the full per-entry figure for the evaluated
    ones, the bodies' 156 B.  The roofline's `achieved` keeps the per-listed-
    interaction figure of the reference's algorithm, which evaluates every
    listed entry; this is the skip's own count beside it."""
    if not int(e.get_option("skip_far")):
        return None
    e.set_option("count_skips", 1)
    e.step(steps)
    skipped, looked = e.get_option("skipped_entries"), e.get_option("entries_looked_at")
    e.set_option("count_skips", 0)
    f = skipped / looked if looked > 0 else 0.0
    per_inter = {"Default": 116, "Avalanche": 148, "StickedLinks": 184}.get(law, 116)
    touched = 156 * nb + 17 * nl + per_inter * (1.0 - f) * nl
    ach = touched / t_force / 1e9 if t_force > 0 else 0.0
    return {"skipped_fraction": f, "steps_counted": steps, "bytes_touched_per_launch": touched,
            "achieved_touched": ach, "frac_touched": ach / PEAK_HBM_GBS}


def force_phase_bytes(nb, nl, law):
    """Algorithmic HBM bytes of one force phase: each body's kinematics read
    once (pos 24 + Q 32 + vel 24 + vrot 24 + group 4 = 108 B) and its force
    and moment written (48 B); per listed interaction its key (12 B), damping
    (8 B) and history read and written (ft, mom: 96 B; Avalanche adds fn and
    dn: 32 B)."""
    per_inter = {"Default": 116, "Avalanche": 148, "StickedLinks": 184}.get(law, 116)
    return 156 * nb + per_inter * nl


def _oracle_rate(deck_path, target_seconds, threads):
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
    return o.nbodies * steps / dt, steps, dt


def lscpu_summary():
    """CPU model and core counts as lscpu reports them."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    model = info.get("Model name")
    if not model:
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
    return {"cpu_model": model or "unknown", "sockets": info.get("Socket(s)"),
            "cores_per_socket": info.get("Core(s) per socket"), "threads_per_core": info.get("Thread(s) per core"),
            "logical_cpus": info.get("CPU(s)")}


def _oracle_rate_child(deck_path, seconds, threads):
    """_oracle_rate in a child process whose OpenMP runtime starts with
    threads threads bound one per core (OMP_PROC_BIND=close,
    OMP_PLACES=cores): the settings must be in the environment before the
    runtime starts, and the parent's runtime has started already."""
    import subprocess
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-child", deck_path, str(seconds),
                        str(threads)], capture_output=True, text=True, timeout=600, env=env)
    if r.returncode != 0:
        raise RuntimeError(r.stderr.strip()[-300:])
    return json.loads(r.stdout.strip().splitlines()[-1])


def cpu_baseline(deck_path, target_seconds, threads=None, repeats=3):
    """The CPU oracle on the cores this process may run on (one OpenMP thread
    per core, bound), repeats runs of target_seconds / repeats each, the
    median reported; and one 1-thread run for scale."""
    import statistics
    if threads is None:
        threads = len(os.sched_getaffinity(0))
    per = max(2.0, target_seconds / repeats)
    runs = [_oracle_rate_child(deck_path, per, threads) for _ in range(repeats)]
    one = _oracle_rate_child(deck_path, max(2.0, per), 1)
    med = statistics.median(r["value"] for r in runs)
    cpu = lscpu_summary()
    out = {"value": med, "unit": "body-steps/s", "cores": threads, "kind": "port", "cpu_model": cpu["cpu_model"],
           "lscpu": cpu, "affinity_cpus": len(os.sched_getaffinity(0)),
           "runs": [r["value"] for r in runs], "value_1_thread": one["value"],
           "sample": "CPU oracle (OpenMP restatement of Rockable's DefaultParallelMode step; %d threads, "
                     "OMP_PROC_BIND=close, OMP_PLACES=cores), the same settled packing, median of %d runs of "
                     "%d-%d steps after 20 warm-up steps (%.1f s each); 1 thread: %d steps, %.1f s"
                     % (threads, repeats, min(r["steps"] for r in runs), max(r["steps"] for r in runs), per,
                        one["steps"], one["seconds"])}
    return out


def copy_bandwidth_child(mib=1024):
    """copy_bandwidth in a child process: torch brings its own HIP runtime,
    which must not share a process with the engine's."""
    import subprocess
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--copy-rate", str(mib)], capture_output=True,
                       text=True, timeout=300)
    if r.returncode != 0:
        raise RuntimeError(r.stderr.strip()[-300:])
    return float(json.loads(r.stdout.strip().splitlines()[-1])["copy_GBs"])


def copy_bandwidth(mib=1024, reps=10):
    """Device-to-device copy rate (read + write bytes per second), the
    STREAM-copy figure the roofline is also quoted against."""
    import torch
    n = mib << 20
    src = torch.empty(n, dtype=torch.uint8, device="cuda:0")
    dst = torch.empty_like(src)
    src.fill_(1)
    dst.copy_(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    rate = 2.0 * n * reps / (time.perf_counter() - t0) / 1e9
    del src, dst
    torch.cuda.empty_cache()
    return rate


def batched_form(a, log):
    """Batched form: a.packings independent packings of uneven sizes (body
    counts log-uniform in [500, 20000]) of the CPU-runnable configuration's
    spheres (R = 0.5 mm on a 2.2 R lattice dropping into a floor and four
    walls, velocities in [-0.01, 0.01] m/s), each with its own Default-law mu
    and en, all in one engine: one launch per kernel and step covers every
    packing.  The launch shapes and the graph length are picked by the
    autotuner on a first load of the batch, then a fresh load runs the
    configuration's a.batched_steps steps from the start, timed (so that the
    autotuned and the default shapes time the same work).  Beside it, a
    sample of the same packings run one engine after another (the unbatched
    way)."""
    import numpy as np
    from rockable_amd import decks as D
    from rockable_amd.engine import Engine
    d = tempfile.mkdtemp(prefix="rk_batch_")
    t0 = time.perf_counter()
    shapes, deck, info = D.batched_sphere_packings(n_packings=a.packings, steps=a.batched_steps)
    path = D.write_deck(d, shapes, deck)
    log("batched deck: %d packings, %d bodies, written in %.1f s" % (len(info), sum(p["bodies"] for p in info),
                                                                     time.perf_counter() - t0))
    e = Engine(path, device=0)
    timed_options(e)
    e.update_nl()
    ntune, choice = form_autotune(e, a)
    log("autotune: %d steps, %s" % (ntune, choice))
    e.close()
    t_load = time.perf_counter()
    e = Engine(path, device=0)
    load_s = time.perf_counter() - t_load
    timed_options(e)
    for k, v in choice.items():
        e.set_option(k, v)
    e.update_nl()
    c0 = e.counts()
    e.synchronize()
    t0 = time.perf_counter()
    e.step(a.batched_steps)
    e.synchronize()
    wall = time.perf_counter() - t0
    c1 = e.counts()
    nb = c1["nbodies"]
    roof = form_roofline(e, min(500, a.batched_steps), "Default")
    e.close()
    # unbatched sample: every 32nd packing as an engine of its own, default shapes
    rng = np.random.default_rng(0)
    sample = sorted(rng.choice(len(info), size=min(a.unbatched_sample, len(info)), replace=False).tolist())
    work, swall = 0, 0.0
    for p in sample:
        s2, d2 = D.single_packing_of(shapes, deck, info, p)
        sp = D.write_deck(os.path.join(d, "p%d" % p), s2, d2)
        q = Engine(sp, device=0)
        timed_options(q)
        q.update_nl()
        q.step(100)
        q.synchronize()
        t1 = time.perf_counter()
        q.step(a.unbatched_steps)
        q.synchronize()
        swall += time.perf_counter() - t1
        work += q.counts()["nbodies"] * a.unbatched_steps
        q.close()
    return {"value": nb * a.batched_steps / wall, "unit": "body-steps/s", "ms_per_step": 1e3 * wall / a.batched_steps,
            "steps": a.batched_steps, "bodies": nb, "packings": len(info),
            "workload": "%d packings of 500-20000 spheres (R 0.5 mm, 2.2 R lattice, floor + 4 walls each), "
                        "per-packing Default-law mu in [0.2, 0.8] and en in [0.1, 0.9], linkCells, "
                        "dt 1e-6 s, NL every 100 steps" % len(info),
            "autotuned": choice, "autotune_steps": ntune, "roofline": roof, "deck_load_s": load_s,
            "nl_rebuilds": c1["nl_updates"] - c0["nl_updates"],
            "graph_launches": c1["graph_launches"] - c0["graph_launches"],
            "unbatched_sample": {"packings": sample, "value": work / swall, "unit": "body-steps/s",
                                 "steps_each": a.unbatched_steps,
                                 "note": "the same packings, one engine after another, default launch shapes"}}


def bonds_form(a, log):
    """Breakable-bonds workload: a.bond_clusters clusters of 5 x 5 x 4
    sphero-cubes (5 mm, Minkowski radius 0.25 mm) bonded face to face at
    their coincident vertices (one StickedLinks interface per touching pair,
    fn0 and ft0 lognormal with mean 50 N and cv 0.3, ParamsInInterfaces 1),
    0.1 mm above a rigid plane inclined at 35 deg and falling onto it at
    2 m/s, dt 1e-7 s, linkCells.  The launch shapes are autotuned on a first
    run of the deck (through its impact); the deck is then loaded again and
    its first a.bond_steps steps are timed: the fall and the impact, with the
    bonds breaking (a step that breaks one freezes the later steps of its
    batch on the device; the host accounts the broken interfaces and
    rebuilds the list after that step, as the reference does)."""
    from rockable_amd import decks as D
    from rockable_amd.engine import Engine
    d = tempfile.mkdtemp(prefix="rk_bonds_")
    t0 = time.perf_counter()
    shapes, deck = D.bonded_clusters_deck(n_clusters=a.bond_clusters, clearance=1e-4)
    path = D.write_deck(d, shapes, deck)
    log("bonded-cluster deck: %d clusters, written in %.1f s" % (a.bond_clusters, time.perf_counter() - t0))
    e = Engine(path, device=0)
    timed_options(e)
    e.update_nl()
    ntune, choice = form_autotune(e, a)
    log("autotune: %d steps, %s" % (ntune, choice))
    e.close()
    t_load = time.perf_counter()
    e = Engine(path, device=0)
    load_s = time.perf_counter() - t_load
    timed_options(e)
    for k, v in choice.items():
        e.set_option(k, v)
    e.update_nl()
    c0 = e.counts()
    e.synchronize()
    t1 = time.perf_counter()
    e.step(a.bond_steps)
    e.synchronize()
    wall = time.perf_counter() - t1
    c1 = e.counts()
    st = e.stats()
    nb = c1["nbodies"]
    roof = form_roofline(e, 200, "StickedLinks")
    e.close()
    return {"value": nb * a.bond_steps / wall, "unit": "body-steps/s", "ms_per_step": 1e3 * wall / a.bond_steps,
            "steps": a.bond_steps, "bodies": nb, "autotuned": choice, "autotune_steps": ntune, "roofline": roof,
            "deck_load_s": load_s,
            "interfaces_before": c0["ninterfaces"], "interfaces_after": c1["ninterfaces"],
            "broken_area": c1["total_broken_area"], "nl_rebuilds": c1["nl_updates"] - c0["nl_updates"],
            "active_per_body": c1["nactive"] / nb, "kinetic_energy": st["kinetic_translational"] +
            st["kinetic_rotational"],
            "workload": "%d clusters of 5x5x4 bonded sphero-cubes (%d bodies) falling at 2 m/s onto a 35 deg "
                        "plane, StickedLinks, dt 1e-7 s; autotuned on a first run (%d steps), then the first %d "
                        "steps of a fresh load timed: the fall and the impact" % (a.bond_clusters, nb, ntune,
                                                                                a.bond_steps)}


def avalanche_forms(a, log, settled_path):
    """Learned-rebuild workload: the settled plain-form packing (50 005
    bodies: the hull polyhedra poured and settled in their 40 x 40-size box)
    tiled along x into a bed of about a.avalanche_bodies bodies and released
    on a 30 deg slope (gravity tilted; the box's downslope wall moved 0.5 m
    past the bed's end; 1 m/s added downslope to every body), dt 2e-7 s,
    ForceLaw Avalanche, the neighbour list rebuilt by the displacement check
    alone (dynamicUpdateNL, dispUpdateNL = DVerlet / 4, angleUpdateNL = 1 rad,
    the reference default).  The launch shapes are autotuned on the running
    avalanche, a.avalanche_warmup more steps are run, and that state is saved;
    then three forms each load it and run the same a.avalanche_steps steps:
    "plain" (the check read by the host after every step, as the reference
    runs it, every kernel launched on its own), "fixed_batches" (the engine's
    default: batches of 16 steps between host reads of the check, captured
    graphs when the autotuner keeps them) and "batched_learned" (batches sized
    by the rebuild-interval network, asynchronous polls of the check).  The
    three final states must be identical."""
    import math
    import numpy as np
    from rockable_amd import decks as D
    from rockable_amd.engine import Engine, Mlp
    d = tempfile.mkdtemp(prefix="rk_aval_")
    t0 = time.perf_counter()
    per_tile = D.read_conf_bodies(settled_path)[1]["names"].size
    nd0 = 5
    tiles = max(1, int(math.ceil(a.avalanche_bodies / float(per_tile - nd0))))
    pool = D.hull_shape_pool(64, a.seed)
    path = D.slope_from_settled(settled_path, pool, d, tiles=tiles, slope_deg=30.0, flow_speed=1.0, runout=0.5,
                                tile_gap=2e-3)
    log("avalanche deck: %d tiles of the settled packing, written in %.1f s" % (tiles, time.perf_counter() - t0))
    mlp = Mlp(device=0)
    mlp.set_params(np.load(os.path.join(ROOT, "rockable_amd", "data", "rebuild_mlp.npy")))
    e = Engine(path, device=0)
    timed_options(e)
    e.update_nl()
    ntune, choice = form_autotune(e, a)
    e.step(a.avalanche_warmup)
    start = ntune + a.avalanche_warmup
    log("autotune: %d steps, %s" % (ntune, choice))
    start_conf = os.path.join(d, "start.conf")
    e.save_conf(start_conf)
    e.close()
    forms, finals = {}, {}
    for name in ("plain", "fixed_batches", "batched_learned"):
        e = Engine(start_conf, device=0)
        timed_options(e)
        for k, v in choice.items():
            e.set_option(k, v)
        if name == "plain":
            e.set_option("guard_batch", 1)
            e.set_option("graph_steps", 0)
        elif name == "fixed_batches":
            e.set_option("guard_batch", 16)
        else:
            e.set_rebuild_model(mlp)
        e.update_nl()
        if name == "plain":
            st0 = e.stats()
        c0 = e.counts()
        e.synchronize()
        t1 = time.perf_counter()
        e.step(a.avalanche_steps)
        e.synchronize()
        wall = time.perf_counter() - t1
        c1 = e.counts()
        forms[name] = {"value": c1["nbodies"] * a.avalanche_steps / wall, "unit": "body-steps/s",
                       "ms_per_step": 1e3 * wall / a.avalanche_steps, "steps": a.avalanche_steps,
                       "nl_rebuilds": c1["nl_updates"] - c0["nl_updates"],
                       "guard_reads": c1["guard_reads"] - c0["guard_reads"],
                       "frozen_steps": c1["guard_wasted"] - c0["guard_wasted"],
                       "guard_polls": c1["guard_polls"] - c0["guard_polls"],
                       "graph_launches": c1["graph_launches"] - c0["graph_launches"]}
        b = e.bodies()
        finals[name] = {k: b[k] for k in ("pos", "vel", "Q", "vrot")}
        if name == "plain":
            stats, c = e.stats(), c1
            roof = form_roofline(e, 200, "Avalanche")
        e.close()
        log("avalanche %s: %.3f ms/step" % (name, forms[name]["ms_per_step"]))
    mlp.close()
    nb = c["nbodies"]
    out = dict(forms)
    out["batched_learned_over_plain"] = forms["batched_learned"]["value"] / forms["plain"]["value"]
    out["batched_learned_over_fixed_batches"] = forms["batched_learned"]["value"] / forms["fixed_batches"]["value"]
    out["identical_final_state"] = all(np.array_equal(finals["plain"][k], finals[f][k])
                                       for f in ("fixed_batches", "batched_learned") for k in finals["plain"])
    out["bodies"] = nb
    out["listed_per_body"] = c["nlisted"] / nb
    out["active_per_body"] = c["nactive"] / nb
    out["coordination_at_start"] = st0["coordination"]
    out["coordination_at_end"] = stats["coordination"]
    out["kinetic_energy"] = stats["kinetic_translational"] + stats["kinetic_rotational"]
    out["autotuned"] = choice
    out["roofline"] = roof
    out["workload"] = ("%d hull polyhedra: the settled plain-form packing tiled %d times and released down a "
                       "30 deg slope at 1 m/s, dt 2e-7 s, dynamicUpdateNL (disp DVerlet/4, 1 rad); %d steps "
                       "(autotuning and warm-up) untimed, then the same %d steps timed in each form from the "
                       "saved state" % (nb, tiles, start, a.avalanche_steps))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=500)
    ap.add_argument("--bodies", type=int, default=50000)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--settle", type=int, default=45000)
    ap.add_argument("--relax", type=int, default=4000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=None,
                    help="CPU baseline threads (default: the cores this process may run on)")
    ap.add_argument("--cpu-child", nargs=3, default=None, metavar=("DECK", "SECONDS", "THREADS"),
                    help=argparse.SUPPRESS)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--deck-dir", default=None)
    ap.add_argument("--packings", type=int, default=256)
    ap.add_argument("--batched-steps", type=int, default=5000)
    ap.add_argument("--unbatched-sample", type=int, default=8)
    ap.add_argument("--unbatched-steps", type=int, default=1000)
    ap.add_argument("--avalanche-bodies", type=int, default=1000000)
    ap.add_argument("--avalanche-steps", type=int, default=3000, help="timed steps per form")
    ap.add_argument("--avalanche-warmup", type=int, default=2000)
    ap.add_argument("--tune-trial", type=int, default=50)
    ap.add_argument("--bond-clusters", type=int, default=2000)
    ap.add_argument("--bond-steps", type=int, default=2000, help="timed steps of the bonds form")
    ap.add_argument("--forms", default="batched,avalanche,bonds",
                    help="extra forms reported beside the plain one (comma list; empty for none)")
    ap.add_argument("--forms-only", action="store_true",
                    help="run the --forms alone, without the headline (profiled runs of one form)")
    ap.add_argument("--prepare-only", action="store_true",
                    help="settle the packing into --deck-dir and stop (later runs reuse it)")
    ap.add_argument("--copy-rate", type=int, default=0, metavar="MIB",
                    help="only print the device copy rate of MIB MiB buffers (run as a child process)")
    ap.add_argument("--no-copy-rate", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--set", default="", metavar="K=V,...",
                    help="launch options of the headline engine, in place of autotuning")
    ap.add_argument("--skip-far", type=int, default=None, choices=(0, 1),
                    help="the engine's skip_far option on every timed engine (default: the engine's default)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device 3 times (calibrates the HBM byte counters)")
    a = ap.parse_args()
    if a.cpu_child:
        v, steps, sec = _oracle_rate(a.cpu_child[0], float(a.cpu_child[1]), int(a.cpu_child[2]))
        print(json.dumps({"value": v, "steps": steps, "seconds": sec}), flush=True)
        return
    if a.copy_rate > 0:
        print(json.dumps({"copy_GBs": copy_bandwidth(a.copy_rate)}), flush=True)
        return
    if a.copy_calibration > 0:
        import torch
        n = a.copy_calibration << 20
        src = torch.empty(n, dtype=torch.uint8, device="cuda:0")
        dst = torch.empty_like(src)
        src.fill_(1)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n, "copies": 3}), flush=True)
        return
    global SKIP_FAR
    SKIP_FAR = a.skip_far
    if a.gpus != 1:
        print("bench.py: the engine runs on one GPU; --gpus %d is not supported" % a.gpus, file=sys.stderr)
        sys.exit(2)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    from rockable_amd.engine import Engine
    d = a.deck_dir or tempfile.mkdtemp(prefix="rk_bench_")
    os.makedirs(d, exist_ok=True)
    meta_path = os.path.join(d, "settled.json")
    meta = {"bodies": a.bodies, "seed": a.seed, "settle": a.settle, "relax": a.relax}
    old = json.load(open(meta_path)) if os.path.exists(meta_path) else None
    if old is not None and all(old.get(k) == v for k, v in meta.items()):
        path, dropped = os.path.join(d, "settled.conf"), old["dropped"]
        log("reusing the settled packing in %s" % d)
    else:
        path, dropped = make_settled_deck(a.bodies, a.seed, d, a.settle, a.relax, log)
        meta["dropped"] = dropped
        json.dump(meta, open(meta_path, "w"))
    if a.prepare_only:
        log("settled packing in %s" % d)
        return
    if a.forms_only:
        # the forms alone (a profiled run of one form); no headline line
        forms = [f for f in a.forms.split(",") if f]
        res = {}
        if "batched" in forms:
            res["batched"] = batched_form(a, log)
        if "avalanche" in forms:
            res["avalanche"] = avalanche_forms(a, log, path)
        if "bonds" in forms:
            res["bonds"] = bonds_form(a, log)
        print(json.dumps({"forms": res}), flush=True)
        return

    e = Engine(path, device=0)
    timed_options(e)
    e.update_nl()
    t0 = time.perf_counter()
    if a.set:
        # launch options given explicitly (a profiled run repeating what the
        # autotuner chose): no autotuning
        choice = {}
        for kv in a.set.split(","):
            k, v = kv.split("=")
            e.set_option(k.strip(), float(v))
            choice[k.strip()] = int(float(v))
        tuned = 0
    else:
        tuned, choice = (0, None) if a.no_autotune else e.autotune(steps_per_trial=a.tune_trial)
    e.step(max(0, a.warmup - tuned))
    e.synchronize()
    log("warm-up: %d steps (%d of them autotuning: %s) in %.2f s" % (max(a.warmup, tuned), tuned, choice,
                                                                   time.perf_counter() - t0))

    # timed region: whole periods of the neighbour-list rebuild (the rebuild
    # is part of the step loop, so the timed steps are --steps rounded up to
    # a multiple of interVerlet), device state resident; HIP events on the
    # engine stream bracket the force phase and the neighbour-list rebuilds
    period = interverlet_steps(e)
    K = max(period, -(-a.steps // period) * period)
    e.set_option("profile", 1)
    e.reset_kernel_time()
    c0 = e.counts()
    e.synchronize()
    t0 = time.perf_counter()
    e.step(K)
    e.synchronize()
    wall = time.perf_counter() - t0
    c1 = e.counts()
    nb = c1["nbodies"]
    kt = {k: e.kernel_time(k) for k in ("force", "nl")}
    value = nb * K / wall
    rebuilds = c1["nl_updates"] - c0["nl_updates"]
    if rebuilds < K // period:
        raise RuntimeError("timed region of %d steps holds %d rebuilds, expected %d" % (K, rebuilds, K // period))
    # per-kernel breakdown over the next K steps (one event pair per kernel)
    e.set_option("profile", 2)
    e.reset_kernel_time()
    e.step(K)
    e.synchronize()
    kt2 = {k: e.kernel_time(k) for k in ("predict", "interact", "gather", "breakage", "nl")}
    e.set_option("profile", 0)
    # and without events, to show what the brackets cost
    e.synchronize()
    t1 = time.perf_counter()
    e.step(K)
    e.synchronize()
    wall_noev = time.perf_counter() - t1
    log("timed: %d steps (%d rebuilds), %.4f ms/step with force-phase events, %.4f ms/step without" % (
        K, rebuilds, 1e3 * wall / K, 1e3 * wall_noev / K))

    skip_far = int(e.get_option("skip_far"))
    # roofline of the force phase (interact + gather), per launch
    n_force, ms_force = kt["force"]
    nl_listed = 0.5 * (c0["nlisted"] + c1["nlisted"])
    bytes_per = force_phase_bytes(nb, nl_listed, "Avalanche")
    t_force = ms_force / max(n_force, 1) * 1e-3
    achieved = bytes_per / t_force / 1e9

    out = {
        "metric": "body-steps/s", "value": value, "unit": "body-steps/s", "n_gpus": 1, "steps": K,
        "steps_requested": a.steps, "warmup": max(a.warmup, tuned), "ms_per_step": 1e3 * wall / K,
        "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "plain GPU form: %d convex sphero-polyhedra (hulls of 8-14 points on a 1:0.8:0.6 "
                               "ellipsoid, 4-6 mm), settled pour in a 40x40x200-size box, linkCells, ForceLaw "
                               "Avalanche, dt 2e-7 s, interVerlet 100 steps" % nb,
                   "bodies": nb, "bodies_dropped_after_settling": dropped,
                   "listed_per_body": c1["nlisted"] / nb, "active_per_body": c1["nactive"] / nb,
                   "nl_rebuilds_in_timed_region": rebuilds, "interverlet_steps": period,
                   "nl_candidate_pairs": c1["last_candidates"],
                   "force_phase_ms": 1e3 * t_force,
                   "nl_rebuild_ms": kt["nl"][1] / max(kt["nl"][0], 1),
                   "kernel_ms_per_launch": {k: (v[1] / max(v[0], 1)) for k, v in kt2.items() if v[0]},
                   "ms_per_step_without_events": 1e3 * wall_noev / K,
                   "autotuned": choice,
                   "skip_far": skip_far,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": TRAFFIC_BYTES_PER_LAUNCH[skip_far],
                     "kernel": "force phase (k_interact_seg with the far-entry skip, else k_interact or k_geom + "
                               "k_law; then k_gather + k_heavy_finish)",
                     "algorithmic_bytes_per_launch": bytes_per},
    }
    sk = skip_accounting(e, period, nb, nl_listed, "Avalanche", t_force)
    if sk is not None:
        out["roofline"]["far_entry_skip"] = sk
    e.close()
    if not a.no_copy_rate:
        try:
            bw = copy_bandwidth_child()
            out["roofline"]["measured_copy_GBs"] = bw
            out["roofline"]["frac_of_measured_copy"] = achieved / bw
        except Exception as exc:
            log("copy bandwidth not measured: %s" % exc)
    forms = [f for f in a.forms.split(",") if f]
    if forms:
        out["forms"] = {}
    if "batched" in forms:
        out["forms"]["batched"] = batched_form(a, log)
    if "avalanche" in forms:
        out["forms"]["avalanche"] = avalanche_forms(a, log, path)
    if "bonds" in forms:
        out["forms"]["bonds"] = bonds_form(a, log)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(path, a.cpu_seconds, a.cpu_threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
