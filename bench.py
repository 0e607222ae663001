#!/usr/bin/env python3
"""Benchmark of the event-driven hard-disk engine on one MI355X.

Headline (`value`): aggregate valid collisions per second of batched
replicas -- 4096 independent systems, N uniform in [256, 2048], phi uniform
in [0.50, 0.75], diameter 1, E/N = 1, per-replica seed, FP64 -- all in one
persistent launch.  Starts are triangular lattices fitted to the square box
(a square lattice of N sites overlaps above phi ~ 0.70 for most N).  One step = every
replica advances by --coll-per-step valid collisions (default 10^4), so the
default 10 timed steps are the workload's 10^5 collisions per replica.

Reported beside it, in the same run:
  plain        single large system: N = 2^20 disks at phi = 0.80 on a
               triangular lattice, the same event loop on one system
               (collisions/s), and the batched/plain ratio;
  sweep        the full prediction sweep of that system (kernel k_sweep_tiled):
               GB/s of algorithmic traffic, 80 B per disk, against the
               MI355X HBM peak (the `roofline` object);
  adjoint      derivatives of the event map: 512 replicas of N = 400 at
               phi = 0.60 (seeds base + id), 10^4 collisions each, checkpoint
               every 256 events: forward run with checkpoints, adjoint of the
               time-averaged collision virial with recomputation of every
               segment, tangent pass (finite-time growth of a random unit
               tangent), and the same with the whole tape kept in HBM;
  dense        dense polydisperse batch with thermostat: 64 replicas of
               N = 65536, diameters N(1, 0.1) truncated at +-3 sigma,
               phi = 0.80, velocity-rescaling thermostat to T = 1 every 0.1
               time units, started by growth from radius 0 (random centres,
               growth rate 0.2: GROWSTOP at t = 5, about 31 collisions per
               disk; the reference's own rate for phi = 0.8 is 0.0125); the
               growth phase and then 2 * 10^5 thermostatted collisions per
               replica are timed separately;
  cpu_baseline the reference's own CPU event loop (built with its shipped
               flags) on a bounded sample of replicas, one core.

Before each leg's timed region every batch is autotuned
(edmd_batch_autotune: each sweep tile shape timed, and for batches whose
systems fit one workgroup's LDS a trial run from HBM and from LDS; state
saved and restored); --no-tune keeps the defaults.  The choices are
reported under "tuning".
"""
import argparse
import importlib.util
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# HBM bytes per launch of k_sweep_tiled at N = 2^20, measured with rocprofv3
# FETCH_SIZE / WRITE_SIZE counters and calibrated against a copy of known
# size; the summary holding it (with the commands) is committed here:
SWEEP_TRAFFIC_PROFILE = os.path.join(HERE, "profiles", "k_sweep_tiled_N1048576_traffic.md")
PKG = os.path.join(HERE, "graphical-edmd_amd")
HBM_PEAK_GBS = 8000.0
SWEEP_BYTES_PER_DISK = 80  # x y vx vy t (40) + cell + coll + cell table (12) + slots (28)


def load_binding():
    spec = importlib.util.spec_from_file_location("edmd_hip", os.path.join(PKG, "edmd_hip.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["edmd_hip"] = mod
    spec.loader.exec_module(mod)
    return mod


def sweep_traffic():
    """HBM bytes per sweep launch from the committed counter summary (its
    line 'traffic_bytes_per_launch: <value>'), or None."""
    try:
        with open(SWEEP_TRAFFIC_PROFILE) as f:
            for line in f:
                if line.strip().startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def copy_calibration():
    """A plain device copy of known size (1 GiB read, 1 GiB written, three
    times), run under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE to scale those
    counters to bytes."""
    import torch
    n = 256 * 1024 * 1024
    x = torch.ones(n, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print("copied 3 x 1 GiB", flush=True)


def replica_params(nrep, seed):
    rng = np.random.default_rng(seed)
    ns = rng.integers(256, 2049, size=nrep)
    phis = rng.uniform(0.50, 0.75, size=nrep)
    return ns, phis


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores():
    """Cores this process may use, at most 16 (one GPU's share of a box)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def _write_dump(E, st, path):
    args = [st[f].ctypes.data for f in ("x", "y", "vx", "vy", "rad", "m")]
    E.check(E.lib().edmd_dump_write(path.encode(), 0, 0.0, len(st["x"]), st["Lx"], st["Ly"],
                                    *args, 1))


def cpu_baseline(E, ns, phis, base_seed, coll, seconds, plain_n):
    """The reference's event loop on the host cores, in the same run: one
    process per core, each on its own replica of the batched workload
    (oracle/_ref/ref_fast, the reference's sources built with its Makefile
    flags), each stopping after `coll` collisions or `seconds`; then one
    process on the N = plain_n, phi = 0.80 triangular-lattice system of the
    plain GPU leg.  Falls back to the oracle (one core) where the reference
    build is absent."""
    ref = os.path.join(HERE, "oracle", "_ref", "ref_fast")
    tmp = tempfile.mkdtemp(prefix="edmd_cpu_")
    if not os.path.exists(ref):
        sys.path.insert(0, os.path.join(HERE, "tests"))
        import oracle_bind
        st = E.init_lattice(int(ns[0]), float(phis[0]), rad=0.5, lattice=2, seed=base_seed)
        o = oracle_bind.Oracle(st)
        t0 = time.perf_counter()
        o.run_collisions(coll)
        el = time.perf_counter() - t0
        v = o.observables()["ncol"] / el
        return {"value": v, "unit": "collisions/s", "cores": 1, "kind": "port",
                "per_core": v, "aggregate": v, "nproc": os.cpu_count(), "cpu_model": cpu_model(),
                "sample": "replica 0 of the batched workload, %d collisions, the oracle" % coll}
    cores = host_cores()
    procs = []
    for k in range(cores):
        st = E.init_lattice(int(ns[k]), float(phis[k]), rad=0.5, lattice=2, seed=base_seed + k)
        path = os.path.join(tmp, "r%d.dump" % k)
        _write_dump(E, st, path)
        procs.append(subprocess.Popen([ref, "bench", path, str(coll), str(seconds)],
                                      stdout=subprocess.DEVNULL, stderr=subprocess.PIPE,
                                      text=True))
    rates = []
    for pr in procs:
        _, err = pr.communicate(timeout=10 * seconds + 120)
        if pr.returncode != 0:
            raise RuntimeError("reference CPU run failed: " + err[-500:])
        rates.append(json.loads(err.strip().splitlines()[-1])["coll_per_s"])
    plain = None
    if plain_n > 0:
        st = E.init_lattice(plain_n, 0.80, rad=0.5, lattice=1, seed=base_seed)
        path = os.path.join(tmp, "plain.dump")
        _write_dump(E, st, path)
        r = subprocess.run([ref, "bench", path, str(1 << 40), str(seconds)], capture_output=True,
                           text=True, timeout=20 * seconds + 300, check=True)
        plain = json.loads(r.stderr.strip().splitlines()[-1])["coll_per_s"]
        os.remove(path)
    agg = float(sum(rates))
    return {"value": agg, "unit": "collisions/s", "cores": cores, "kind": "reference",
            "per_core": agg / cores, "per_core_min": min(rates), "per_core_max": max(rates),
            "aggregate": agg, "nproc": os.cpu_count(), "cpu_model": cpu_model(),
            "plain_cpu_collisions_per_s": plain,
            "sample": "%d processes at once, one per core, each on its own replica of the "
                      "batched workload (N~U[256,2048], phi~U[0.5,0.75]) for %d collisions or "
                      "%g s; plain: one process on the N=%d phi=0.80 triangular lattice for %g s"
                      % (cores, coll, seconds, plain_n, seconds)}


def tune(b, args, sweep_reps, trial_events):
    """Autotunes batch b unless --no-tune; returns what was kept."""
    if args.no_tune:
        rep = None
    else:
        rep = b.autotune(sweep_reps=sweep_reps, trial_events=trial_events)
    sh = b.sweep_shape()
    out = {"sweep_tile": [sh["tile_x"], sh["tile_y"]], "sweep_threads": sh["threads"],
           "lds_resident": b.residency()["lds"], "loop_systems_per_group": b.loop_waves()}
    if rep is not None:
        out["sweep_ms_by_shape"] = [round(x, 5) for x in rep["sweep_ms"]]
        if rep["loop_ms_hbm"] >= 0:
            out["trial_ms_hbm_lds"] = [round(rep["loop_ms_hbm"], 3), round(rep["loop_ms_lds"], 3)]
            out["trial_ms_hbm_2_4_per_group"] = [round(x, 3) for x in rep["loop_ms_packed"]]
    return out


def adjoint_leg(E, args):
    """Forward run with checkpoints, adjoint and tangent passes with
    recomputation, then the whole-tape variant, on the same replicas."""
    nrep, n, ncoll = args.adj_replicas, args.adj_n, args.adj_collisions
    states = [E.init_lattice(n, args.adj_phi, rad=0.5, lattice=2, seed=args.seed + k)
              for k in range(nrep)]
    rng = np.random.default_rng(args.seed)
    tin = []
    for _ in range(nrep):
        v = rng.standard_normal((n, 4))
        tin.append(v / np.linalg.norm(v))
    res = {"workload": "%d replicas, N=%d, phi=%.2f, %d collisions each"
                       % (nrep, n, args.adj_phi, ncoll),
           "checkpoint_every": args.adj_every}
    for mode, every, cap in (("checkpointed", args.adj_every, 0), ("whole_tape", 0, ncoll)):
        b = E.Batch(states)
        res.setdefault("tuning", tune(b, args, 2, 2000))
        b.deriv_begin(checkpoint_every=every, tape_capacity=cap)
        t0 = time.perf_counter()
        b.deriv_run(max_collisions=ncoll, max_events=1 << 62)
        t1 = time.perf_counter()
        W = np.array([b.deriv_values(k)[0] for k in range(nrep)])
        T = np.array([b.deriv_values(k)[1] for k in range(nrep)])
        t2 = time.perf_counter()
        grad = b.adjoint(lw=1.0 / T, lt=-W / T ** 2)
        t3 = time.perf_counter()
        tout, dw, dtl = b.tangent(tin)
        t4 = time.perf_counter()
        growth = np.array([np.log(np.linalg.norm(v)) for v in tout]) / T
        res[mode] = {"forward_s": t1 - t0, "adjoint_s": t3 - t2, "tangent_s": t4 - t3,
                     "collisions_per_s_forward": nrep * ncoll / (t1 - t0),
                     "collisions_per_s_forward_plus_adjoint": nrep * ncoll / (t1 - t0 + t3 - t2)}
        log("adjoint/%s: forward %.3f s, adjoint %.3f s, tangent %.3f s"
            % (mode, t1 - t0, t3 - t2, t4 - t3))
        res["virial_time_average_mean"] = float(np.mean(W / T))
        res["tangent_growth_rate_mean"] = float(np.mean(growth))
        res["gradient_norm_mean"] = float(np.mean([np.linalg.norm(g) for g in grad]))
        b.close()
    return res


def dense_leg(E, args):
    """Growth start, then the thermostatted dense dynamics, of 64 large
    polydisperse replicas in one launch."""
    nrep, n, vr = args.dense_replicas, args.dense_n, args.dense_vr
    t0 = time.perf_counter()
    states = [E.init_random(n, 0.80, rad=0.5, poly=0.1, seed=args.seed + 7919 * (k + 1),
                            e_per=0.05) for k in range(nrep)]
    b = E.Batch(states, vr=vr, noise=2, dtnoise=0.1, temperature=1.0, einit=1.0)
    log("dense: %d replicas of N=%d built in %.1f s" % (nrep, n, time.perf_counter() - t0))
    res = {"workload": "%d replicas, N=%d, polydisperse N(1,0.1) diameters, phi=0.80, "
                       "thermostat T=1 every 0.1, growth start at rate %g"
                       % (nrep, n, vr), "growth_rate": vr}
    # growth: up to just before GROWSTOP (stopGrow resets the counters)
    t0 = time.perf_counter()
    b.run(tmax=(1.0 / vr) * (1 - 1e-12), max_events=1 << 62)
    t1 = time.perf_counter()
    obs = [b.observables(k) for k in range(nrep)]
    gcol = sum(o["ncol"] for o in obs)
    gev = sum(o["nevents"] for o in obs)
    res["growth"] = {"seconds": t1 - t0, "collisions": gcol, "events": gev,
                     "collisions_per_s": gcol / (t1 - t0), "events_per_s": gev / (t1 - t0)}
    log("dense growth: %d collisions in %.2f s = %.4e coll/s" % (gcol, t1 - t0, gcol / (t1 - t0)))
    b.run(tmax=1.0 / vr + 1e-6, max_events=1 << 62)  # GROWSTOP: radii reach their targets
    res["tuning"] = tune(b, args, 3, 0)
    obs = [b.observables(k) for k in range(nrep)]
    c0 = sum(o["ncol"] for o in obs)
    e0 = sum(o["nevents"] for o in obs)
    t0 = time.perf_counter()
    b.run(max_collisions=args.dense_collisions, max_events=1 << 62)
    t1 = time.perf_counter()
    obs = [b.observables(k) for k in range(nrep)]
    ncol = sum(o["ncol"] for o in obs) - c0
    nev = sum(o["nevents"] for o in obs) - e0
    res["thermostatted"] = {"seconds": t1 - t0, "collisions": ncol, "events": nev,
                            "collisions_per_replica": args.dense_collisions,
                            "collisions_per_s": ncol / (t1 - t0), "events_per_s": nev / (t1 - t0),
                            "e_per_n_mean": float(np.mean([o["E"] / n for o in obs]))}
    log("dense thermostatted: %d collisions in %.2f s = %.4e coll/s"
        % (ncol, t1 - t0, ncol / (t1 - t0)))
    b.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--replicas", type=int, default=4096)
    ap.add_argument("--coll-per-step", type=int, default=10000)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--plain-n", type=int, default=1 << 20)
    ap.add_argument("--plain-collisions", type=int, default=300000)
    ap.add_argument("--sweep-reps", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-collisions", type=int, default=20000000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default sweep shapes and residency (no autotuning)")
    ap.add_argument("--adj-replicas", type=int, default=512)
    ap.add_argument("--adj-n", type=int, default=400)
    ap.add_argument("--adj-phi", type=float, default=0.60)
    ap.add_argument("--adj-collisions", type=int, default=10000)
    ap.add_argument("--adj-every", type=int, default=256)
    ap.add_argument("--dense-replicas", type=int, default=64)
    ap.add_argument("--dense-n", type=int, default=65536)
    ap.add_argument("--dense-vr", type=float, default=0.2)
    ap.add_argument("--dense-collisions", type=int, default=200000)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run the 1 GiB copy used to calibrate memory counters")
    ap.add_argument("--sweep-only", action="store_true",
                    help="only build the N = 2^20 system, tune it and time --sweep-reps sweeps "
                         "(the workload of the sweep profiles)")
    ap.add_argument("--melt", type=float, default=0.0,
                    help="with --sweep-only: first run the system for this many valid "
                         "collisions per disk (the melt of the N = 2^20 configuration is 50)")
    args = ap.parse_args()
    if args.calibrate_copy:
        copy_calibration()
        return
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; replicas would shard across GPUs with no "
              "exchange, but --gpus %d is not supported here" % args.gpus, file=sys.stderr)
        sys.exit(2)

    E = load_binding()
    if E.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.sweep_only:
        big = E.Batch([E.init_lattice(args.plain_n, 0.80, rad=0.5, lattice=1, seed=args.seed)])
        out = {"n": args.plain_n}
        if args.melt > 0:
            target = int(args.melt * args.plain_n)
            chunk = 5000000
            t0 = time.perf_counter()
            done = 0
            while done < target:
                done = min(target, done + chunk)
                big.run(max_collisions=done, max_events=1 << 62)
                log("melt: %d of %d collisions, %.0f s" % (done, target, time.perf_counter() - t0))
            el = time.perf_counter() - t0
            ob = big.observables(0)
            out["melt"] = {"collisions": int(ob["ncol"]), "events": int(ob["nevents"]),
                           "seconds": el, "collisions_per_s": ob["ncol"] / el, "t": ob["t"]}
            big.sweep_predict()
        out["tuning"] = tune(big, args, 5, 0)
        out["sweep_ms"] = big.time_sweep(args.sweep_reps)
        print(json.dumps(out), flush=True)
        return

    # ---- batched replicas -------------------------------------------------
    ns, phis = replica_params(args.replicas, args.seed)
    t0 = time.perf_counter()
    states = [E.init_lattice(int(n), float(p), rad=0.5, lattice=2, seed=args.seed + k)
              for k, (n, p) in enumerate(zip(ns, phis))]
    b = E.Batch(states)
    log("batch of %d replicas (%d disks) built in %.1f s" % (len(states), int(ns.sum()),
                                                            time.perf_counter() - t0))
    tuning = {"batched": tune(b, args, 2, 2000)}
    cps = args.coll_per_step
    for w in range(args.warmup):
        b.run(max_collisions=cps * (w + 1), max_events=1 << 62)
    ncol0 = sum(b.observables(k)["ncol"] for k in range(len(states)))
    nev0 = sum(b.observables(k)["nevents"] for k in range(len(states)))
    t0 = time.perf_counter()
    for s in range(args.steps):
        b.run(max_collisions=cps * (args.warmup + s + 1), max_events=1 << 62)
    t1 = time.perf_counter()
    obs = [b.observables(k) for k in range(len(states))]
    ncol = sum(o["ncol"] for o in obs) - ncol0
    nev = sum(o["nevents"] for o in obs) - nev0
    elapsed = t1 - t0
    value = ncol / elapsed
    log("batched: %d collisions (%d events) in %.3f s = %.4e coll/s" % (ncol, nev, elapsed, value))
    b.close()
    del states

    # ---- plain single large system: sweep roofline and event loop ---------
    st2 = E.init_lattice(args.plain_n, 0.80, rad=0.5, lattice=1, seed=args.seed)
    big = E.Batch([st2])
    tuning["plain"] = tune(big, args, 5, 0)
    sweep_ms = big.time_sweep(args.sweep_reps)
    sweep_bytes = SWEEP_BYTES_PER_DISK * args.plain_n
    sweep_gbs = sweep_bytes / (sweep_ms * 1e-3) / 1e9
    log("sweep: N=%d  %.4f ms  %.1f GB/s" % (args.plain_n, sweep_ms, sweep_gbs))
    big.run(max_events=20000)  # first events off the lattice
    o0 = big.observables(0)
    t0 = time.perf_counter()
    big.run(max_collisions=o0["ncol"] + args.plain_collisions, max_events=1 << 62)
    t1 = time.perf_counter()
    o1 = big.observables(0)
    plain = (o1["ncol"] - o0["ncol"]) / (t1 - t0)
    plain_ev = (o1["nevents"] - o0["nevents"]) / (t1 - t0)
    log("plain: %d collisions in %.3f s = %.4e coll/s" % (o1["ncol"] - o0["ncol"], t1 - t0, plain))
    # the sweep's conservative root screen on the state the loop left:
    # timed against the unscreened sweep, then counted (roots skipped, and
    # misses: skipped roots the minimum would have taken)
    screen = {"ms_unscreened": big.time_sweep(args.sweep_reps)}
    big.set_screen(1)
    screen["ms_screened"] = big.time_sweep(args.sweep_reps)
    big.set_screen(2)
    big.sweep_predict()
    marks, solved, misses = big.screen_stats()
    screen.update(marked=marks, solved=solved, misses=misses,
                  solved_fraction=solved / marks if marks else None)
    log("sweep screen: %.4f ms against %.4f ms unscreened, %d of %d roots solved, %d misses"
        % (screen["ms_screened"], screen["ms_unscreened"], solved, marks, misses))
    big.close()

    adj = None
    if args.adj_replicas > 0:
        adj = adjoint_leg(E, args)

    dense = None
    if args.dense_replicas > 0:
        dense = dense_leg(E, args)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(E, ns, phis, args.seed, args.cpu_collisions, args.cpu_seconds,
                           args.plain_n)
        log("cpu baseline: %.4e coll/s aggregate on %d cores (%s)"
            % (cpu["value"], cpu["cores"], cpu["kind"]))

    out = {
        "metric": "collisions/sec, batched replicas vs plain single system; prediction GB/s",
        "value": value,
        "unit": "collisions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "batched replicas: %d systems, N~U[256,2048], phi~U[0.50,0.75], "
                        "diameter 1, E/N=1, %d collisions per replica per step"
                        % (args.replicas, cps),
            "replicas": args.replicas,
            "disks": int(ns.sum()),
            "plain_system": "N=%d, phi=0.80, triangular lattice start" % args.plain_n,
        },
        "events_per_s": nev / elapsed,
        "plain_collisions_per_s": plain,
        "plain_events_per_s": plain_ev,
        "batched_over_plain": value / plain,
        "sweep": {"n": args.plain_n, "ms": sweep_ms, "GB_per_s": sweep_gbs, "screen": screen,
                  "tile": tuning["plain"]["sweep_tile"]},
        "tuning": tuning,
        "roofline": {"kernel": "k_sweep_tiled", "bound": "hbm", "achieved": sweep_gbs,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": sweep_gbs / HBM_PEAK_GBS,
                     "traffic": sweep_traffic() if args.plain_n == 1 << 20 else None},
        "adjoint": adj,
        "dense": dense,
        "cpu_baseline": cpu,
        "plain_over_cpu_core": (plain / cpu["plain_cpu_collisions_per_s"]
                                if cpu and cpu.get("plain_cpu_collisions_per_s") else None),
        "batched_over_cpu_aggregate": value / cpu["aggregate"] if cpu else None,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
