"""Writes the golden vectors of tests/golden/*.npz from the reference.

Runs the trace harness built from the reference's own sources
(oracle/_ref/ref_strict, strict IEEE build, see oracle/Makefile) on seeded
initial states and stores inputs and outputs as compressed numpy archives:

  inputs     x y vx vy rad m (float64), Lx Ly
  sweep_*    the initial prediction sweep (eventListInit, EDMD.c:1935-2012):
             crossing time/direction and collision time/partner/counter
  trace_*    the first events popped by the dispatch loop (EDMD.c:534-610):
             kind, i, j, valid flag, time
  state_*    the state after the trace: x y vx vy t_local coll cell
  glob_*     t, virial sums, E, px, py, ncol, ncross after the trace
  thermo     the reference's own thermo rows (saveThermo, EDMD.c:5232-5586)
  pcf_*      calculate_pcf (pcf.c:16-75) on the input positions
  grow_*     growth starts (the radii stored are the targets; vr the rate):
             trace through the GROWSTOP event (stopGrow, EDMD.c:4740-4785)
             and the final state with every radius
  rescale_*  velocity-rescale thermostat (addNoise, EDMD.c:4828-4923): trace
             through every rescale, final state and radii, E after each
             rescale
  cli_*      the reference's own starts from its command line (random
             growth start, text load): the disks right after particlesInit,
             trace, final state and radii

Run from the repository root after `make -C oracle ref`:
    python tests/golden/make_golden.py
"""
from __future__ import annotations

import math
import os
import sys
import tempfile

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import edmd_fixtures as F  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def poly_state(n, phi, seed, mean_rad=0.5, rel_sd=0.1, trunc=2.0):
    """Square-lattice start with Gaussian radii (mean_rad, rel_sd) truncated at
    +-trunc sd; the lattice spacing exceeds the largest contact distance."""
    rng = np.random.default_rng(seed)
    rad = np.empty(0)
    while len(rad) < n:
        r = rng.normal(mean_rad, rel_sd * mean_rad, 2 * n)
        r = r[np.abs(r - mean_rad) <= trunc * rel_sd * mean_rad]
        rad = np.concatenate([rad, r])
    rad = rad[:n]
    L = math.sqrt(np.pi * np.sum(rad * rad) / phi)
    side = int(math.ceil(math.sqrt(n)))
    a = L / side
    assert a > 2 * rad.max(), "lattice too dense for these radii"
    k = np.arange(n)
    st = F.lattice_state(n, phi, seed)  # for the velocities
    st["x"] = (k % side + 0.5) * a
    st["y"] = (k // side + 0.5) * a
    st["rad"] = rad
    st["Lx"] = st["Ly"] = L
    return st


CASES = [
    # name, state builder, trace length
    ("sq2000_phi070", lambda: F.lattice_state(2000, 0.70, seed=1), 20000),
    ("sq2000_phi070_r1", lambda: F.scale_state(F.lattice_state(2000, 0.70, seed=1), 2.0), 10000),
    ("n400_phi060", lambda: F.lattice_state(400, 0.60, seed=2), 20000),
    ("tri1024_phi080", lambda: F.lattice_state(1024, 0.80, seed=5, lattice="tri"), 10000),
    ("poly500_phi050", lambda: poly_state(500, 0.50, seed=6), 10000),
    ("tiny_ncx3", lambda: F.lattice_state(20, 0.30, seed=9), 4000),
    # Boxes of one or two cells per side: the reference's crossing times wrap
    # wrongly once a cell spans half the box (PBCinsideCellX, EDMD.c:5922-5936)
    # and its event loop aborts on an overlap, so only the initial sweep (a
    # 3x3 stencil visiting wrapped duplicates) is kept for these.
    ("tiny_ncx2", lambda: F.lattice_state(12, 0.30, seed=7), 0),
    ("tiny_ncx1", lambda: F.lattice_state(4, 0.35, seed=8), 0),
]

GROWTH_CASES = [
    # name, state builder, growth rate vr, trace length (past t = 1/vr)
    ("grow_mono200", lambda: F.growth_state(200, 0.50, seed=1), 0.1, 20000),
    ("grow_poly150", lambda: F.growth_state(150, 0.72, seed=3, poly=0.1), 0.1 / 8, 32000),
]

NOISE_CASES = [
    # name, state builder, T, dtnoise, growth rate vr (0: no growth), extra
    # constantInit options, trace length.  addNoise's velocity-rescale branch
    # (EDMD.c:4880-4901) run by the reference: a hot start rescaled to a lower
    # T; inelastic disks (-r 0.9) so that every rescale moves E; a growth
    # start thermostatted while it grows (addNoise through the growth
    # pointers) and past GROWSTOP.
    ("rescale_n400", lambda: F.lattice_state(400, 0.60, seed=77, temperature=2.0), 1.5, 0.05,
     0.0, [], 25000),
    ("rescale_res09_n400", lambda: F.lattice_state(400, 0.60, seed=78), 1.0, 0.1, 0.0,
     ["-r", "0.9"], 25000),
    ("rescale_grow_poly400", lambda: F.growth_state(400, 0.70, seed=21, poly=0.1), 1.0, 0.5,
     0.05, [], 50000),
]

def text_state():
    """A text-format start (radius-1 units, as the reference's own): 196
    disks on a square lattice of spacing 2.6, every third one of radius 0.5
    (type 0, mass 0.125), the others of radius 1 (type 1, mass 1)."""
    side, a = 14, 2.6
    k = np.arange(side * side)
    big = (k % 3) != 2
    st = {"x": (k % side + 0.5) * a, "y": (k // side + 0.5) * a,
          "rad": np.where(big, 1.0, 0.5), "m": np.where(big, 1.0, 0.125),
          "Lx": side * a, "Ly": side * a}
    return st, big.astype(int)


CLI_CASES = [
    # name, command line (no -l dump: the reference's own start), trace length
    # random growth start, default binary mixture (sizeratio 0.4, 30 % small)
    ("cli_n300_phi060", ["-N", "300", "-p", "0.6", "-v", "7"], 40000),
    # monodisperse growth start in a box twice as high as wide
    ("cli_n256_phi055_aspect2", ["-N", "256", "-p", "0.55", "-x", "0", "-a", "2", "-v", "3"],
     40000),
    # a box given by its width (-X): N follows from phi
    ("cli_lx40_phi050", ["-X", "40", "-p", "0.5", "-x", "0.2", "-q", "0.6", "-v", "11"], 30000),
    # the text load path: velocities from the stream, normalizePhysicalQ
    ("cli_text196", ["-l", "@TEXT@", "-v", "5", "-E", "1.3"], 20000),
]

THERMO_CASE = ("thermo64", lambda: F.lattice_state(64, 0.50, seed=4), 1300.0, ["-o", "100"])

# Long run of the N = 2000, phi = 0.70 square-lattice system with the reference built with its shipped
# flags (-O3 -ffast-math; its trajectory leaves the strict one after ~10^4
# collisions): thermo rows every time unit to t = 60 (about 2 x 10^6
# collisions) and the states at t = 30, 40, 50, 60, for statistical
# comparisons of the reduced pressure and g(r).  The first thermo row is at
# t = 1 (the harness moves the reference's fixed firstThermo = 1000).
LONGRUN_CASE = ("sq2000_phi070_longrun", lambda: F.lattice_state(2000, 0.70, seed=1), [30.0, 40.0, 50.0, 60.0])


def state_arrays(st):
    return {"x": st["x"], "y": st["y"], "vx": st["vx"], "vy": st["vy"], "rad": st["rad"],
            "m": st["m"], "Lx": np.float64(st["Lx"]), "Ly": np.float64(st["Ly"])}


def pack_state(prefix, tr):
    s = tr["state"]
    out = {prefix + "_" + k: s[k] for k in s.dtype.names}
    g = tr["glob"]
    for k, v in g.items():
        out["glob_" + k] = np.asarray(v)
    return out


def main():
    tmp = tempfile.mkdtemp(prefix="edmd_golden_")
    for name, build, nev in CASES:
        st = build()
        dump = os.path.join(tmp, name + ".dump")
        F.write_dump(dump, st)
        sw = os.path.join(tmp, name + ".sweep")
        tr = os.path.join(tmp, name + ".trace")
        F.run_ref(["sweep", dump, sw])
        s = F.parse_sweep(sw)
        arrs = state_arrays(st)
        rec = s["rec"]
        for k in rec.dtype.names:
            arrs["sweep_" + k] = rec[k]
        arrs["grid"] = np.array([s["ncx"], s["ncy"]], dtype=np.int32)
        if nev == 0:
            np.savez_compressed(os.path.join(OUT, name + ".npz"), **arrs)
            print(name, "N=%d sweep only" % len(st["x"]))
            continue
        F.run_ref(["trace", dump, tr, nev])
        t = F.parse_trace(tr)
        trace = t["trace"]
        for k in trace.dtype.names:
            arrs["trace_" + k] = trace[k]
        arrs.update(pack_state("state", t))
        if name == "sq2000_phi070":
            pcf = os.path.join(tmp, name + ".pcf")
            F.run_ref(["pcf", dump, pcf, 0.1, st["Lx"] / 2])
            r, g = F.parse_pcf(pcf)
            arrs["pcf_r"] = r
            arrs["pcf_g"] = g
        np.savez_compressed(os.path.join(OUT, name + ".npz"), **arrs)
        print(name, "N=%d events=%d ncol=%d" % (len(st["x"]), nev, t["glob"]["ncol"]))

    for name, build, vr, nev in GROWTH_CASES:
        st = build()
        dump = os.path.join(tmp, name + ".dump")
        F.write_dump(dump, st)
        out = os.path.join(tmp, name + ".grow")
        F.run_ref(["grow", dump, out, nev, repr(vr)])
        t = F.parse_grow(out)
        arrs = state_arrays(st)
        arrs["vr"] = np.float64(vr)
        for k in t["trace"].dtype.names:
            arrs["trace_" + k] = t["trace"][k]
        arrs.update(pack_state("state", t))
        arrs["state_rad"] = t["rad"]
        assert (t["trace"]["type"] == F.GROWSTOP).sum() == 1, name
        np.savez_compressed(os.path.join(OUT, name + ".npz"), **arrs)
        print(name, "N=%d events=%d ncol=%d" % (len(st["x"]), nev, t["glob"]["ncol"]))

    make_noise(tmp)
    make_cli(tmp)
    make_longrun(tmp)

    name, build, tmax, extra = THERMO_CASE
    st = build()
    dump = os.path.join(tmp, name + ".dump")
    F.write_dump(dump, st)
    out = os.path.join(tmp, name + ".state")
    th = os.path.join(tmp, name + ".thermo")
    F.run_ref(["run", dump, out, tmax, th] + extra)
    t = F.parse_trace(out)
    rows = [list(map(float, line.split())) for line in open(th) if line.strip()]
    arrs = state_arrays(st)
    arrs.update(pack_state("state", t))
    arrs["thermo_rows"] = np.array([r[:8] for r in rows])
    arrs["tmax"] = np.float64(tmax)
    arrs["dtime_thermo"] = np.float64(float(extra[1]))
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **arrs)
    print(name, "rows=%d ncol=%d" % (len(rows), t["glob"]["ncol"]))


def make_noise(tmp):
    for name, build, temp, dtn, vr, extra, nev in NOISE_CASES:
        st = build()
        dump = os.path.join(tmp, name + ".dump")
        F.write_dump(dump, st)
        out = os.path.join(tmp, name + ".noise")
        F.run_ref(["noise", dump, out, nev, repr(temp), repr(dtn), repr(vr)] + extra)
        t = F.parse_noise(out)
        arrs = state_arrays(st)
        arrs["temperature"] = np.float64(temp)
        arrs["dtnoise"] = np.float64(dtn)
        arrs["vr"] = np.float64(vr)
        arrs["res"] = np.float64(float(extra[1]) if extra[:1] == ["-r"] else 1.0)
        for k in t["trace"].dtype.names:
            arrs["trace_" + k] = t["trace"][k]
        arrs.update(pack_state("state", t))
        arrs["state_rad"] = t["rad"]
        arrs["rescale_E"] = t["rescale_E"]
        nres = int((t["trace"]["type"] == F.ADDINGNOISE).sum())
        assert nres == len(t["rescale_E"]) and nres >= 20, (name, nres)
        np.savez_compressed(os.path.join(OUT, name + ".npz"), **arrs)
        print(name, "N=%d events=%d rescales=%d ncol=%d" % (len(st["x"]), nev, nres,
                                                            t["glob"]["ncol"]))


def make_cli(tmp):
    for name, args, nev in CLI_CASES:
        arrs = {}
        if "@TEXT@" in args:
            st, types = text_state()
            path = os.path.join(tmp, name + ".txt")
            F.write_text_state(path, st, types)
            args = [path if a == "@TEXT@" else a for a in args]
            arrs["text_x"], arrs["text_y"] = st["x"], st["y"]
            arrs["text_rad"], arrs["text_m"] = st["rad"], st["m"]
            arrs["text_type"] = types.astype(np.int32)
            arrs["text_Lx"], arrs["text_Ly"] = np.float64(st["Lx"]), np.float64(st["Ly"])
        out = os.path.join(tmp, name + ".cli")
        F.run_ref(["cli", out, nev] + args)
        t = F.parse_cli(out)
        arrs["args"] = np.array(" ".join(a for a in args if not a.endswith(".txt")))
        arrs["Lx"], arrs["Ly"], arrs["vr"] = np.float64(t["Lx"]), np.float64(t["Ly"]), np.float64(t["vr"])
        for k, v in t["init"].items():
            arrs["init_" + k] = v
        for k in t["trace"].dtype.names:
            arrs["trace_" + k] = t["trace"][k]
        arrs.update(pack_state("state", t))
        arrs["state_rad"] = t["rad"]
        np.savez_compressed(os.path.join(OUT, name + ".npz"), **arrs)
        ngs = int((t["trace"]["type"] == F.GROWSTOP).sum())
        print(name, "N=%d Lx=%.3f Ly=%.3f vr=%g events=%d growstop=%d ncol=%d"
              % (len(t["init"]["x"]), t["Lx"], t["Ly"], t["vr"], nev, ngs, t["glob"]["ncol"]))


def make_longrun(tmp):
    name, build, times = LONGRUN_CASE
    st = build()
    dump = os.path.join(tmp, name + ".dump")
    F.write_dump(dump, st)
    arrs = state_arrays(st)
    for k, tm in enumerate(times):
        out = os.path.join(tmp, "%s_%d.state" % (name, k))
        th = os.path.join(tmp, "%s_%d.thermo" % (name, k))
        os.environ["EDMD_HARNESS_FIRST_THERMO"] = "1"
        try:
            F.run_ref(["run", dump, out, tm, th, "-o", "1"], kind="fast")
        finally:
            del os.environ["EDMD_HARNESS_FIRST_THERMO"]
        t = F.parse_trace(out)
        s = t["state"]
        for f in ("x", "y", "vx", "vy", "t"):
            arrs["snap%d_%s" % (k, f)] = s[f]
        arrs["snap%d_time" % k] = np.float64(t["glob"]["t"])
        if k == len(times) - 1:
            rows = [list(map(float, line.split())) for line in open(th) if line.strip()]
            arrs["thermo_rows"] = np.array([r[:8] for r in rows])
    arrs["snap_tmax"] = np.array(times)
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **arrs)
    print(name, "rows=%d" % len(arrs["thermo_rows"]))


if __name__ == "__main__":
    if sys.argv[1:] == ["longrun"]:
        make_longrun(tempfile.mkdtemp(prefix="edmd_golden_"))
    elif sys.argv[1:] == ["noise"]:
        make_noise(tempfile.mkdtemp(prefix="edmd_golden_"))
    elif sys.argv[1:] == ["cli"]:
        make_cli(tempfile.mkdtemp(prefix="edmd_golden_"))
    else:
        main()
