"""LDS-resident event loops (edmd_batch_set_residency): a batch whose systems
run with their particles, event slots, calendar, ranks and cell lists in
the workgroup's LDS evolves bit for bit as the same batch run from device
memory, and as the oracle evolves each system alone.

The reference runs one system in host memory (mainLoop, EDMD.c:5595-5630);
residency is this library's choice of where the batched loop keeps its
arrays, so every comparison is against the HBM loop or the oracle."""
import numpy as np
import pytest

import edmd_fixtures as F
from conftest import bits

pytestmark = pytest.mark.gpu


def small_states(k0=0, count=6):
    rng = np.random.default_rng(31 + k0)
    out = []
    for k in range(count):
        n = int(rng.integers(100, 700))
        phi = float(rng.uniform(0.5, 0.75))
        out.append(F.lattice_state(n, phi, seed=300 + k0 + k, lattice="fit"))
    return out


def polydisperse(n, seed):
    """Square lattice at phi 0.5 (spacing 1.25) with radii uniform in
    [0.4, 0.6]: no overlaps."""
    st = F.lattice_state(n, 0.5, seed=seed, lattice="fit")
    rng = np.random.default_rng(seed)
    st["rad"] = rng.uniform(0.4, 0.6, n)
    st["m"] = (st["rad"] / 0.5) ** 3
    return st


def pair(gpu, states, **kw):
    """The same batch twice: LDS-resident and device-memory loops."""
    a = gpu.Batch(states, **kw)
    a.set_residency(gpu.RES_LDS)
    b = gpu.Batch(states, **kw)
    b.set_residency(gpu.RES_HBM)
    assert a.residency()["lds"] and not b.residency()["lds"]
    return a, b


def assert_batches_equal(a, b, nsys, traces=True):
    for s in range(nsys):
        sa, sb = a.state(s), b.state(s)
        for f in ("x", "y", "vx", "vy", "t"):
            np.testing.assert_array_equal(bits(sa[f]), bits(sb[f]), err_msg="system %d %s" % (s, f))
        np.testing.assert_array_equal(sa["coll"], sb["coll"])
        np.testing.assert_array_equal(sa["cx"], sb["cx"])
        np.testing.assert_array_equal(sa["cy"], sb["cy"])
        oa, ob = a.observables(s), b.observables(s)
        for k in ("t", "ncol", "ncross", "nevents"):
            assert oa[k] == ob[k], (s, k, oa[k], ob[k])
        ea, eb = a.events(s), b.events(s)
        for k in ea:
            np.testing.assert_array_equal(ea[k], eb[k], err_msg="system %d %s" % (s, k))
        ca, cb = a.calendar(s), b.calendar(s)
        for k in ca:
            np.testing.assert_array_equal(ca[k], cb[k], err_msg="system %d calendar %s" % (s, k))
        if traces:
            ta, tb = a.trace(s), b.trace(s)
            assert len(ta) == len(tb)
            for k in ("type", "i", "j", "valid"):
                np.testing.assert_array_equal(ta[k], tb[k])
            np.testing.assert_array_equal(bits(ta["t"]), bits(tb["t"]))


def test_residency_query_and_limits(gpu):
    """Auto mode picks LDS for a small batch that is resident at once and
    HBM for systems whose arrays exceed one workgroup's LDS; forcing LDS on
    those is an argument error that leaves the batch usable."""
    small = gpu.Batch(small_states(count=4))
    r = small.residency()
    print("small batch residency", r, flush=True)
    assert r["lds"] and r["lds_bytes"] > 0 and r["lds_slots"] >= 4
    big = gpu.Batch([F.lattice_state(4000, 0.7, seed=5, lattice="fit")])
    r = big.residency()
    print("4000-disk residency", r, flush=True)
    assert not r["lds"] and r["lds_slots"] == 0
    with pytest.raises(gpu.EdmdError):
        big.set_residency(gpu.RES_LDS)
    assert not big.residency()["lds"]
    big.run(max_events=2000)
    assert big.observables(0)["nevents"] == 2000
    with pytest.raises(gpu.EdmdError):
        small.set_residency(7)


def test_lds_loop_equals_hbm_loop(gpu):
    """Traces, states, event slots, calendars and counters agree bit for bit
    after runs split over several launches (arrays go back to device memory
    and come in again between them)."""
    states = small_states()
    nev = 12000
    a, b = pair(gpu, states, trace_capacity=3 * nev)
    for k in (1, 2, 3):
        a.run(max_events=nev * k)
        b.run(max_events=nev * k)
    assert_batches_equal(a, b, len(states))
    # bounded by collisions and by time as well
    a.run(max_collisions=a.observables(0)["ncol"] + 5000)
    b.run(max_collisions=b.observables(0)["ncol"] + 5000)
    t = a.observables(1)["t"] + 0.5
    a.run(tmax=t)
    b.run(tmax=t)
    assert_batches_equal(a, b, len(states), traces=False)


def test_lds_loop_matches_oracle(gpu, oracle):
    """A system of 400 disks at phi 0.60 in an LDS-resident loop:
    trace and state bit for bit as the oracle's over 10^4 events."""
    st = F.lattice_state(400, 0.6, seed=4, lattice="fit")
    nev = 10000
    b = gpu.Batch([st, F.lattice_state(256, 0.72, seed=9, lattice="fit")], trace_capacity=nev)
    b.set_residency(gpu.RES_LDS)
    b.run(max_events=nev)
    o = oracle.Oracle(st)
    code, tr = o.run(nev)
    assert code == 0
    g = b.trace(0)
    assert len(g) == len(tr)
    for k in ("type", "i", "j", "valid"):
        np.testing.assert_array_equal(g[k], tr[k])
    np.testing.assert_array_equal(bits(g["t"]), bits(tr["t"]))
    so, sb = o.state(), b.state(0)
    for f in ("x", "y", "vx", "vy", "t"):
        np.testing.assert_array_equal(bits(sb[f]), bits(so[f]), err_msg=f)


def test_lds_growth_and_thermostat(gpu):
    """Growing systems (radii grow until GROWSTOP, growth rates held in LDS)
    and thermostat rescales with thermo rows: LDS and HBM loops agree."""
    states = [F.growth_state(300, 0.55, 21), F.growth_state(256, 0.7, 22, poly=0.1),
              F.lattice_state(350, 0.6, seed=23, lattice="fit")]
    vrs = [0.2, 0.05, 0.0]
    kw = dict(per_system={"vr": vrs}, trace_capacity=30000)
    a, b = pair(gpu, states, **kw)
    a.run(max_events=30000)
    b.run(max_events=30000)
    assert_batches_equal(a, b, len(states))
    for s in range(2):
        np.testing.assert_array_equal(bits(a.radii(s)), bits(b.radii(s)))
    states = small_states(50, 3)
    kw = dict(noise=2, dtnoise=0.05, temperature=1.5, first_thermo=0.1, dtime_thermo=0.1,
              thermo_capacity=256)
    a, b = pair(gpu, states, **kw)
    a.run(max_events=20000)
    b.run(max_events=20000)
    assert_batches_equal(a, b, len(states), traces=False)
    for s in range(len(states)):
        ra, rb = a.thermo(s), b.thermo(s)
        assert len(ra) == len(rb) > 5
        for k in ra.dtype.names:
            np.testing.assert_array_equal(ra[k], rb[k])


def test_lds_epochs_and_taped_adjoint(gpu):
    """hipGraph epochs and a taped run with checkpoints and the adjoint pass
    give the same counters, tapes and gradients in both residencies."""
    states = small_states(80, 4)
    a, b = pair(gpu, states)
    ra = a.run_epochs(4, events_per_epoch=3000)
    rb = b.run_epochs(4, events_per_epoch=3000)
    for k in ra.dtype.names:
        np.testing.assert_array_equal(ra[k], rb[k])
    assert_batches_equal(a, b, len(states), traces=False)
    # the whole tape kept
    for x in (a, b):
        x.deriv_begin(checkpoint_every=0, tape_capacity=6500)
        x.deriv_run(max_events=12000 + 6000)
    for s in range(len(states)):
        ta, tb = a.tape(s), b.tape(s)
        assert len(ta) == len(tb) > 0
        for k in ta.dtype.names:
            np.testing.assert_array_equal(ta[k], tb[k])
        assert a.deriv_values(s) == b.deriv_values(s)
    ga, gb = a.adjoint(lw=1.0), b.adjoint(lw=1.0)
    for s in range(len(states)):
        np.testing.assert_array_equal(bits(ga[s]), bits(gb[s]))
    # checkpoints every 256 events, segments recomputed in the adjoint pass
    for x in (a, b):
        x.deriv_end()
        x.deriv_begin(checkpoint_every=256)
        x.deriv_run(max_events=18000 + 6000)
    ga, gb = a.adjoint(lw=1.0, lt=-0.5), b.adjoint(lw=1.0, lt=-0.5)
    for s in range(len(states)):
        assert a.deriv_values(s) == b.deriv_values(s)
        np.testing.assert_array_equal(bits(ga[s]), bits(gb[s]))
    a.deriv_end()
    b.deriv_end()


def test_every_sweep_shape_predicts_identically(gpu):
    """Every tile shape of the prediction sweep gives the same event slots
    and calendar, bit for bit: unit-diameter disks at phi 0.7 and 0.8, the
    reference CLI's radius 1, a polydisperse system and a box three cells
    wide, after 500 events; and boxes two cells wide (the halo wraps onto
    the tile itself) at the start.  Events in two-cell boxes are left out:
    the reference's own dynamics break down there (the oracle reaches
    negative times and overlaps within a few hundred events)."""
    moving = [F.lattice_state(2000, 0.7, seed=61, lattice="fit"),
              F.lattice_state(3000, 0.8, seed=62, lattice="fit"),
              F.lattice_state(900, 0.5, seed=63, rad=1.0, lattice="fit"),
              polydisperse(700, 64),
              F.lattice_state(20, 0.4, seed=4, lattice="fit")]
    still = [F.lattice_state(12, 0.3, seed=2, lattice="fit"),
             F.lattice_state(14, 0.35, seed=6, lattice="fit"),
             F.lattice_state(3000, 0.8, seed=67, lattice="fit")]
    ref = None
    for shape in range(8):
        got = []
        for states, nev in ((moving, 500), (still, 0)):
            b = gpu.Batch(states)
            if nev:
                b.run(max_events=nev)
            b.set_sweep_shape(shape)
            sh = b.sweep_shape()
            assert sh["shape"] == shape and sh["tile_x"] > 0 and sh["threads"] % 64 == 0
            b.sweep_predict()
            got += [(b.events(s), b.calendar(s)) for s in range(len(states))]
        if ref is None:
            ref = got
            continue
        for s in range(len(got)):
            for part in (0, 1):
                for k in ref[s][part]:
                    np.testing.assert_array_equal(got[s][part][k], ref[s][part][k],
                                                  err_msg="shape %d system %d %s" % (shape, s, k))
    with pytest.raises(gpu.EdmdError):
        b.set_sweep_shape(8)


def test_autotune_restores_state(gpu):
    """The autotuner times every sweep shape and both residencies on the
    batch itself; afterwards the batch continues exactly as an untuned twin
    (traces, states and calendars bit for bit)."""
    states = small_states(120, 5)
    a = gpu.Batch(states, trace_capacity=8000)
    b = gpu.Batch(states, trace_capacity=8000)
    a.run(max_events=3000)
    b.run(max_events=3000)
    rep = a.autotune(sweep_reps=3, trial_events=1500)
    print("autotune report", rep, flush=True)
    assert all(ms > 0 for ms in rep["sweep_ms"])
    assert rep["loop_ms_hbm"] > 0 and rep["loop_ms_lds"] > 0
    assert rep["sweep_shape"] == a.sweep_shape()["shape"]
    assert rep["in_lds"] == a.residency()["lds"]
    # the kept shape is the fastest, or the previous one within 3% of it
    assert rep["sweep_ms"][rep["sweep_shape"]] <= min(rep["sweep_ms"]) / 0.97
    assert_batches_equal(a, b, len(states))
    a.run(max_events=8000)
    b.run(max_events=8000)
    assert_batches_equal(a, b, len(states))
    a.sweep_predict()
    b.sweep_predict()
    assert_batches_equal(a, b, len(states), traces=False)
    # a system too large for LDS: sweep shapes only
    # a system too large for LDS: HBM configurations only
    big = gpu.Batch([F.lattice_state(5000, 0.75, seed=8, lattice="fit")])
    rep = big.autotune(sweep_reps=2, trial_events=1000)
    assert rep["loop_ms_lds"] < 0 and not rep["in_lds"]
    assert rep["loop_ms_hbm"] > 0 and all(ms > 0 for ms in rep["loop_ms_packed"])
    assert rep["loop_systems_per_group"] == big.loop_waves() in (1, 2, 4)
    assert rep["in_lds"] or rep["loop_systems_per_group"] in (1, 2, 4)


def test_systems_per_workgroup_give_identical_results(gpu, oracle):
    """The HBM loop with 1, 2 or 4 systems per workgroup (one wavefront
    each, the last workgroup partly empty: 13 systems, two of them growing
    and run by their own launch) gives the same traces, states and
    calendars bit for bit, and every system matches the oracle."""
    states = small_states(140, 11)
    vrs = [0.0] * 11
    for k, n in ((0, 300), (1, 420)):
        states.insert(3 * k + 2, F.growth_state(n, 0.6, seed=900 + k))
        vrs.insert(3 * k + 2, 0.1)
    nev = 12000
    out = []
    for wpb in (1, 2, 4):
        b = gpu.Batch(states, trace_capacity=nev, per_system={"vr": vrs})
        b.set_residency(gpu.RES_HBM)
        b.set_loop_waves(wpb)
        assert b.loop_waves() == wpb
        b.run(max_events=nev // 3)
        b.run(max_events=nev)
        out.append(b)
    for b in out[1:]:
        assert_batches_equal(out[0], b, len(states), traces=False)
    # traces are drained when read: each batch's once
    tr = [[b.trace(s) for s in range(len(states))] for b in out]
    for s in range(len(states)):
        assert len(tr[0][s]) == nev
        for w in (1, 2):
            np.testing.assert_array_equal(tr[w][s], tr[0][s], err_msg="system %d" % s)
    for k in (0, 2, 5, 12):
        o = oracle.Oracle(states[k], vr=vrs[k])
        code, tr = o.run(nev)
        assert code == 0
        so, sb = o.state(), out[2].state(k)
        for f in ("x", "y", "vx", "vy", "t"):
            np.testing.assert_array_equal(bits(sb[f]), bits(so[f]), err_msg="%d %s" % (k, f))
    walled = gpu.Batch(states[:2], wall_y=1)
    with pytest.raises(gpu.EdmdError):
        walled.set_loop_waves(2)
    with pytest.raises(gpu.EdmdError):
        out[0].set_loop_waves(3)
