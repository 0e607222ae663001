"""GPU growth start (the reference's default start: disks grow from radius 0
until GROWSTOP, then stopGrow normalises and the Normal dynamics begin)
against the reference's golden vectors and the oracle, bit for bit."""
import numpy as np
import pytest

import edmd_fixtures as F
from conftest import GROWN, bits, golden_state, load_golden

pytestmark = pytest.mark.gpu


def assert_trace_equal(got, ref):
    """In raw order, including the orientation of mirrored pair events
    scheduled more than 5 time units ahead (frequent at the start of a
    growth, when the disks are small), which the reference's overflow list
    reorders when its bucket ring wraps (EDMD.c:2287-2307)."""
    a, b = got, ref
    for k in ("type", "i", "j", "valid"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    np.testing.assert_array_equal(bits(a["t"]), bits(b["t"]))


@pytest.mark.parametrize("name", GROWN)
def test_growth_matches_reference(gpu, name):
    g = load_golden(name)
    nev = len(g["trace_t"])
    s = gpu.System(golden_state(g), vr=float(g["vr"]), trace_capacity=nev)
    s.run(max_events=nev)
    ref = np.zeros(nev, dtype=gpu.TRACE_DT)
    for k in ("type", "i", "j", "valid", "t"):
        ref[k] = g["trace_" + k]
    assert_trace_equal(s.trace(), ref)
    st = s.state()
    for k in ("x", "y", "vx", "vy", "t"):
        np.testing.assert_array_equal(bits(st[k]), bits(g["state_" + k]), err_msg=k)
    np.testing.assert_array_equal(st["coll"], g["state_coll"].astype(np.uint32))
    np.testing.assert_array_equal(bits(s.radii()), bits(g["state_rad"]))
    ob = s.observables()
    for k in ("t", "E", "px", "py"):
        assert bits(ob[k]) == bits(g["glob_" + k]), k
    assert ob["ncol"] == int(g["glob_ncol"]) and ob["ncross"] == int(g["glob_ncross"])


def test_mixed_batch_growing_and_normal_match_oracle(gpu, oracle):
    """Growing and non-growing systems in one batch; growth rates differ so
    GROWSTOP falls at different times; runs cross GROWSTOP in the middle of
    a call and continue in the Normal loop."""
    cases = [(300, 0.55, 11, 0.0, 0.2), (256, 0.72, 12, 0.1, 0.0125), (400, 0.6, 13, 0.0, 0.0),
             (150, 0.5, 14, 0.1, 0.1), (500, 0.65, 15, 0.1, 0.05)]
    states, vrs = [], []
    for n, phi, seed, poly, vr in cases:
        if vr > 0:
            states.append(F.growth_state(n, phi, seed, poly=poly))
        else:
            states.append(F.lattice_state(n, phi, seed, lattice="fit"))
        vrs.append(vr)
    nev = 40000
    b = gpu.Batch(states, trace_capacity=nev, per_system={"vr": vrs})
    b.run(max_events=nev // 2)
    b.run(max_events=nev)
    for k, (st, vr) in enumerate(zip(states, vrs)):
        o = oracle.Oracle(st, vr=vr)
        code, tr = o.run(nev)
        assert code == 0
        assert_trace_equal(b.trace(k), tr)
        so, sb = o.state(), b.state(k)
        for f in ("x", "y", "vx", "vy", "t"):
            np.testing.assert_array_equal(bits(sb[f]), bits(so[f]), err_msg="%d %s" % (k, f))
        np.testing.assert_array_equal(bits(b.radii(k)), bits(o.radii()))
        assert b.observables(k)["ncol"] == o.observables()["ncol"]


def test_growth_under_thermostat_matches_oracle(gpu, oracle):
    """Velocity rescaling during the growth uses the growth free flight and
    growth predictions (addNoise through the dispatch pointers)."""
    st = F.growth_state(400, 0.7, seed=21, poly=0.1)
    kw = dict(vr=0.05, noise=2, dtnoise=0.5, temperature=1.0)
    nev = 50000
    s = gpu.System(st, trace_capacity=nev, **kw)
    s.run(max_events=nev)
    o = oracle.Oracle(st, **kw)
    code, tr = o.run(nev)
    assert code == 0
    assert_trace_equal(s.trace(), tr)
    so, sg = o.state(), s.state()
    for f in ("x", "y", "vx", "vy", "t"):
        np.testing.assert_array_equal(bits(sg[f]), bits(so[f]), err_msg=f)
    np.testing.assert_array_equal(bits(s.radii()), bits(o.radii()))


def test_growth_blocks_derivatives(gpu):
    b = gpu.Batch([F.growth_state(200, 0.5, seed=2)], vr=0.1)
    with pytest.raises(gpu.EdmdError) as e:
        b.deriv_begin()
    assert e.value.code == 22
    b.run(max_events=200000, tmax=11.0)  # past GROWSTOP at t = 10
    b.deriv_begin()


def test_cli_default_start_is_the_growth_start(gpu, tmp_path):
    """`edmd -N 300 -p 0.6 -v 7 -t 11.5 -D 1 --full-precision` without -l or
    --lattice starts as the reference does by default: a binary mixture
    (30% of radius 0.4, the rest radius 1) at random centres, growing at the
    rate optimizeGrowConstant picks (0.1 at phi = 0.6, so GROWSTOP at t = 10);
    screenshot times below 1/vr move by 1/vr, so the one frame is at t = 11.
    It equals the library's run from the disks particlesInit makes for the
    same command line (the reference's MT19937 start).  Both are seeded by
    -v: without it the reference seeds from the clock's seconds (EDMD.c:500),
    and the two processes could read different seconds."""
    import ctypes as C
    from test_start_reference import start
    import os
    import subprocess
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    cli = os.path.join(repo, "graphical-edmd_amd", "bin", "edmd")
    r = subprocess.run([cli, "-N", "300", "-p", "0.6", "-v", "7", "-t", "11.5", "-D", "1",
                        "--full-precision"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    dumps = [f for f in os.listdir(tmp_path / "dump") if f.endswith(".dump")]
    assert len(dumps) == 1
    frame = F.read_dump(str(tmp_path / "dump" / dumps[0]))
    assert gpu.growth_rate(0.6) == 0.1
    L = gpu.lib()
    assert start(L, ["-N", "300", "-p", "0.6", "-v", "7", "--no-output"]) == 0
    st = {k: np.zeros(300) for k in ("x", "y", "vx", "vy", "rad", "m", "vr")}
    L.edmd_compat_particles.argtypes = [C.c_void_p] * 7
    assert L.edmd_compat_particles(*[st[k].ctypes.data for k in
                                     ("x", "y", "vx", "vy", "rad", "m", "vr")]) == 0
    st["Lx"] = C.c_double.in_dll(L, "Lx").value
    st["Ly"] = C.c_double.in_dll(L, "Ly").value
    L.freeArrays()
    np.testing.assert_array_equal(st["vr"], 0.1 * st["rad"])
    del st["vr"]
    s = gpu.System(st, vr=0.1, first_screen=1.0, dtime=1.0, first_thermo=1000.0,
                   dtime_thermo=100.0, einit=1.0)
    reason, kind = s.run_until_kind(1 << 2)  # stop right after the screenshot
    assert kind == 2
    so = s.state()
    assert s.observables()["t"] == 11.0
    for k in ("x", "y", "vx", "vy"):
        np.testing.assert_array_equal(bits(frame[k]), bits(so[k]), err_msg=k)
    np.testing.assert_array_equal(frame["rad"], st["rad"])
    assert np.sum(st["rad"] == 0.4) == 90


def test_sweep_after_monodisperse_growth_matches_oracle(gpu, oracle):
    """A monodisperse growth start ends with radii that differ in their last
    bits (each disk's radius is advanced by its own flights), so the sweep
    after GROWSTOP must use each pair's own contact term DIST2(r_i, r_j)
    (EDMD.c:35-43), not the owner's 4 r_i r_i.  The box has interior tiles
    for every tile shape; every disk is compared."""
    st = F.growth_state(3000, 0.5, seed=31)
    vr = 0.1
    o = oracle.Oracle(st, vr=vr)
    code, done = o.run_until(10.5)
    assert code == 0 and o.t > 10.0
    assert len(np.unique(o.radii())) > 1
    o.sync_all(o.t)
    from test_gpu_parity import assert_sweep_equals_oracle
    for v in (1, 5):
        b = gpu.Batch([st], vr=vr)
        b.run(max_events=done)
        assert b.observables(0)["nevents"] == done
        b.set_sweep_shape(v)
        b.sweep_predict()
        sb, so = b.state(0), o.state()
        for k in ("x", "y", "vx", "vy"):
            np.testing.assert_array_equal(bits(sb[k]), bits(so[k]), err_msg=k)
        np.testing.assert_array_equal(bits(b.radii(0)), bits(o.radii()))
        assert_sweep_equals_oracle(b.events(0), o, "shape %d" % v)
