"""The oracle against the reference's own outputs (tests/golden/*.npz,
written by tests/golden/make_golden.py from the reference built strict).

Bit-exact: sweep slots, event traces (kind, i, j, valid, time, including the
orientation of mirrored pairs), final states, virial sums and counters."""
import numpy as np
import pytest

from conftest import (GROWN, RESCALED, SWEEP_ONLY, TRACED, bits, golden_state, load_golden,
                      rescale_params)
import edmd_fixtures as F


@pytest.mark.parametrize("name", TRACED + SWEEP_ONLY)
def test_sweep_matches_reference(oracle, name):
    g = load_golden(name)
    o = oracle.Oracle(golden_state(g))
    ncx, ncy, _, _ = o.grid()
    assert (ncx, ncy) == tuple(g["grid"])
    ev = o.events()
    np.testing.assert_array_equal(bits(ev["t_cross"]), bits(g["sweep_t_cross"]))
    np.testing.assert_array_equal(bits(ev["t_coll"]), bits(g["sweep_t_coll"]))
    np.testing.assert_array_equal(ev["dir"], g["sweep_dir"])
    np.testing.assert_array_equal(ev["partner"], g["sweep_partner"])
    np.testing.assert_array_equal(ev["coll_actual"], g["sweep_coll_actual"])
    s = o.state()
    np.testing.assert_array_equal(s["cx"], g["sweep_cx"])
    np.testing.assert_array_equal(s["cy"], g["sweep_cy"])


@pytest.mark.parametrize("name", TRACED)
def test_trace_matches_reference(oracle, name):
    g = load_golden(name)
    nev = len(g["trace_t"])
    o = oracle.Oracle(golden_state(g))
    st, tr = o.run(nev)
    assert st == 0
    for k in ("type", "i", "j", "valid"):
        np.testing.assert_array_equal(tr[k], g["trace_" + k], err_msg=k)
    np.testing.assert_array_equal(bits(tr["t"]), bits(g["trace_t"]))
    s = o.state()
    for k in ("x", "y", "vx", "vy", "t"):
        np.testing.assert_array_equal(bits(s[k]), bits(g["state_" + k]), err_msg=k)
    np.testing.assert_array_equal(s["coll"], g["state_coll"])
    np.testing.assert_array_equal(s["cx"], g["state_cx"])
    np.testing.assert_array_equal(s["cy"], g["state_cy"])
    ob = o.observables()
    for k in ("t", "virXX", "virYY", "virXY", "virYX", "E", "px", "py"):
        assert bits(ob[k]) == bits(g["glob_" + k]), k
    assert ob["ncol"] == int(g["glob_ncol"])
    assert ob["ncross"] == int(g["glob_ncross"])


def test_thermo_run_matches_reference(oracle):
    """Main's loop bound (t <= tmax) through 13 screenshots (free flight of all
    disks, EDMD.c:4659-4661) and 4 thermo events: final state bit-exact, the
    reference's printed thermo rows reproduced to their printed digits."""
    g = load_golden("thermo64")
    o = oracle.Oracle(golden_state(g), dtime_thermo=float(g["dtime_thermo"]))
    st, _ = o.run_until(float(g["tmax"]))
    assert st == 0
    s = o.state()
    for k in ("x", "y", "vx", "vy", "t"):
        np.testing.assert_array_equal(bits(s[k]), bits(g["state_" + k]), err_msg=k)
    ob = o.observables()
    assert ob["ncol"] == int(g["glob_ncol"])
    rows = o.thermo_rows()
    ref = g["thermo_rows"]
    assert len(rows) == len(ref)
    for r, q in zip(rows, ref):
        got = [r["t"], r["ncol"], r["e_per_n"], r["pressure"], r["pxx"], r["pyy"], r["pxy"],
               r["pyx"]]
        # %lf prints 6 decimals (pxy, pyx: %.10lf)
        np.testing.assert_allclose(got[:6], q[:6], rtol=0, atol=6e-7)
        np.testing.assert_allclose(got[6:], q[6:], rtol=0, atol=6e-11)


def test_pcf_matches_reference(oracle):
    g = load_golden("sq2000_phi070")
    r, gr = oracle.pcf(g["x"], g["y"], float(g["Lx"]), float(g["Ly"]), 0.1, float(g["Lx"]) / 2)
    np.testing.assert_array_equal(bits(r), bits(g["pcf_r"]))
    np.testing.assert_array_equal(bits(gr), bits(g["pcf_g"]))


def test_calendar_shape_does_not_change_pop_order(oracle):
    """The bucket count only changes how the calendar stores events, never
    which event pops next: a 3-bucket calendar gives the same trace."""
    g = load_golden("n400_phi060")
    nev = len(g["trace_t"])
    o = oracle.Oracle(golden_state(g), bucket_count=3)
    st, tr = o.run(nev)
    assert st == 0
    np.testing.assert_array_equal(bits(tr["t"]), bits(g["trace_t"]))
    lo = np.minimum(tr["i"], tr["j"])
    glo = np.minimum(g["trace_i"], g["trace_j"])
    np.testing.assert_array_equal(lo, glo)


@pytest.mark.parametrize("name", GROWN)
def test_growth_matches_reference(oracle, name):
    """Growth start from radius 0 (monodisperse and polydisperse targets):
    the trace through GROWSTOP and stopGrow's re-prediction, the final state
    with every radius, the counters stopGrow resets."""
    g = load_golden(name)
    nev = len(g["trace_t"])
    o = oracle.Oracle(golden_state(g), vr=float(g["vr"]))
    st, tr = o.run(nev)
    assert st == 0
    ref = np.zeros(nev, dtype=tr.dtype)
    for k in ("type", "i", "j", "valid", "t"):
        ref[k] = g["trace_" + k]
    ct, cr = tr, ref
    for k in ("type", "i", "j", "valid"):
        np.testing.assert_array_equal(ct[k], cr[k], err_msg=k)
    np.testing.assert_array_equal(bits(ct["t"]), bits(cr["t"]))
    assert (tr["type"] == F.GROWSTOP).sum() == 1
    s = o.state()
    for k in ("x", "y", "vx", "vy", "t"):
        np.testing.assert_array_equal(bits(s[k]), bits(g["state_" + k]), err_msg=k)
    np.testing.assert_array_equal(s["coll"], g["state_coll"])
    np.testing.assert_array_equal(bits(o.radii()), bits(g["state_rad"]))
    ob = o.observables()
    for k in ("t", "E", "px", "py"):
        assert bits(ob[k]) == bits(g["glob_" + k]), k
    assert ob["ncol"] == int(g["glob_ncol"])
    assert ob["ncross"] == int(g["glob_ncross"])


@pytest.mark.parametrize("name", RESCALED)
def test_rescale_thermostat_matches_reference(oracle, name):
    """The velocity-rescale thermostat (addNoise, EDMD.c:4828-4923) against
    the reference itself: a hot start rescaled to a lower T, inelastic disks
    whose every rescale moves E, and a growth start thermostatted while it
    grows.  Raw trace order, every rescale's E, final state and radii bit
    for bit."""
    g = load_golden(name)
    nev = len(g["trace_t"])
    o = oracle.Oracle(golden_state(g), **rescale_params(g))
    st, tr = o.run(nev)
    assert st == 0
    ref = np.zeros(nev, dtype=tr.dtype)
    for k in ("type", "i", "j", "valid", "t"):
        ref[k] = g["trace_" + k]
    for k in ("type", "i", "j", "valid"):
        np.testing.assert_array_equal(tr[k], ref[k], err_msg=k)
    np.testing.assert_array_equal(bits(tr["t"]), bits(ref["t"]))
    s = o.state()
    for k in ("x", "y", "vx", "vy", "t"):
        np.testing.assert_array_equal(bits(s[k]), bits(g["state_" + k]), err_msg=k)
    np.testing.assert_array_equal(s["coll"], g["state_coll"])
    np.testing.assert_array_equal(bits(o.radii()), bits(g["state_rad"]))
    ob = o.observables()
    for k in ("t", "E", "px", "py", "virXX", "virYY"):
        assert bits(ob[k]) == bits(g["glob_" + k]), k
    assert ob["ncol"] == int(g["glob_ncol"])
    assert ob["ncross"] == int(g["glob_ncross"])
    # E/N = T right after every rescale, in the reference's own arithmetic
    e = g["rescale_E"] / len(g["x"])
    np.testing.assert_allclose(e, float(g["temperature"]), rtol=1e-13)
