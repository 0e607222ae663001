"""GPU velocity-rescale thermostat (addNoise, EDMD.c:4828-4923) against the
reference's own golden vectors (tests/golden/rescale_*.npz, written by the
reference built strict through the trace harness's noise mode): a hot start
rescaled to a lower T, inelastic disks (restitution 0.9) whose every rescale
moves E, and a polydisperse growth start thermostatted while it grows.
Traces, final states, radii, counters and E after each rescale bit for bit."""
import numpy as np
import pytest

import edmd_fixtures as F
from conftest import RESCALED, bits, golden_state, load_golden, rescale_params

pytestmark = pytest.mark.gpu


def _ref_trace(gpu, g):
    nev = len(g["trace_t"])
    ref = np.zeros(nev, dtype=gpu.TRACE_DT)
    for k in ("type", "i", "j", "valid", "t"):
        ref[k] = g["trace_" + k]
    return ref


def _assert_trace(got, ref, growing):
    """Raw order, growing or not: pair events parked more than 5 time units
    ahead during a growth pop in the reference's orientation too."""
    for k in ("type", "i", "j", "valid"):
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
    np.testing.assert_array_equal(bits(got["t"]), bits(ref["t"]))


def _assert_state(st, rad, ob, g):
    for k in ("x", "y", "vx", "vy", "t"):
        np.testing.assert_array_equal(bits(st[k]), bits(g["state_" + k]), err_msg=k)
    np.testing.assert_array_equal(st["coll"], g["state_coll"].astype(np.uint32))
    np.testing.assert_array_equal(bits(rad), bits(g["state_rad"]))
    for k in ("t", "E", "px", "py"):
        assert bits(ob[k]) == bits(g["glob_" + k]), k
    assert bits(ob["vir_xx"]) == bits(g["glob_virXX"])
    assert bits(ob["vir_yy"]) == bits(g["glob_virYY"])
    assert ob["ncol"] == int(g["glob_ncol"]) and ob["ncross"] == int(g["glob_ncross"])


@pytest.mark.parametrize("name", RESCALED)
def test_rescale_matches_reference(gpu, name):
    g = load_golden(name)
    nev = len(g["trace_t"])
    s = gpu.System(golden_state(g), trace_capacity=nev, **rescale_params(g))
    s.run(max_events=nev)
    _assert_trace(s.trace(), _ref_trace(gpu, g), float(g["vr"]) > 0)
    _assert_state(s.state(), s.radii(), s.observables(), g)


@pytest.mark.parametrize("name", RESCALED)
def test_energy_after_each_rescale_matches_reference(gpu, name):
    """The loop stopped right after each of the first rescales: E (the
    ordered sum of the reference's physicalQ, EDMD.c:5968-5990) bit for bit
    with the reference's E after the same rescale."""
    g = load_golden(name)
    s = gpu.System(golden_state(g), **rescale_params(g))
    ref = g["rescale_E"]
    for k in range(8):
        _, kind = s.run_until_kind(1 << F.ADDINGNOISE, max_events=10 ** 6)
        assert kind == F.ADDINGNOISE
        assert bits(s.observables()["E"]) == bits(ref[k]), "rescale %d" % k


def test_rescale_batch_matches_reference(gpu):
    """The three thermostatted systems in one launch (growing and normal
    loops, different T, periods, restitution): each equals its golden
    vector."""
    gs = [load_golden(n) for n in RESCALED]
    nev = max(len(g["trace_t"]) for g in gs)
    per = {k: [rescale_params(g)[k] for g in gs] for k in ("dtnoise", "temperature", "vr", "res")}
    b = gpu.Batch([golden_state(g) for g in gs], trace_capacity=nev, noise=2, per_system=per)
    # every system runs the longest horizon; a shorter golden trace is the
    # prefix of its system's trace
    b.run(max_events=nev)
    for k, g in enumerate(gs):
        ref = _ref_trace(gpu, g)
        got = b.trace(k)[:len(ref)]
        _assert_trace(got, ref, float(g["vr"]) > 0)
