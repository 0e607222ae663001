"""The replica workloads at their own scale on the GPU.

* 4096 replicas in one launch, N ~ U{256..2048}, phi ~ U[0.50, 0.75], the
  fitted triangular lattice start of the batched benchmark: every 64th
  replica against the oracle bit for bit (trace, state), every replica's
  energy, momentum, counters and calendar.
* 64 replicas of N = 65536 polydisperse disks (diameters N(1, 0.1)
  truncated at +-3 sigma) at phi = 0.80, growth start with the velocity-
  rescale thermostat to T = 1 every 0.1: two replicas against the oracle
  (traces in raw order, states and radii bitwise) over their first
  events, then all 64 through GROWSTOP and two more rescales, with the
  thermostat's E/N, the momentum stopGrow removed and every calendar.
* 512 replicas of N = 400 at phi = 0.60, 10^4 collisions each, taped with a
  checkpoint every 256 events: three replicas' virial sums, tangents and
  gradients against the oracle bit for bit, and for every replica the
  adjoint/tangent duality <grad, e> = lw dW + lt dt_last and a positive
  finite-time growth rate.
"""
import numpy as np
import pytest

import edmd_fixtures as F
from calendar_check import check_calendar
from conftest import bits

pytestmark = pytest.mark.gpu


def replica_params(nrep, seed):
    """The batched workload's sizes and packing fractions."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(256, 2049, size=nrep)
    phis = rng.uniform(0.50, 0.75, size=nrep)
    return ns, phis


def _assert_state(sg, so, label):
    for k in ("x", "y", "vx", "vy", "t"):
        np.testing.assert_array_equal(bits(sg[k]), bits(so[k]), err_msg="%s %s" % (label, k))
    np.testing.assert_array_equal(sg["coll"], so["coll"].astype(np.uint32), err_msg=label)
    np.testing.assert_array_equal(sg["cx"], so["cx"], err_msg=label)
    np.testing.assert_array_equal(sg["cy"], so["cy"], err_msg=label)


def test_4096_replicas_one_launch(gpu, oracle):
    nrep, nev, seed = 4096, 10000, 1000
    ns, phis = replica_params(nrep, seed)
    states = [gpu.init_lattice(int(n), float(p), rad=0.5, lattice=2, seed=seed + k)
              for k, (n, p) in enumerate(zip(ns, phis))]
    checked = list(range(0, nrep, 64))
    cap = [nev if k % 64 == 0 else 0 for k in range(nrep)]
    b = gpu.Batch(states, per_system={"trace_capacity": cap})
    e0 = [0.5 * float(np.sum(st["m"] * (st["vx"] ** 2 + st["vy"] ** 2))) for st in states]
    b.run(max_events=nev)
    for k in checked:
        o = oracle.Oracle(states[k])
        code, tr = o.run(nev)
        assert code == 0
        gt = b.trace(k)
        for f in ("type", "i", "j", "valid"):
            np.testing.assert_array_equal(gt[f], tr[f], err_msg="replica %d %s" % (k, f))
        np.testing.assert_array_equal(bits(gt["t"]), bits(tr["t"]), err_msg="replica %d" % k)
        _assert_state(b.state(k), o.state(), "replica %d" % k)
        ob, oo = b.observables(k), o.observables()
        assert ob["ncol"] == oo["ncol"] and ob["ncross"] == oo["ncross"]
        assert bits(ob["vir_xx"]) == bits(oo["virXX"]) and bits(ob["vir_yy"]) == bits(oo["virYY"])
    for k in range(nrep):
        ob = b.observables(k)
        n = int(ns[k])
        assert ob["status"] == 0 and ob["nevents"] == nev, k
        assert 0.3 * nev < ob["ncol"] < nev, k
        assert abs(ob["E"] - e0[k]) <= 1e-10 * e0[k], k
        assert abs(ob["px"]) < 1e-9 * n and abs(ob["py"]) < 1e-9 * n, k
        bad = check_calendar(b.calendar(k), b.events(k))
        assert bad == [], (k, bad[:3])


def test_dense_polydisperse_growth_thermostat_64x65536(gpu, oracle):
    nrep, n, vr, seed = 64, 65536, 0.2, 1000
    kw = dict(vr=vr, noise=2, dtnoise=0.1, temperature=1.0, einit=1.0)
    states = [gpu.init_random(n, 0.80, rad=0.5, poly=0.1, seed=seed + 7919 * (k + 1),
                              e_per=0.05) for k in range(nrep)]
    for st in states:
        d = 2 * st["rad"]
        assert d.min() >= 0.7 and d.max() <= 1.3
    checked = (0, 37)
    nev1 = 1200000
    cap = [nev1 if k in checked else 0 for k in range(nrep)]
    b = gpu.Batch(states, per_system={"trace_capacity": cap}, **kw)
    # first: the growing, thermostatted dynamics event by event
    b.run(max_events=nev1)
    for k in checked:
        o = oracle.Oracle(states[k], **kw)
        code, tr = o.run(nev1)
        assert code == 0
        assert np.sum(tr["type"] == F.ADDINGNOISE) >= 3
        ct, cg = tr, b.trace(k)
        for f in ("type", "i", "j", "valid"):
            np.testing.assert_array_equal(cg[f], ct[f], err_msg="replica %d %s" % (k, f))
        np.testing.assert_array_equal(bits(cg["t"]), bits(ct["t"]), err_msg="replica %d" % k)
        _assert_state(b.state(k), o.state(), "replica %d" % k)
        np.testing.assert_array_equal(bits(b.radii(k)), bits(o.radii()))
    # then everyone through GROWSTOP (t = 1/vr = 5) and the rescales at 5.1, 5.2
    b.run(tmax=5.25)
    for k in range(nrep):
        ob = b.observables(k)
        assert ob["status"] == 0, k
        assert 5.2 <= ob["t"] <= 5.3, k
        assert ob["ncol"] > 0, k  # counted afresh from GROWSTOP
        assert abs(ob["E"] / n - 1.0) < 1e-10, (k, ob["E"] / n)
        assert abs(ob["px"]) < 1e-9 * n and abs(ob["py"]) < 1e-9 * n, k
        # the radii reached their targets through each disk's own flights
        np.testing.assert_allclose(b.radii(k), states[k]["rad"], rtol=1e-12, err_msg="%d" % k)
        bad = check_calendar(b.calendar(k), b.events(k))
        assert bad == [], (k, bad[:3])


def test_adjoint_batch_512x400(gpu, oracle):
    nrep, n, ncoll, every, seed = 512, 400, 10000, 256, 1000
    states = [gpu.init_lattice(n, 0.60, rad=0.5, lattice=2, seed=seed + k) for k in range(nrep)]
    rng = np.random.default_rng(seed)
    tin = [rng.standard_normal((n, 4)) for _ in range(nrep)]
    b = gpu.Batch(states)
    b.deriv_begin(checkpoint_every=every)
    b.deriv_run(max_collisions=ncoll)
    vals = [b.deriv_values(k) for k in range(nrep)]
    W = np.array([v[0] for v in vals])
    T = np.array([v[1] for v in vals])
    lw, lt = 1.0 / T, -W / T ** 2
    grad = b.adjoint(lw=lw, lt=lt)
    tout, dw, dtl = b.tangent(tin)
    for k in (0, 255, 511):
        o = oracle.Oracle(states[k])
        s0 = o.state()
        o.tape_enable(ncoll + 16)
        o.run_collisions(ncoll)
        tape = o.tape()
        assert bits(W[k]) == bits(o.tape_wsum()) and T[k] == tape["t"][-1], k
        tan = tin[k].copy()
        rdw, rdtl = oracle.tangent(states[k]["m"], s0["t"], tape, o.t, tan)
        np.testing.assert_array_equal(bits(tout[k]), bits(tan), err_msg="tangent %d" % k)
        assert bits(dw[k]) == bits(rdw) and bits(dtl[k]) == bits(rdtl), k
        g = oracle.adjoint(states[k]["m"], s0["t"], tape, o.t, lw[k], lt[k])
        np.testing.assert_array_equal(bits(grad[k]), bits(g), err_msg="adjoint %d" % k)
    for k in range(nrep):
        assert np.all(np.isfinite(grad[k])) and np.all(np.isfinite(tout[k])), k
        lhs = float((grad[k] * tin[k]).sum())
        rhs = lw[k] * dw[k] + lt[k] * dtl[k]
        assert abs(lhs - rhs) <= 1e-10 * max(abs(lhs), abs(lw[k] * dw[k])), k
        growth = np.log(np.linalg.norm(tout[k]) / np.linalg.norm(tin[k])) / T[k]
        assert 1.0 < growth < 20.0, (k, growth)
