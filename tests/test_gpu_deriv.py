"""GPU derivatives of the event map (tape, tangent, adjoint, checkpointed
recomputation) against the oracle, and the adjoint against finite
differences through a torch.autograd.Function that exists only here.

Tapes, tangents and gradients are compared bit for bit with the oracle's:
the kernels perform the oracle's operations in the oracle's order.  Flat
walls (x, y and the mid wall) enter the tape as wall records.
"""
import numpy as np
import pytest

import edmd_fixtures as F
from conftest import bits, golden_state, load_golden

pytestmark = pytest.mark.gpu

TAPE_FIELDS = ("t", "ax", "ay", "cx", "cy", "ff", "i", "j", "g")
NCOLL = 10000


def states_mixed():
    out = [F.lattice_state(400, 0.60, seed=200 + k, lattice="fit") for k in range(6)]
    out.append(golden_state(load_golden("poly500_phi050")))
    out.append(F.lattice_state(1024, 0.80, seed=5, lattice="tri"))
    return out


def oracle_taped(oracle, st, ncoll, **params):
    o = oracle.Oracle(st, **params)
    s0 = o.state()
    o.tape_enable(ncoll + 16)
    o.run_collisions(ncoll)
    return o, s0, o.tape()


def tangent_inputs(states, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal((len(st["x"]), 4)) for st in states]


@pytest.fixture(scope="module")
def full_tape(gpu, oracle):
    states = states_mixed()
    b = gpu.Batch(states)
    b.deriv_begin(checkpoint_every=0, tape_capacity=NCOLL + 16)
    b.deriv_run(max_collisions=NCOLL)
    refs = [oracle_taped(oracle, st, NCOLL) for st in states]
    return states, b, refs


def test_tape_matches_oracle(full_tape):
    states, b, refs = full_tape
    for k, (o, s0, tape) in enumerate(refs):
        got = b.tape(k)  # raw order, mirrored pairs in the oracle's orientation
        assert len(got) == len(tape) == NCOLL, k
        for f in TAPE_FIELDS:
            np.testing.assert_array_equal(got[f], tape[f], err_msg="system %d %s" % (k, f))
        w, tl, nrec = b.deriv_values(k)
        assert nrec == NCOLL
        assert bits(w) == bits(o.tape_wsum())
        assert tl == tape["t"][-1]


def test_tangent_and_adjoint_match_oracle(full_tape, oracle):
    states, b, refs = full_tape
    tin = tangent_inputs(states, 1)
    tout, dw, dtl = b.tangent(tin)
    rng = np.random.default_rng(2)
    lw = rng.uniform(0.5, 2.0, len(states))
    lt = rng.uniform(-3.0, 3.0, len(states))
    lam = [rng.standard_normal((len(st["x"]), 4)) for st in states]
    grad = b.adjoint(lw, lt, lam)
    for k, (st, (o, s0, tape)) in enumerate(zip(states, refs)):
        te = o.t
        tan = tin[k].copy()
        rdw, rdtl = oracle.tangent(st["m"], s0["t"], tape, te, tan)
        np.testing.assert_array_equal(bits(tout[k]), bits(tan), err_msg="tangent %d" % k)
        assert bits(dw[k]) == bits(rdw) and bits(dtl[k]) == bits(rdtl), k
        g = oracle.adjoint(st["m"], s0["t"], tape, te, lw[k], lt[k], lam[k])
        np.testing.assert_array_equal(bits(grad[k]), bits(g), err_msg="adjoint %d" % k)


def test_context_adjoint_matches_oracle(gpu, oracle, tmp_path):
    """edmd_load_dump + edmd_adjoint_*: the single-system context path gives
    the oracle's tape values, tangent and gradient bit for bit, here through
    checkpointed recomputation (every 300 events)."""
    st = golden_state(load_golden("poly500_phi050"))
    dump = str(tmp_path / "init.dump")
    F.write_dump(dump, st)  # 17 significant digits: the state round-trips exactly
    s = gpu.System.from_dump(dump)
    ncoll = 3000
    s.adjoint_begin(checkpoint_every=300)
    s.adjoint_run(max_collisions=ncoll)
    o, s0, tape = oracle_taped(oracle, st, ncoll)
    w, tl, _ = s.adjoint_values()
    assert bits(w) == bits(o.tape_wsum()) and tl == tape["t"][-1]
    rng = np.random.default_rng(3)
    tin = rng.standard_normal((len(st["x"]), 4))
    tout, dw, dtl = s.adjoint_tangent(tin)
    tan = tin.copy()
    rdw, rdtl = oracle.tangent(st["m"], s0["t"], tape, o.t, tan)
    np.testing.assert_array_equal(bits(tout), bits(tan))
    assert bits(dw) == bits(rdw) and bits(dtl) == bits(rdtl)
    lam = rng.standard_normal((len(st["x"]), 4))
    g = s.adjoint_gradient(1.25, -0.75, lam)
    rg = oracle.adjoint(st["m"], s0["t"], tape, o.t, 1.25, -0.75, lam)
    np.testing.assert_array_equal(bits(g), bits(rg))
    s.adjoint_end()
    with pytest.raises(gpu.EdmdError):
        s.adjoint_values()


@pytest.mark.parametrize("every", [256, 97])
def test_checkpointed_recomputation_equals_full_tape(gpu, full_tape, every):
    """Checkpoint every K events, recompute each segment's tape in the
    passes: identical results, and the state is left as the run left it."""
    states, bfull, refs = full_tape
    b = gpu.Batch(states, trace_capacity=4000)
    b.deriv_begin(checkpoint_every=every)
    b.deriv_run(max_collisions=NCOLL)
    before = [b.state(k) for k in range(len(states))]
    tin = tangent_inputs(states, 1)
    t_seg, dw_seg, dtl_seg = b.tangent(tin)
    t_full, dw_full, dtl_full = bfull.tangent(tin)
    lam = [np.full((len(st["x"]), 4), 0.25) for st in states]
    g_seg = b.adjoint(1.5, -0.5, lam)
    g_full = bfull.adjoint(1.5, -0.5, lam)
    for k in range(len(states)):
        np.testing.assert_array_equal(bits(t_seg[k]), bits(t_full[k]))
        np.testing.assert_array_equal(bits(g_seg[k]), bits(g_full[k]))
        assert bits(dw_seg[k]) == bits(dw_full[k]) and bits(dtl_seg[k]) == bits(dtl_full[k])
        w, tl, _ = b.deriv_values(k)
        wf, tlf, _ = bfull.deriv_values(k)
        assert bits(w) == bits(wf) and tl == tlf
        after = b.state(k)
        for f in ("x", "y", "vx", "vy", "t", "coll"):
            np.testing.assert_array_equal(after[f], before[k][f], err_msg="%d %s" % (k, f))
    # the system goes on exactly as one that never took derivatives
    b.deriv_end()
    ref = gpu.Batch(states, trace_capacity=4000)
    ref.run(max_collisions=NCOLL)
    b.run(max_collisions=NCOLL + 2000)
    ref.run(max_collisions=NCOLL + 2000)
    for k in range(len(states)):
        sb, sr = b.state(k), ref.state(k)
        for f in ("x", "y", "vx", "vy", "t"):
            np.testing.assert_array_equal(bits(sb[f]), bits(sr[f]))


def test_adjoint_matches_finite_differences_through_autograd(gpu, oracle):
    """The time-averaged collision virial W / t_last over 600 collisions of
    N=400 at phi=0.60, as a torch.autograd.Function whose backward is the
    GPU adjoint (checkpoint every 256 events); checked against central
    differences (step 1e-7, extended precision) of the oracle's replay."""
    torch = pytest.importorskip("torch")
    st = F.lattice_state(400, 0.60, seed=31, lattice="fit")
    ncoll = 600

    class VirialAverage(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x0):
            s = dict(st)
            a = x0.detach().cpu().numpy()
            for c, key in enumerate(("x", "y", "vx", "vy")):
                s[key] = a[:, c].copy()
            b = gpu.Batch([s])
            b.deriv_begin(checkpoint_every=256)
            b.deriv_run(max_collisions=ncoll)
            W, T, _ = b.deriv_values(0)
            ctx.b, ctx.W, ctx.T = b, W, T
            return torch.tensor(W / T, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gJ):
            g = float(gJ)
            grad = ctx.b.adjoint(lw=g / ctx.T, lt=-g * ctx.W / ctx.T ** 2)[0]
            return torch.from_numpy(grad)

    x0 = torch.tensor(np.stack([st["x"], st["y"], st["vx"], st["vy"]], axis=1),
                      dtype=torch.float64, requires_grad=True)
    J = VirialAverage.apply(x0)
    J.backward()
    grad = x0.grad.numpy()

    o, s0, tape = oracle_taped(oracle, st, ncoll)
    W, T = o.tape_wsum(), tape["t"][-1]
    assert float(J.detach()) == W / T
    s0.update(rad=st["rad"], m=st["m"], Lx=st["Lx"], Ly=st["Ly"])
    rng = np.random.default_rng(8)
    for _ in range(3):
        e = rng.standard_normal((400, 4))
        fd = oracle.replay_fd(s0, tape, o.t, e, 1e-7, 1 / T, -W / T ** 2)
        ad = (grad * e).sum()
        assert abs(fd - ad) <= 1e-6 * abs(ad)


def test_tape_capacity_is_enforced(gpu):
    b = gpu.Batch([F.lattice_state(400, 0.6, seed=1, lattice="fit")])
    b.deriv_begin(checkpoint_every=0, tape_capacity=100)
    with pytest.raises(gpu.EdmdError) as e:
        b.deriv_run(max_collisions=1000)
    assert e.value.code == 8


def test_thermostatted_systems_are_rejected(gpu):
    b = gpu.Batch([F.lattice_state(400, 0.6, seed=1, lattice="fit")], noise=2)
    with pytest.raises(gpu.EdmdError) as e:
        b.deriv_begin()
    assert e.value.code == 22


# ---- flat walls on the tape

def _osmosis(n, L, seed):
    rng = np.random.default_rng(seed)
    xs, ys = [], []
    while len(xs) < n:
        x, y = rng.uniform(1.5, L - 1.5), rng.uniform(1.5, L - 1.5)
        if abs(x - L / 2) > 1 and all((x - a) ** 2 + (y - b) ** 2 > 1.2 ** 2
                                      for a, b in zip(xs, ys)):
            xs.append(x)
            ys.append(y)
    xs, ys = np.array(xs), np.array(ys)
    v = rng.standard_normal((n, 2))
    return {"x": xs, "y": ys, "vx": v[:, 0].copy(), "vy": v[:, 1].copy(),
            "rad": np.full(n, 0.5), "m": np.ones(n), "Lx": L, "Ly": L,
            "type": ((xs < L / 2) & (np.arange(n) % 2 == 0)).astype(np.int32)}


WALL_CASES = [
    ("channel", lambda: F.walled_state(400, 0.55, 21, wall_x=True), dict(wall_x=1), 4000),
    ("box", lambda: F.walled_state(400, 0.55, 22, wall_x=True, wall_y=True),
     dict(wall_x=1, wall_y=1, res_wall=0.9, res_tan=0.7), 4000),
    ("mid", lambda: _osmosis(160, 30.0, 7), dict(wall_x=1, wall_y=1, mid_wall=1, res_tan=1.0),
     3000),
    ("gravity box", lambda: F.walled_state(400, 0.55, 26, wall_x=True, wall_y=True),
     dict(wall_x=1, wall_y=1, add_field=1, field=-0.1, res_wall=0.9, res_tan=0.5), 4000),
    ("gravity channel", lambda: F.walled_state(400, 0.55, 27, wall_y=True),
     dict(wall_y=1, add_field=1, field=-0.2, res_tan=1.0), 4000),
    ("damped box", lambda: F.walled_state(400, 0.55, 28, wall_x=True, wall_y=True),
     dict(wall_x=1, wall_y=1, damping=1, gamma=0.01, res_wall=0.9, res_tan=0.5), 4000),
    ("damped periodic", lambda: F.lattice_state(400, 0.60, seed=29, lattice="fit"),
     dict(damping=1, gamma=0.02), 4000),
]


@pytest.mark.parametrize("name,make,params,ncoll", WALL_CASES)
def test_walled_tape_tangent_adjoint_match_oracle(gpu, oracle, name, make, params, ncoll):
    """Wall hits taped on the GPU as the oracle tapes them; tangent and
    gradient bit for bit, from the whole tape and with checkpoints every
    97 events (each segment's tape recomputed)."""
    st = make()
    o, s0, tape = oracle_taped(oracle, st, ncoll, **params)
    nwall = int(np.sum(tape["j"] < 0))
    assert nwall >= (20 if "wall_x" in params or "wall_y" in params else 0), (name, nwall)
    gam = params.get("gamma", 0.0) if params.get("damping") else 0.0
    rng = np.random.default_rng(41)
    tin = rng.standard_normal((len(st["x"]), 4))
    lam = rng.standard_normal((len(st["x"]), 4))
    tan = tin.copy()
    rdw, rdtl = oracle.tangent(st["m"], s0["t"], tape, o.t, tan, gamma=gam)
    rg = oracle.adjoint(st["m"], s0["t"], tape, o.t, 0.8, -1.5, lam, gamma=gam)
    for every in (0, 97):
        b = gpu.Batch([st], **params)
        b.deriv_begin(checkpoint_every=every, tape_capacity=ncoll + 16)
        b.deriv_run(max_collisions=ncoll)
        w, tl, nrec = b.deriv_values(0)
        assert bits(w) == bits(o.tape_wsum()), (name, every)
        assert tl == tape["t"][tape["j"] >= 0][-1], (name, every)
        if every == 0:
            assert nrec == len(tape)
            got, ref = b.tape(0), tape
            for f in TAPE_FIELDS:
                np.testing.assert_array_equal(got[f], ref[f], err_msg="%s %s" % (name, f))
        tout, dw, dtl = b.tangent([tin])
        np.testing.assert_array_equal(bits(tout[0]), bits(tan), err_msg="%s %d" % (name, every))
        assert bits(dw[0]) == bits(rdw) and bits(dtl[0]) == bits(rdtl), (name, every)
        g = b.adjoint(0.8, -1.5, [lam])
        np.testing.assert_array_equal(bits(g[0]), bits(rg), err_msg="%s %d" % (name, every))


@pytest.mark.parametrize("field,gamma", [(0.0, 0.0), (-0.1, 0.0), (0.0, 0.01)])
def test_walled_gradient_matches_finite_differences(gpu, oracle, field, gamma):
    """W / t_last over 600 events in a closed box with inelastic walls,
    plain, with gravity and damped: the GPU gradient against central
    differences of the oracle's replay."""
    st = F.walled_state(400, 0.55, 24, wall_x=True, wall_y=True)
    params = dict(wall_x=1, wall_y=1, res_wall=0.9, res_tan=0.7)
    if field:
        params.update(add_field=1, field=field)
    if gamma:
        params.update(damping=1, gamma=gamma)
    b = gpu.Batch([st], **params)
    b.deriv_begin(checkpoint_every=256)
    b.deriv_run(max_collisions=600)
    W, T, _ = b.deriv_values(0)
    grad = b.adjoint(lw=1 / T, lt=-W / T ** 2)[0]
    o, s0, tape = oracle_taped(oracle, st, 600, **params)
    assert np.sum(tape["j"] < 0) >= 8
    s0.update(rad=st["rad"], m=st["m"], Lx=st["Lx"], Ly=st["Ly"])
    rng = np.random.default_rng(6)
    for _ in range(3):
        e = rng.standard_normal((400, 4))
        fd = oracle.replay_fd(s0, tape, o.t, e, 1e-7, 1 / T, -W / T ** 2, field=field,
                              gamma=gamma)
        ad = (grad * e).sum()
        assert abs(fd - ad) <= 1e-6 * abs(ad)


def test_unsupported_models_are_rejected_for_taping(gpu):
    st = F.walled_state(400, 0.55, 25, wall_x=True, wall_y=True)
    for kw in (dict(add_field=1, damping=1, gamma=0.01), dict(thermal_wall=1),
               dict(wall_y=1, shear_wall=1)):
        b = gpu.Batch([st], wall_x=1, **kw)
        with pytest.raises(gpu.EdmdError) as e:
            b.deriv_begin()
        assert e.value.code == 22
