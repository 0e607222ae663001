"""Pins of the oracle's derivatives of the event map (tangent, adjoint and
the fixed-sequence replay they differentiate).

The reference has no derivatives, so these checks come from the
mathematics: closed forms for a head-on collision, central finite
differences (step 1e-7) of the nonlinear fixed-sequence replay, the
adjoint/tangent duality, and the invariants the collision rule keeps
(momentum, energy, centre of mass) carried over to the linearised map.
"""
import numpy as np
import pytest

import oracle_bind as ob
from edmd_fixtures import lattice_state

KEYS = ("x", "y", "vx", "vy")


def two_disks(dx=2.0, v=1.0, L=20.0, dy=0.0):
    st = {"x": np.array([5.0, 5.0 + dx]), "y": np.array([5.0, 5.0 + dy]),
          "vx": np.array([v, -v]), "vy": np.array([0.0, 0.0]),
          "rad": np.array([0.5, 0.5]), "m": np.array([1.0, 1.0]), "Lx": L, "Ly": L}
    return st


def taped(st, ncoll, cap=None, **params):
    o = ob.Oracle(st, **params)
    s0 = o.state()
    s0.update(rad=st["rad"], m=st["m"], Lx=st["Lx"], Ly=st["Ly"])
    o.tape_enable(cap or ncoll + 10)
    o.run_collisions(ncoll)
    return o, s0, o.tape()


def perturbed(s0, e, h):
    s = dict(s0)
    for c, k in enumerate(KEYS):
        s[k] = s0[k] + h * e[:, c]
    return s


def test_head_on_closed_form():
    st = two_disks()
    o, s0, tape = taped(st, 1)
    assert len(tape) == 1 and tape["t"][0] == 0.5
    T = 1.0
    o.sync_all(T)
    # W = m m ff |a|^2 = (a.c) for unit masses and diameter: -2
    assert o.tape_wsum() == -2.0
    m, tl0 = st["m"], s0["t"]
    # d W / d(state): only the normal relative velocity enters
    g = ob.adjoint(m, tl0, tape, T, 1.0)
    np.testing.assert_allclose(g, [[0, 0, -1, 0], [0, 0, 1, 0]], atol=1e-15)
    # d t_c / d(state): t_c = (x1 - x0 - 1) / (v0 - v1)
    g = ob.adjoint(m, tl0, tape, T, 0.0, 1.0)
    np.testing.assert_allclose(g, [[-0.5, 0, -0.25, 0], [0.5, 0, 0.25, 0]], atol=1e-15)
    # final x0 = x1 - 1 + v1 T: independent of x0 and v0
    lam = np.zeros((2, 4))
    lam[0, 0] = 1.0
    g = ob.adjoint(m, tl0, tape, T, 0.0, 0.0, lam)
    np.testing.assert_allclose(g, [[0, 0, 0, 0], [1, 0, T, 0]], atol=1e-15)
    # final v0 = v1
    lam = np.zeros((2, 4))
    lam[0, 2] = 1.0
    g = ob.adjoint(m, tl0, tape, T, 0.0, 0.0, lam)
    np.testing.assert_allclose(g, [[0, 0, 0, 0], [0, 0, 1, 0]], atol=1e-15)
    # tangent: moving disk 0 forward by 1 makes the collision 0.5 earlier
    tan = np.zeros((2, 4))
    tan[0, 0] = 1.0
    dW, dtl = ob.tangent(m, tl0, tape, T, tan)
    assert dW == 0.0 and dtl == -0.5
    np.testing.assert_allclose(tan, [[0, 0, 0, 0], [1, 0, 0, 0]], atol=1e-15)


def test_oblique_closed_form():
    # impact parameter b: W = (a.c) = -2 sqrt(1 - b^2), dW/db = 2 b / sqrt(1 - b^2)
    b = 0.3
    st = two_disks(dy=b)
    o, s0, tape = taped(st, 1)
    g = ob.adjoint(st["m"], s0["t"], tape, 1.0, 1.0)
    dWdb = 2 * b / np.sqrt(1 - b * b)
    np.testing.assert_allclose(g[1, 1], dWdb, rtol=1e-13)
    np.testing.assert_allclose(g[0, 1], -dWdb, rtol=1e-13)
    np.testing.assert_allclose(o.tape_wsum(), -2 * np.sqrt(1 - b * b), rtol=1e-14)


@pytest.fixture(scope="module")
def dense400():
    st = lattice_state(400, 0.60, 7, lattice="fit")
    o, s0, tape = taped(st, 800)
    return st, o, s0, tape


def test_replay_reproduces_engine(dense400):
    st, o, s0, tape = dense400
    te = o.t
    S, tk, fin = ob.replay(s0, tape, te)
    assert tk == pytest.approx(tape["t"][-1], abs=1e-12)
    assert S == pytest.approx(o.tape_wsum(), rel=1e-12)
    o2 = ob.Oracle(st)
    o2.run_collisions(800)
    o2.sync_all(te)
    ref = o2.state()
    d = fin["x"] - ref["x"]
    d -= st["Lx"] * np.round(d / st["Lx"])
    assert np.abs(d).max() < 1e-11
    assert np.abs(fin["vx"] - ref["vx"]).max() < 1e-11


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_adjoint_matches_finite_differences(dense400, seed):
    """Time-averaged collision virial W / t_last plus a random linear
    functional of the final state; central differences with step 1e-7 of
    the replay (in extended precision), linearised at the base point:
    d(W / t_last) = dW / t_last - W dt_last / t_last^2."""
    st, o, s0, tape = dense400
    te = o.t
    S, tk, _ = ob.replay(s0, tape, te)
    rng = np.random.default_rng(seed)
    lam = rng.standard_normal((400, 4))
    lw, lt = 1 / tk, -S / tk ** 2
    g = ob.adjoint(st["m"], s0["t"], tape, te, lw, lt, lam)
    e = rng.standard_normal((400, 4))
    fd = ob.replay_fd(s0, tape, te, e, 1e-7, lw, lt, lam)
    ad = (g * e).sum()
    assert abs(fd - ad) <= 1e-6 * abs(ad)


def test_gradient_components_by_finite_differences(dense400):
    """Single components of dW/d(state) for a few disks, each by FD."""
    st, o, s0, tape = dense400
    te = o.t
    g = ob.adjoint(st["m"], s0["t"], tape, te, 1.0)
    worst = 0.0
    for p in (0, 17, 203, 399):
        for c in range(4):
            e = np.zeros((400, 4))
            e[p, c] = 1.0
            fd = ob.replay_fd(s0, tape, te, e)
            worst = max(worst, abs(fd - g[p, c]) / abs(g[p, c]))
    assert worst <= 1e-6


def test_polydisperse_finite_differences():
    """Unequal masses and radii: a swapped m_i / m_j would show here."""
    from conftest import golden_state, load_golden
    st = golden_state(load_golden("poly500_phi050"))
    o, s0, tape = taped(st, 600)
    te = o.t
    rng = np.random.default_rng(5)
    lam = rng.standard_normal((len(st["x"]), 4))
    g = ob.adjoint(st["m"], s0["t"], tape, te, 1.0, 0.7, lam)
    for _ in range(3):
        e = rng.standard_normal((len(st["x"]), 4))
        fd = ob.replay_fd(s0, tape, te, e, 1e-7, 1.0, 0.7, lam)
        assert abs(fd - (g * e).sum()) <= 1e-6 * abs((g * e).sum())


def test_adjoint_tangent_duality_long_horizon():
    """10^4 collisions at N=400, phi=0.60: finite differences lose meaning
    (the tangent grows ~1e11), but <grad, e> must equal the tangent."""
    st = lattice_state(400, 0.60, 11, lattice="fit")
    o, s0, tape = taped(st, 10000)
    te = o.t
    rng = np.random.default_rng(9)
    lam = rng.standard_normal((400, 4))
    g = ob.adjoint(st["m"], s0["t"], tape, te, 0.3, -2.0, lam)
    for _ in range(2):
        e = rng.standard_normal((400, 4))
        tan = e.copy()
        dW, dtl = ob.tangent(st["m"], s0["t"], tape, te, tan)
        lhs = (g * e).sum()
        rhs = 0.3 * dW - 2.0 * dtl + (lam * tan).sum()
        assert abs(lhs - rhs) <= 1e-11 * abs(lhs)
        assert np.linalg.norm(tan) > 1e6 * np.linalg.norm(e)


def test_tangent_conserves_linearised_invariants(dense400):
    st, o, s0, tape = dense400
    te = o.t
    m = st["m"]
    rng = np.random.default_rng(4)
    e = rng.standard_normal((400, 4))
    tan = e.copy()
    ob.tangent(m, s0["t"], tape, te, tan)
    # total momentum
    for c in (2, 3):
        p0 = (m * e[:, c]).sum()
        p1 = (m * tan[:, c]).sum()
        assert abs(p1 - p0) < 1e-9 * np.abs(m * tan[:, c]).sum()
    # centre of mass moves with the total momentum
    for c in (0, 1):
        com0 = (m * (e[:, c] + (te - s0["t"]) * e[:, c + 2])).sum()
        com1 = (m * tan[:, c]).sum()
        assert abs(com1 - com0) < 1e-9 * np.abs(m * tan[:, c]).sum()
    # kinetic energy: d(1/2 m v^2) = m v . dv is conserved by elastic collisions
    o.sync_all(te)
    fin = o.state()
    dE0 = (m * (s0["vx"] * e[:, 2] + s0["vy"] * e[:, 3])).sum()
    dE1 = (m * (fin["vx"] * tan[:, 2] + fin["vy"] * tan[:, 3])).sum()
    assert abs(dE1 - dE0) < 1e-9 * np.abs(m * fin["vx"] * tan[:, 2]).sum()
