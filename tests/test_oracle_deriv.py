"""Pins of the oracle's derivatives of the event map (tangent, adjoint and
the fixed-sequence replay they differentiate).

The reference has no derivatives, so these checks come from the
mathematics: closed forms for a head-on collision, central finite
differences (step 1e-7) of the nonlinear fixed-sequence replay, the
adjoint/tangent duality, and the invariants the collision rule keeps
(momentum, energy, centre of mass) carried over to the linearised map.
Flat walls (x, y and mid walls) enter the tape as wall records; a single
bounce is checked against its closed form, walled systems by finite
differences and duality.
"""
import numpy as np
import pytest

import oracle_bind as ob
from edmd_fixtures import lattice_state, walled_state

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


# ---- flat walls on the tape

def test_wall_bounce_closed_form():
    """One disk reflected by the left x wall (resW 0.8, resT 0.5) and
    flown to T = 4: x(T) = r - rw vx (T - th), y(T) = y0 + vy th + rt vy
    (T - th) with th = (x0 - r) / (-vx); their partial derivatives."""
    x0, y0, vx, vy, r, T, rw, rt = 3.0, 10.0, -1.0, 0.6, 0.5, 4.0, 0.8, 0.5
    st = {"x": np.array([x0, 15.0]), "y": np.array([y0, 5.0]),
          "vx": np.array([vx, 0.0]), "vy": np.array([vy, 0.0]),
          "rad": np.array([r, r]), "m": np.array([1.0, 1.0]), "Lx": 20.0, "Ly": 20.0}
    o, s0, tape = taped(st, 1, wall_x=1, res_wall=rw, res_tan=rt)
    th = (x0 - r) / (-vx)
    assert len(tape) == 1 and tape["j"][0] == -1 and tape["t"][0] == th
    assert tape["ax"][0] == vx and tape["ay"][0] == vy and tape["ff"][0] == r
    m, tl0 = st["m"], s0["t"]
    want = {  # output component: gradient row of disk 0 (x0, y0, vx, vy)
        0: [-rw, 0, -rw * T, 0],
        1: [(1 - rt) * vy / (-vx), 1, (1 - rt) * vy * (x0 - r) / vx ** 2, th + rt * (T - th)],
        2: [0, 0, -rw, 0],
        3: [0, 0, 0, rt],
    }
    for c, row in want.items():
        lam = np.zeros((2, 4))
        lam[0, c] = 1.0
        g = ob.adjoint(m, tl0, tape, T, 0.0, 0.0, lam)
        np.testing.assert_allclose(g[0], row, rtol=1e-14, atol=1e-15, err_msg=str(c))
        np.testing.assert_array_equal(g[1], 0.0)
        tan = np.zeros((2, 4))
        tan[0] = row  # the tangent along the gradient row gives its square norm
        ob.tangent(m, tl0, tape, T, tan)
        assert tan[0, c] == pytest.approx(np.dot(row, row), rel=1e-14)
    # the wall is no collision: no virial, and t_last stays unset
    assert o.tape_wsum() == 0.0
    tan = np.ones((2, 4))
    dW, dtl = ob.tangent(m, tl0, tape, T, tan)
    assert dW == 0.0 and dtl == 0.0


def _last_pair_time(tape):
    return tape["t"][tape["j"] >= 0][-1]


WALLED = [
    ("channel", dict(wall_x=1), dict(wall_x=True)),
    ("box", dict(wall_x=1, wall_y=1, res_wall=0.9, res_tan=0.7), dict(wall_x=True, wall_y=True)),
    ("box_rt0", dict(wall_x=1, wall_y=1), dict(wall_x=True, wall_y=True)),
]


@pytest.mark.parametrize("name,params,walls", WALLED)
def test_walled_adjoint_matches_finite_differences(name, params, walls):
    """W / t_last plus a random functional of the final state over 600
    events of N = 400 at phi = 0.55 between walls, by central differences
    of the replay; the unperturbed replay reproduces the engine."""
    st = walled_state(400, 0.55, 21, **walls)
    o, s0, tape = taped(st, 600, **params)
    assert np.sum(tape["j"] < 0) >= 8, name
    te = o.t
    S, tk, fin = ob.replay(s0, tape, te)
    assert tk == pytest.approx(_last_pair_time(tape), abs=1e-12)
    assert S == pytest.approx(o.tape_wsum(), rel=1e-12)
    o.sync_all(te)
    ref = o.state()
    for k in ("x", "y", "vx", "vy"):
        d = fin[k] - ref[k]
        if k == "x" and not walls.get("wall_x") or k == "y" and not walls.get("wall_y"):
            d -= st["Lx"] * np.round(d / st["Lx"])
        assert np.abs(d).max() < 1e-10, (name, k)
    rng = np.random.default_rng(3)
    lam = rng.standard_normal((400, 4))
    lw, lt = 1 / tk, -S / tk ** 2
    g = ob.adjoint(st["m"], s0["t"], tape, te, lw, lt, lam)
    for _ in range(2):
        e = rng.standard_normal((400, 4))
        fd = ob.replay_fd(s0, tape, te, e, 1e-7, lw, lt, lam)
        ad = (g * e).sum()
        assert abs(fd - ad) <= 1e-6 * abs(ad), name


def test_mid_wall_adjoint_matches_finite_differences():
    """The solute's hits of the mid wall are wall records at x = Lx/2 -/+ r."""
    rng = np.random.default_rng(7)
    L, n = 30.0, 160
    xs, ys = [], []
    while len(xs) < n:
        x, y = rng.uniform(1.5, L - 1.5), rng.uniform(1.5, L - 1.5)
        if abs(x - L / 2) > 1 and all((x - a) ** 2 + (y - b) ** 2 > 1.2 ** 2
                                      for a, b in zip(xs, ys)):
            xs.append(x)
            ys.append(y)
    xs, ys = np.array(xs), np.array(ys)
    v = rng.standard_normal((n, 2))
    st = {"x": xs, "y": ys, "vx": v[:, 0].copy(), "vy": v[:, 1].copy(),
          "rad": np.full(n, 0.5), "m": np.ones(n), "Lx": L, "Ly": L,
          "type": ((xs < L / 2) & (np.arange(n) % 2 == 0)).astype(np.int32)}
    o, s0, tape = taped(st, 250, wall_x=1, wall_y=1, mid_wall=1, res_tan=1.0)
    mid = (tape["j"] == -1) & (np.abs(tape["ff"] - L / 2) < 1)
    assert mid.sum() >= 3
    te = o.t
    S, tk, _ = ob.replay(s0, tape, te)
    lw, lt = 1 / tk, -S / tk ** 2
    lam = rng.standard_normal((n, 4))
    g = ob.adjoint(st["m"], s0["t"], tape, te, lw, lt, lam)
    e = rng.standard_normal((n, 4))
    fd = ob.replay_fd(s0, tape, te, e, 1e-7, lw, lt, lam)
    assert abs(fd - (g * e).sum()) <= 1e-6 * abs((g * e).sum())


def test_walled_duality_long_horizon():
    """5000 events in a closed box with inelastic walls: <grad, e> equals
    the tangent's functional to rounding."""
    st = walled_state(400, 0.55, 23, wall_x=True, wall_y=True)
    o, s0, tape = taped(st, 5000, wall_x=1, wall_y=1, res_wall=0.95, res_tan=0.8)
    te = o.t
    rng = np.random.default_rng(12)
    lam = rng.standard_normal((400, 4))
    g = ob.adjoint(st["m"], s0["t"], tape, te, 0.3, -2.0, lam)
    e = rng.standard_normal((400, 4))
    tan = e.copy()
    dW, dtl = ob.tangent(st["m"], s0["t"], tape, te, tan)
    lhs = (g * e).sum()
    rhs = 0.3 * dW - 2.0 * dtl + (lam * tan).sum()
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs)


# ---- the field on the tape

def test_bounce_under_the_field_closed_form():
    """A disk dropped from rest (vy = 0, vx = 0.3) onto the bottom wall
    under the field f < 0 (resW 0.9, resT 0.5), flown to T after the hit:
    th = sqrt(2 (y0 - r)/|f|), the bounce speed resW |f| th, and the
    partial derivatives of x(T), y(T), vx(T), vy(T) that follow; the
    field's pull over the shifted hit time enters the velocity's."""
    x0, y0, vx, r, f, rw, rt, T = 5.0, 6.0, 0.3, 0.5, -0.5, 0.9, 0.5, 6.0
    af = -f
    st = {"x": np.array([x0, 15.0]), "y": np.array([y0, 15.0]),
          "vx": np.array([vx, 0.0]), "vy": np.array([0.0, 0.0]),
          "rad": np.array([r, r]), "m": np.array([1.0, 1.0]), "Lx": 20.0, "Ly": 20.0}
    o, s0, tape = taped(st, 1, wall_y=1, add_field=1, field=f, res_wall=rw, res_tan=rt)
    th = np.sqrt(2 * (y0 - r) / af)
    assert len(tape) == 1 and tape["j"][0] == -2 and tape["g"][0] == f
    assert tape["t"][0] == pytest.approx(th, rel=1e-13)
    want = {
        0: [1, (1 - rt) * vx / (af * th), th + rt * (T - th), (1 - rt) * vx / af],
        1: [0, ((1 + rw) * (T - th) - rw * th) / th, 0, (T - th) - rw * th],
        2: [0, 0, rt, 0],
        3: [0, (1 + rw) / th, 0, 1],
    }
    m, tl0 = st["m"], s0["t"]
    for c, row in want.items():
        lam = np.zeros((2, 4))
        lam[0, c] = 1.0
        g = ob.adjoint(m, tl0, tape, T, 0.0, 0.0, lam)
        np.testing.assert_allclose(g[0], row, rtol=1e-12, atol=1e-13, err_msg=str(c))
        tan = np.zeros((2, 4))
        tan[0] = row
        ob.tangent(m, tl0, tape, T, tan)
        assert tan[0, c] == pytest.approx(np.dot(row, row), rel=1e-12)


FIELD_CASES = [
    ("gravity box", dict(wall_x=1, wall_y=1, add_field=1, field=-0.1, res_wall=0.9, res_tan=0.5),
     dict(wall_x=True, wall_y=True)),
    ("gravity channel", dict(wall_y=1, add_field=1, field=-0.2, res_tan=1.0), dict(wall_y=True)),
    ("periodic fall", dict(add_field=1, field=-0.05), dict(wall_x=False)),
]


@pytest.mark.parametrize("name,params,walls", FIELD_CASES)
def test_field_adjoint_matches_finite_differences(name, params, walls):
    st = walled_state(400, 0.55, 31, **walls)
    o, s0, tape = taped(st, 600, **params)
    f = params["field"]
    te = o.t
    S, tk, fin = ob.replay(s0, tape, te, field=f)
    assert tk == pytest.approx(_last_pair_time(tape), abs=1e-12)
    assert S == pytest.approx(o.tape_wsum(), rel=1e-11)
    o.sync_all(te)
    ref = o.state()
    for k in ("x", "y", "vx", "vy"):
        d = fin[k] - ref[k]
        if k == "x" and not walls.get("wall_x") or k == "y" and not walls.get("wall_y"):
            d -= st["Lx"] * np.round(d / st["Lx"])
        assert np.abs(d).max() < 1e-9, (name, k)
    rng = np.random.default_rng(5)
    lam = rng.standard_normal((400, 4))
    lw, lt = 1 / tk, -S / tk ** 2
    g = ob.adjoint(st["m"], s0["t"], tape, te, lw, lt, lam)
    for _ in range(2):
        e = rng.standard_normal((400, 4))
        fd = ob.replay_fd(s0, tape, te, e, 1e-7, lw, lt, lam, field=f)
        ad = (g * e).sum()
        assert abs(fd - ad) <= 1e-6 * abs(ad), name


# ---- damping on the tape

def test_damped_head_on_closed_form():
    """Two unit disks head on under damping at rate g: the relative speed
    decays with the common factor E(t) = exp(-g t), contact comes when the
    displacement (v0 - v1)(1 - E)/g closes the gap x1 - x0 - 1, so
    E = 1 - g gap/(v0 - v1) and W = (a.c) = (v1 - v0) + g gap, and
    t_c = -log(E)/g.  Undamped, W would not depend on the positions."""
    g = 0.05
    st = two_disks(dx=2.0, v=1.0)
    o, s0, tape = taped(st, 1, damping=1, gamma=g)
    assert len(tape) == 1
    gap, dv = 1.0, 2.0
    E = 1 - g * gap / dv
    assert o.tape_wsum() == pytest.approx(-dv + g * gap, rel=1e-13)
    assert tape["t"][0] == pytest.approx(-np.log(E) / g, rel=1e-13)
    m, tl0 = st["m"], s0["t"]
    G = ob.adjoint(m, tl0, tape, 1.0, 1.0, gamma=g)
    np.testing.assert_allclose(G, [[-g, 0, -1, 0], [g, 0, 1, 0]], atol=1e-14)
    G = ob.adjoint(m, tl0, tape, 1.0, 0.0, 1.0, gamma=g)
    dtc = 1 / (E * dv)
    np.testing.assert_allclose(G[:, 0], [-dtc, dtc], rtol=1e-13)


def test_damped_wall_bounce_closed_form():
    """A disk damped at rate g meets the left x wall (resW 0.8, resT 0.5)
    and is flown to T: its velocity at the hit is vx + g (x0 - r), so
    vx(T) = -resW vx e^{-gT} and x(T) = r - resW (vh - vx e^{-gT})/g,
    vy(T) = resT vy e^{-gT}: the undamped partials with T replaced by
    (1 - e^{-gT})/g where the flight enters."""
    g, x0, y0, vx, vy, r, T, rw, rt = 0.1, 3.0, 10.0, -1.0, 0.6, 0.5, 4.0, 0.8, 0.5
    st = {"x": np.array([x0, 15.0]), "y": np.array([y0, 5.0]),
          "vx": np.array([vx, 0.0]), "vy": np.array([vy, 0.0]),
          "rad": np.array([r, r]), "m": np.array([1.0, 1.0]), "Lx": 20.0, "Ly": 20.0}
    o, s0, tape = taped(st, 1, wall_x=1, res_wall=rw, res_tan=rt, damping=1, gamma=g)
    assert len(tape) == 1 and tape["j"][0] == -1
    eT = np.exp(-g * T)
    E = 1 - g * (r - x0) / vx  # e^{-g th}
    phT = (1 - eT) / g
    want = {
        0: [-rw, 0, -rw * phT, 0],
        1: [(1 - rt) * vy / (-vx), 1, (1 - rt) * vy * (x0 - r) / vx ** 2,
            (1 - E) / g + rt * (E - eT) / g],
        2: [0, 0, -rw * eT, 0],
        3: [0, 0, 0, rt * eT],
    }
    m, tl0 = st["m"], s0["t"]
    for c, row in want.items():
        lam = np.zeros((2, 4))
        lam[0, c] = 1.0
        G = ob.adjoint(m, tl0, tape, T, 0.0, 0.0, lam, gamma=g)
        np.testing.assert_allclose(G[0], row, rtol=1e-12, atol=1e-13, err_msg=str(c))


DAMPED_CASES = [
    ("damped periodic", dict(damping=1, gamma=0.02), dict(wall_x=False)),
    ("damped box", dict(damping=1, gamma=0.01, wall_x=1, wall_y=1, res_wall=0.9, res_tan=0.5),
     dict(wall_x=True, wall_y=True)),
]


@pytest.mark.parametrize("name,params,walls", DAMPED_CASES)
def test_damped_adjoint_matches_finite_differences(name, params, walls):
    st = walled_state(400, 0.55, 41, **walls)
    o, s0, tape = taped(st, 600, **params)
    g = params["gamma"]
    te = o.t
    S, tk, fin = ob.replay(s0, tape, te, gamma=g)
    assert tk == pytest.approx(_last_pair_time(tape), abs=1e-11)
    assert S == pytest.approx(o.tape_wsum(), rel=1e-10)
    rng = np.random.default_rng(8)
    lam = rng.standard_normal((400, 4))
    lw, lt = 1 / tk, -S / tk ** 2
    G = ob.adjoint(st["m"], s0["t"], tape, te, lw, lt, lam, gamma=g)
    for _ in range(2):
        e = rng.standard_normal((400, 4))
        fd = ob.replay_fd(s0, tape, te, e, 1e-7, lw, lt, lam, gamma=g)
        ad = (G * e).sum()
        assert abs(fd - ad) <= 1e-6 * abs(ad), name
    e = rng.standard_normal((400, 4))
    tan = e.copy()
    dW, dtl = ob.tangent(st["m"], s0["t"], tape, te, tan, gamma=g)
    assert abs((G * e).sum() - (lw * dW + lt * dtl + (lam * tan).sum())) <= 1e-11 * abs((G * e).sum())
