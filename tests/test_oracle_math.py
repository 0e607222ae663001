"""The oracle against what the mathematics fixes: closed-form collision and
crossing times, a brute-force neighbour scan, conservation laws, the
collision rule, the hard-disk equation of state and g(r) identities.

These checks do not retype the oracle's formulas: each compares against an
independent statement (an exact root, all pairs, a conservation law, a
published equation of state) that a dropped term, a wrong sign or a
transposed operand would violate."""
import math
from decimal import Decimal, getcontext

import numpy as np
import pytest

import edmd_fixtures as F

NEVER = 1e26


def two_disks(p0, v0, p1, v1, L=20.0, rad=0.5):
    return {"x": np.array([p0[0], p1[0]]), "y": np.array([p0[1], p1[1]]),
            "vx": np.array([v0[0], v1[0]]), "vy": np.array([v0[1], v1[1]]),
            "rad": np.array([rad, rad]), "m": np.ones(2), "Lx": L, "Ly": L}


def test_head_on_closed_form(oracle):
    # gap 1 between surfaces, closing speed 1: contact at exactly t = 1
    o = oracle.Oracle(two_disks((5.0, 5.0), (0.5, 0.0), (7.0, 5.0), (-0.5, 0.0)))
    assert o.collision_time(0, 1) == 1.0
    t, partner = o.predict_collision(0)
    assert (t, partner) == (1.0, 1)
    # equal and opposite impulse on contact: velocities swap for equal masses
    st, n = o.run_collisions(1)
    assert st == 0 and o.t == 1.0
    s = o.state()
    assert s["vx"][0] == -0.5 and s["vx"][1] == 0.5


def exact_root(dx, dy, dvx, dvy, D2):
    getcontext().prec = 80
    dx, dy, dvx, dvy, D2 = map(Decimal, (dx, dy, dvx, dvy, D2))
    b = dx * dvx + dy * dvy
    v2 = dvx * dvx + dvy * dvy
    d = dx * dx + dy * dy - D2
    det = b * b - v2 * d
    if b > 0 or det < 0:
        return None
    return (-b - det.sqrt()) / v2


def test_collision_time_against_exact_root(oracle):
    """Oblique approaches with gaps from 1e-3 to 3 diameters: the FP64 root
    of the reference's formula agrees with the exact root to 1e-12."""
    rng = np.random.default_rng(11)
    checked = 0
    for _ in range(400):
        gap = 10 ** rng.uniform(-3, 0.5)
        ang = rng.uniform(0, 2 * math.pi)
        dist = 1.0 + gap
        p0 = np.array([10.0, 10.0])
        p1 = p0 + dist * np.array([math.cos(ang), math.sin(ang)])
        v0 = rng.normal(size=2)
        v1 = rng.normal(size=2)
        o = oracle.Oracle(two_disks(p0, v0, p1, v1))
        t = o.collision_time(0, 1)
        dx = p1[0] - p0[0]
        dy = p1[1] - p0[1]
        ex = exact_root(dx, dy, v1[0] - v0[0], v1[1] - v0[1], 1.0)
        if ex is None:
            assert t == NEVER
            continue
        checked += 1
        assert abs(Decimal(t) - ex) <= Decimal(1e-12) * abs(ex)
    assert checked > 100


def test_receding_grazing_and_overlap(oracle):
    o = oracle.Oracle(two_disks((5.0, 5.0), (-1.0, 0.0), (7.0, 5.0), (1.0, 0.0)))
    assert o.collision_time(0, 1) == NEVER  # b > 0
    o = oracle.Oracle(two_disks((5.0, 5.0), (1.0, 0.0), (7.0, 6.5), (1.0, -0.1)))
    assert o.collision_time(0, 1) == NEVER  # misses: det < 0
    o = oracle.Oracle(two_disks((5.0, 5.0), (1.0, 0.0), (5.9, 5.0), (-1.0, 0.0)))
    assert o.status == 3  # overlap found by the initial sweep, EDMD.c:2708-2717


def test_minimum_image_across_boundary(oracle):
    L = 20.0
    # centres 1.5 apart through the periodic boundary, closing speed 1: t = 0.5
    a = oracle.Oracle(two_disks((0.8, 5.0), (-0.5, 0.0), (L - 0.7, 5.0), (0.5, 0.0), L=L))
    b = oracle.Oracle(two_disks((10.8, 5.0), (-0.5, 0.0), (9.3, 5.0), (0.5, 0.0), L=L))
    ta, tb = a.collision_time(0, 1), b.collision_time(0, 1)
    assert abs(ta - 0.5) < 1e-12 and abs(tb - 0.5) < 1e-12


def test_crossing_closed_form(oracle):
    L = 20.0  # 10 cells of width 2
    o = oracle.Oracle(two_disks((1.0, 7.0), (2.0, 0.25), (15.0, 15.0), (0.0, -0.5), L=L))
    t, d = o.predict_crossing(0)
    assert (t, d) == (0.5, 2)  # (2 - 1) / 2, +x
    t, d = o.predict_crossing(1)
    assert (t, d) == (2.0, 3)  # vx = 0 gives +inf in x; (14 - 15) / -0.5, -y
    # equal x and y times: the y crossing is chosen (EDMD.c:2476-2480)
    o = oracle.Oracle(two_disks((1.0, 1.0), (1.0, 1.0), (15.0, 15.0), (0.0, 1.0), L=L))
    t, d = o.predict_crossing(0)
    assert (t, d) == (1.0, 4)


def test_pair_time_symmetric_when_synchronised(oracle):
    """Mirrored predictions from a synchronised state tie bit for bit, for
    every pair a 3x3 scan can see.  (Exactly at a separation of L/2 the
    minimum image, >= L/2 versus < -L/2, is not symmetric; such pairs are
    never neighbours.)"""
    st = F.lattice_state(400, 0.7, seed=3)
    o = oracle.Oracle(st)
    L = st["Lx"]
    n = 0
    for i in range(0, 400, 7):
        dx = st["x"] - st["x"][i]
        dy = st["y"] - st["y"][i]
        dx -= L * np.round(dx / L)
        dy -= L * np.round(dy / L)
        for j in np.nonzero((dx * dx + dy * dy < 25) & (np.arange(400) != i))[0]:
            assert F.bits_equal(o.collision_time(i, j), o.collision_time(j, i))
            n += 1
    assert n > 500


def random_disks(n, L, rng, rad=0.5):
    pts = []
    while len(pts) < n:
        p = rng.uniform(0, L, 2)
        ok = True
        for q in pts:
            d = p - q
            d -= L * np.round(d / L)
            if d @ d < (2 * rad * 1.02) ** 2:
                ok = False
                break
        if ok:
            pts.append(p)
    pts = np.array(pts)
    v = rng.normal(size=(n, 2))
    return {"x": pts[:, 0], "y": pts[:, 1], "vx": v[:, 0], "vy": v[:, 1],
            "rad": np.full(n, rad), "m": np.ones(n), "Lx": L, "Ly": L}


def all_pair_times(st, i):
    """Collision times of disk i with every other disk, all synchronised."""
    L = st["Lx"]
    dx = st["x"] - st["x"][i]
    dy = st["y"] - st["y"][i]
    dx = np.where(dx >= L / 2, dx - L, np.where(dx < -L / 2, dx + L, dx))
    dy = np.where(dy >= L / 2, dy - L, np.where(dy < -L / 2, dy + L, dy))
    dvx = st["vx"] - st["vx"][i]
    dvy = st["vy"] - st["vy"][i]
    b = dx * dvx + dy * dvy
    v2 = dvx * dvx + dvy * dvy
    d = dx * dx + dy * dy - 4 * st["rad"] * st["rad"][i]
    det = b * b - v2 * d
    with np.errstate(invalid="ignore", divide="ignore"):
        t = (-b - np.sqrt(np.maximum(det, 0))) / v2
    t = np.where((b > 0) | (det < 0), NEVER, t)
    t[i] = np.inf
    return t


@pytest.mark.parametrize("n,L", [(6, 3.2), (14, 5.5), (30, 7.0), (120, 16.0), (256, 30.0)])
def test_neighbour_scan_brute_force(oracle, n, L):
    """Boxes of 1-2 cells per side visit every disk (with duplicates): the 3x3
    result is then the all-pairs minimum.  Larger boxes: every collision that
    happens before either disk leaves its cell is found by the 3x3 scan."""
    rng = np.random.default_rng(n)
    st = random_disks(n, L, rng)
    o = oracle.Oracle(st)
    ncx = o.grid()[0]
    ev = o.events()
    for i in range(n):
        tt = all_pair_times(st, i)
        j = int(np.argmin(tt))
        if ncx <= 2:
            assert ev["t_coll"][i] == tt[j] and ev["partner"][i] == j
        else:
            if tt[j] < min(ev["t_cross"][i], ev["t_cross"][j]):
                assert ev["t_coll"][i] == tt[j] and ev["partner"][i] == j
            else:
                assert ev["t_coll"][i] >= tt[j]


def test_events_pop_in_time_order(oracle):
    st = F.lattice_state(800, 0.72, seed=21)
    o = oracle.Oracle(st)
    s, tr = o.run(20000)
    assert s == 0
    assert np.all(np.diff(tr["t"]) >= 0)


def test_conservation_over_long_trace(oracle):
    st = F.lattice_state(2000, 0.7, seed=5)
    o = oracle.Oracle(st)
    e0 = o.observables()["E"]
    s, _ = o.run(200000, trace=False)
    assert s == 0
    ob = o.observables()
    assert abs(ob["E"] - e0) <= 1e-12 * e0
    assert abs(ob["px"]) < 1e-11 and abs(ob["py"]) < 1e-11
    assert ob["ncol"] > 100000


def test_collision_rule_reverses_normal_velocity(oracle):
    p0, p1 = np.array([5.0, 5.0]), np.array([6.2, 5.7])
    v0, v1 = np.array([0.9, 0.4]), np.array([-0.8, -0.3])
    o = oracle.Oracle(two_disks(p0, v0, p1, v1))
    t = o.collision_time(0, 1)
    o.run_collisions(1)
    s = o.state()
    assert o.observables()["ncol"] == 1 and o.t == t
    n = (p1 + v1 * t) - (p0 + v0 * t)
    n /= np.linalg.norm(n)
    dv0 = v1 - v0
    dv1 = np.array([s["vx"][1] - s["vx"][0], s["vy"][1] - s["vy"][0]])
    assert abs(dv1 @ n + dv0 @ n) < 1e-13
    tan = np.array([-n[1], n[0]])
    assert abs(dv1 @ tan - dv0 @ tan) < 1e-13
    e0 = 0.5 * (v0 @ v0 + v1 @ v1)
    e1 = 0.5 * (s["vx"] @ s["vx"] + s["vy"] @ s["vy"])
    assert abs(e1 - e0) < 1e-14
    ob = o.observables()
    assert ob["virXX"] < 0 and ob["virYY"] < 0  # m_i m_j ff dx^2 with dx.dv < 0


def henderson_z(phi):
    """Henderson's hard-disk equation of state, Mol. Phys. 30, 971 (1975)."""
    return (1 + phi * phi / 8) / (1 - phi) ** 2


@pytest.mark.parametrize("phi,tol", [(0.02, 0.02), (0.5, 0.03)])
def test_pressure_equation_of_state(oracle, phi, tol):
    """Reduced pressure Z = P A / (N T) from the collision virial
    (EDMD.c:5328) against the hard-disk equation of state."""
    n = 400
    st = F.lattice_state(n, phi, seed=31)
    o = oracle.Oracle(st, first_thermo=20.0, dtime_thermo=60.0)
    s, _ = o.run_until(80.0 + 1e-6)
    assert s == 0
    rows = o.thermo_rows()
    assert len(rows) == 2
    r = rows[1]  # averaged over 20 < t < 80, after melting
    A = st["Lx"] * st["Ly"]
    z = r["pressure"] * A / (n * r["e_per_n"])
    assert abs(z / henderson_z(phi) - 1) < tol


def test_pcf_identities(oracle):
    st = F.lattice_state(900, 0.5, seed=1)
    L = st["Lx"]
    r, g = oracle.pcf(st["x"], st["y"], L, L, 0.05, L / 2)
    assert np.all(g[r < 1.0 - 0.05] == 0)  # no pair closer than a diameter
    # sum over the histogram recovers the number of pairs within max_r
    n = 900
    dr = 0.05
    pairs = np.sum(g * 2 * np.pi * r * dr * (n / (L * L)) * n) / 2
    dx = st["x"][:, None] - st["x"][None, :]
    dy = st["y"][:, None] - st["y"][None, :]
    dx -= L * np.round(dx / L)
    dy -= L * np.round(dy / L)
    rr = np.sqrt(dx * dx + dy * dy)[np.triu_indices(n, 1)]
    assert abs(pairs - np.sum(rr < dr * len(r))) < 1e-6
    # uniform random points: g ~ 1 away from the origin
    rng = np.random.default_rng(2)
    x, y = rng.uniform(0, 40, 3000), rng.uniform(0, 40, 3000)
    r, g = oracle.pcf(x, y, 40.0, 40.0, 1.0, 20.0)
    assert abs(np.mean(g[2:]) - 1) < 0.02


def test_growth_head_on_contact_time(oracle):
    """Two equal disks growing from radius 0 at rate w each, closing at
    speed u along x from gap g: contact when g - u t = 2 w t, so
    t = g / (u + 2 w); the growth rule then separates them at a relative
    normal speed of u + 4 w (restitution 1 against the moving surface)."""
    g, u, w = 3.0, 1.0, 0.25   # vr * target radius = w
    st = {"x": np.array([5.0, 5.0 + g]), "y": np.array([5.0, 5.0]),
          "vx": np.array([u / 2, -u / 2]), "vy": np.zeros(2),
          "rad": np.array([1.0, 1.0]), "m": np.ones(2), "Lx": 20.0, "Ly": 20.0}
    o = oracle.Oracle(st, vr=w)        # vr_i = vr * 1 = w, growth stops at 1/vr = 4
    tc = None
    while tc is None:                  # cell crossings come first
        code, tr = o.run(1)
        assert code == 0
        if tr["type"][0] == 1:
            tc = tr["t"][0]
    assert tc == pytest.approx(g / (u + 2 * w), rel=1e-15)
    s = o.state()
    assert s["vx"][1] - s["vx"][0] == pytest.approx(u + 4 * w, rel=1e-14)
    assert s["vx"][0] + s["vx"][1] == pytest.approx(0.0, abs=1e-15)
    assert o.radii()[0] == pytest.approx(w * tc, rel=1e-15)


def test_growth_stop_normalises(oracle):
    """After GROWSTOP every radius is its target (to rounding of the summed
    growth steps), the total momentum vanishes and E/N = Einit; counters
    restart at zero."""
    st = F.growth_state(120, 0.6, seed=8, poly=0.1)
    o = oracle.Oracle(st, vr=0.2, einit=1.7)
    while o.t < 5.0:
        code, _ = o.run(1, trace=False)
        assert code == 0
    r = o.radii()
    o.sync_all(o.t)
    np.testing.assert_allclose(o.radii(), st["rad"], rtol=1e-12)
    ob = o.observables()
    assert ob["E"] / 120 == pytest.approx(1.7, rel=1e-13)
    assert abs(ob["px"]) < 1e-12 and abs(ob["py"]) < 1e-12
    assert r.min() > 0
