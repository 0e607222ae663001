"""The oracle's hard walls (addWallx / addWally of the reference's graphics
build, EDMD.c:2845-2895, 3321-3330, 3384-3430, 5428-5451, 5896-5913) against
what the mathematics fixes.

The reference's command-line build compiles the walls out (its addWallx and
addWally are constants 0, EDMD.c:290-291), so no reference vector exists for
them; these checks pin the restatement instead: exact wall-contact times and
bounce periods, the reflection rule, conservation laws and the ideal-gas
limit of the wall pressure."""
import ctypes as C

import numpy as np
import pytest

import edmd_fixtures as F

WALL = 5
QUIET = dict(first_screen=1e9, dtime=1e9, first_thermo=1e9)


def disks(pos, vel, L=20.0, rad=0.5):
    pos, vel = np.asarray(pos, float), np.asarray(vel, float)
    return {"x": pos[:, 0].copy(), "y": pos[:, 1].copy(), "vx": vel[:, 0].copy(),
            "vy": vel[:, 1].copy(), "rad": np.full(len(pos), rad), "m": np.ones(len(pos)),
            "Lx": L, "Ly": L}


def test_bounce_times_closed_form(oracle):
    """A disk falling straight onto the y walls of a channel hits them at
    t = y0 - r, then every (L - 2r) / |vy|, alternating walls; the other
    disk rests ten units away."""
    st = disks([(3.0, 7.25), (13.0, 10.0)], [(0.0, -1.0), (0.0, 0.0)])
    o = oracle.Oracle(st, wall_y=1, res_tan=1.0, first_screen=1e9, dtime=1e9,
                      first_thermo=1e9)
    code, tr = o.run(400)
    assert code == 0
    w = tr[tr["type"] == WALL]
    assert len(w) >= 8 and np.all(w["i"] == 0) and np.all(w["j"] == 1)
    k = np.arange(len(w))
    np.testing.assert_allclose(w["t"], 6.75 + 19.0 * k, rtol=1e-13)
    # every hit counts as a collision; velocities flip exactly
    s = o.state()
    assert s["vy"][0] == (-1.0 if len(w) % 2 == 0 else 1.0) and s["vx"][0] == 0.0
    assert o.observables()["ncol"] == len(w)


def test_corner_quirk_and_tangential_damping(oracle):
    """Defaults resW = 1, resT = 0: a hit keeps the normal speed and
    removes the tangential one.  In a corner cell moving towards both walls
    the x branch sets xy = 0 even when the y wall is the earlier
    (EDMD.c:2875-2894): the event comes at the y-wall time and reflects
    vx, and resT = 0 stops the disk in y."""
    st = disks([(1.5, 0.9), (15.0, 15.0)], [(-0.5, -1.0), (0.0, 0.3)])
    o = oracle.Oracle(st, wall_x=1, wall_y=1, first_screen=1e9, dtime=1e9, first_thermo=1e9)
    t, partner = o.predict_collision(0)
    # y wall at 0.4, x wall at 2.0: the y wall time, labelled xy = 0
    assert t == 0.4 and partner == -5
    code, tr = o.run(1)
    assert tr[0]["type"] == WALL and tr[0]["j"] == 0 and tr[0]["t"] == 0.4
    s = o.state()
    assert s["vx"][0] == 0.5 and s["vy"][0] == 0.0
    # a disk in a non-corner cell: tangential velocity removed, normal reversed
    st = disks([(10.0, 0.9), (15.0, 15.0)], [(0.3, -2.0), (0.0, 0.3)])
    o = oracle.Oracle(st, wall_y=1, first_screen=1e9, dtime=1e9, first_thermo=1e9)
    code, tr = o.run(1)
    assert tr[0]["type"] == WALL and tr[0]["j"] == 1 and tr[0]["t"] == 0.2
    s = o.state()
    assert s["vx"][0] == 0.0 and s["vy"][0] == 2.0
    # a negative tangential velocity times resT = 0 gives +0.0, not the
    # reference's -0.0 (whose crossing time would be -inf, EDMD.c:2416-2425)
    st2 = disks([(10.5, 0.9), (15.0, 15.0)], [(-0.3, -2.0), (0.0, 0.3)])
    o = oracle.Oracle(st2, wall_y=1, first_screen=1e9, dtime=1e9, first_thermo=1e9)
    o.run(1)
    s = o.state()
    assert s["vx"][0] == 0.0 and not np.signbit(s["vx"][0])
    t, d = o.predict_crossing(0)
    assert d == 4 and t == pytest.approx(0.2 + (2.0 - 0.5) / 2.0, rel=1e-12)
    # res_wall scales the normal speed
    o = oracle.Oracle(st, wall_y=1, res_wall=0.5, res_tan=1.0, first_screen=1e9, dtime=1e9,
                      first_thermo=1e9)
    o.run(1)
    s = o.state()
    assert s["vx"][0] == 0.3 and s["vy"][0] == 1.0


def test_no_minimum_image_across_walls(oracle):
    """Two disks on either side of a y wall see each other only through the
    box, not across the wall; across the periodic x boundary they do."""
    st = disks([(5.0, 0.6), (5.0, 19.4)], [(0.0, -0.1), (0.0, 0.1)])
    o = oracle.Oracle(st, wall_y=1)
    assert o.collision_time(0, 1) == 1e26
    o = oracle.Oracle(st)
    # across the periodic boundary: gap 0.2 closing at 0.2
    assert o.collision_time(0, 1) == pytest.approx(1.0, rel=1e-12)
    st = disks([(0.6, 5.0), (19.4, 5.0)], [(-0.1, 0.0), (0.1, 0.0)])
    o = oracle.Oracle(st, wall_y=1)
    assert o.collision_time(0, 1) == pytest.approx(1.0, rel=1e-12)


def test_channel_conserves_energy_and_x_momentum(oracle):
    """Specular elastic walls (resW = resT = 1) in y: energy and the x
    momentum stay at their initial values through thousands of wall hits and
    disk collisions; the y momentum does not (the walls take it)."""
    st = F.lattice_state(400, 0.5, seed=11, lattice="fit")
    o = oracle.Oracle(st, wall_y=1, res_tan=1.0, first_screen=1e9, dtime=1e9, first_thermo=1e9)
    e0 = o.observables()
    code, tr = o.run(40000)
    assert code == 0
    assert np.sum(tr["type"] == WALL) > 300
    e1 = o.observables()
    assert abs(e1["E"] / e0["E"] - 1) < 1e-12
    assert abs(e1["px"] - e0["px"]) < 1e-10
    assert abs(e1["py"] - e0["py"]) > 1e-3


def test_wall_pressure_ideal_gas_limit(oracle):
    """A dilute gas (phi = 0.02) between y walls: the wall pressure of the
    thermo rows (momentum to the walls per time per perimeter) approaches
    the ideal-gas value N T / A with T = E / N (2D, E = sum m v^2 / 2), with
    the low-density excluded-volume correction Z = 1 + 2 phi."""
    st = F.lattice_state(300, 0.02, seed=12, lattice="fit")
    o = oracle.Oracle(st, wall_y=1, res_tan=1.0, first_screen=1e9, dtime=1e9,
                      first_thermo=20.0, dtime_thermo=20.0)
    o.run_until(2000.0)
    rows = o.thermo_rows()
    pw = np.array([r["pwall"] for r in rows[1:]])
    assert len(pw) > 50
    A = st["Lx"] * st["Ly"]
    ideal = 300 * o.observables()["E"] / 300 / A * (1 + 2 * 0.02)
    se = pw.std(ddof=1) / np.sqrt(len(pw))
    assert abs(pw.mean() - ideal) < max(4 * se, 0.03 * ideal), (pw.mean(), ideal, se)
    # no walls: the column is zero
    o = oracle.Oracle(st, first_thermo=20.0, dtime_thermo=20.0)
    o.run_until(100.0)
    assert all(r["pwall"] == 0 for r in o.thermo_rows())


def test_field_and_damping_wait_for_growstop(oracle):
    """While the disks grow the Grow dispatch ignores the field and damping:
    straight flights (freeFlyGrow, EDMD.c:4992-5007) and straight crossing
    times without logTime (crossingEventGrow, EDMD.c:2343-2400).  A disk
    at y = 5 rising at 0.5 leaves its cell [4, 6) at t = 2, not at the
    parabola's 2.76 (field -0.1) nor the damped 2.23 (gamma 0.05); after
    GROWSTOP at t = 10 the same disk gets the Normal times."""
    st = disks([(9.0, 5.0), (15.0, 15.0)], [(0.0, 0.5), (0.0, 0.0)])
    for kw in (dict(add_field=1, field=-0.1), dict(damping=1, gamma=0.05)):
        o = oracle.Oracle(st, vr=0.1, **kw, **QUIET)
        t, d = o.predict_crossing(0)
        assert d == 4 and t == pytest.approx(2.0, rel=1e-15), kw
        code, tr = o.run(1)
        assert code == 0 and tr["type"][0] == F.CELLCROSS and tr["t"][0] == t
        oracle.lib().orc_sync_all(o.h, o.t)
        s = o.state()
        assert s["vy"][0] == 0.5 and s["y"][0] == pytest.approx(6.0, rel=1e-15)
    # after GROWSTOP (t = 1 at rate 1): the field's parabola from the state
    # stopGrow left (velocities rescaled to E/N = 1)
    f = -0.1
    o = oracle.Oracle(st, vr=1.0, add_field=1, field=f, **QUIET)
    code, tr = o.run(1)
    assert code == 0 and tr["type"][0] == F.GROWSTOP
    s = o.state()
    y, vy = s["y"][0], s["vy"][0]
    assert 4.0 < y < 6.0 and vy > 0
    t, d = o.predict_crossing(0)
    assert d == 4
    assert t - o.t == pytest.approx((-vy + np.sqrt(vy * vy - 2 * f * (y - 6.0))) / f, rel=1e-12)
    assert abs((t - o.t) - (6.0 - y) / vy) > 1e-3


# ---- the uniform field along y (addField, EDMD.c:2428-2442, 2664-2666,
# 2846-2860, 4957-4963)

def test_field_bounce_closed_form(oracle):
    """A disk dropped from rest at height h above contact in a field f < 0
    between y walls (resW = 1) hits the floor at t = sqrt(2h/|f|) and then
    every 2 sqrt(2h/|f|): the flights are exact parabolas and the bounce
    keeps the speed."""
    f = -0.5
    h = 6.0
    st = disks([(3.0, 0.5 + h), (13.0, 10.0)], [(0.0, 0.0), (0.0, 0.0)])
    o = oracle.Oracle(st, wall_y=1, add_field=1, field=f, res_tan=1.0, **QUIET)
    code, tr = o.run(300)
    assert code == 0
    w = tr[(tr["type"] == WALL) & (tr["i"] == 0)]
    assert len(w) >= 5 and np.all(w["j"] == 1)
    t1 = np.sqrt(2 * h / -f)
    np.testing.assert_allclose(w["t"], t1 * (1 + 2 * np.arange(len(w))), rtol=1e-12)
    # the resting disk falls too and bounces with its own period
    w1 = tr[(tr["type"] == WALL) & (tr["i"] == 1)]
    np.testing.assert_allclose(w1["t"][0], np.sqrt(2 * 9.5 / -f), rtol=1e-12)


def test_field_crossing_closed_form(oracle):
    """Crossing times under the field against the other root formula: up
    through the top edge when vy > sqrt(2 f dy), else down through the
    bottom edge."""
    f = -0.3
    for vy, expect_dir in ((2.0, 4), (0.2, 3), (-1.0, 3)):
        st = disks([(3.0, 4.7), (13.0, 13.0)], [(0.0, vy), (0.0, 0.0)])
        o = oracle.Oracle(st, add_field=1, field=f, **QUIET)
        t, d = o.predict_crossing(0)
        assert d == expect_dir
        # cell 2 spans [4, 6): top edge 1.3 above, bottom edge 0.7 below
        g = 1.3 if d == 4 else -0.7
        # root of g = vy t + f t^2 / 2 in the cancellation-free form
        disc = np.sqrt(vy * vy + 2 * f * g)
        tt = 2 * g / (vy + disc) if d == 4 else 2 * g / (vy - disc)
        assert t == pytest.approx(tt, rel=1e-12)


def test_field_box_conserves_total_energy(oracle):
    """Elastic disks (res 1) in a closed box under the field with specular
    walls (resW = resT = 1): kinetic plus potential energy (-m f y) stays
    constant through collisions, wall hits and parabolic flights."""
    f = -0.05
    st = F.lattice_state(200, 0.4, seed=17, lattice="fit")
    L = st["Lx"]
    st["x"] = 0.51 + st["x"] / L * (L - 1.02)
    st["y"] = 0.51 + st["y"] / L * (L - 1.02)
    o = oracle.Oracle(st, wall_x=1, wall_y=1, res_tan=1.0, add_field=1, field=f, **QUIET)

    def total():
        s = o.state()
        T = o.t
        dt = T - s["t"]
        vy = s["vy"] + dt * f
        y = s["y"] + dt * s["vy"] + 0.5 * f * dt * dt
        return 0.5 * np.sum(s["vx"] ** 2 + vy ** 2) - f * np.sum(y)

    e0 = total()
    code, tr = o.run(30000)
    assert code == 0 and np.sum(tr["type"] == WALL) > 200
    assert abs(total() / e0 - 1) < 1e-11


# ---- damping (damping / gamm, EDMD.c:4965-4975, 5709-5721, 2664-2666)

def test_own_exp_and_log_are_accurate(oracle):
    """The oracle's exp and log (the steps the GPU library also runs) agree
    with exact values to 4 ulp over the ranges the dynamics use (log adds
    e ln2 to log m, which cancels partly for x just below 1)."""
    from decimal import Decimal, getcontext
    getcontext().prec = 40
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(-40, 5, 4000), rng.uniform(-1e-3, 1e-3, 500), [0.0, -1e-300]])
    for x in xs:
        got = oracle.lib().orc_exp(float(x))
        ref = float(Decimal(float(x)).exp())
        assert abs(got - ref) <= 4 * np.spacing(ref), (x, got, ref)
    ys = np.concatenate([rng.uniform(1e-6, 1.0, 4000), rng.uniform(1.0, 1e6, 500), [1.0, 0.5]])
    for y in ys:
        got = oracle.lib().orc_log(float(y))
        ref = float(Decimal(float(y)).ln())
        assert abs(got - ref) <= 4 * np.spacing(abs(ref)) + (5e-324 if ref == 0 else 0), (y, got, ref)
    assert oracle.lib().orc_log(1.0) == 0.0 and oracle.lib().orc_exp(0.0) == 1.0


def test_damped_head_on_and_crossing_closed_form(oracle):
    """Under damping the separation of two disks closing at u shrinks by
    u (1 - e^{-g t}) / g: contact at t = -ln(1 - g gap / u) / g, never when
    g gap >= u; a crossing a distance D ahead comes at -ln(1 - g D / v) / g."""
    g = 0.05
    st = disks([(5.0, 5.0), (7.5, 5.0)], [(0.5, 0.0), (-0.5, 0.0)])  # gap 1.5, closing 1
    o = oracle.Oracle(st, damping=1, gamma=g, **QUIET)
    assert o.collision_time(0, 1) == pytest.approx(-np.log(1 - g * 1.5) / g, rel=1e-13)
    o = oracle.Oracle(st, damping=1, gamma=1.0, **QUIET)
    assert o.collision_time(0, 1) == 1e26
    st = disks([(4.5, 5.0), (15.0, 15.0)], [(0.8, 0.0), (0.0, 0.0)])  # cell [4, 6): 1.5 ahead
    o = oracle.Oracle(st, damping=1, gamma=g, **QUIET)
    t, d = o.predict_crossing(0)
    assert d == 2 and t == pytest.approx(-np.log(1 - g * 1.5 / 0.8) / g, rel=1e-13)


def test_damped_energy_decays_exactly(oracle):
    """Elastic collisions keep the kinetic energy at the instant and every
    flight scales velocities by e^{-g dt}, so the energy of the disks flown
    to a common time T is E0 e^{-2 g T}."""
    g = 0.03
    st = F.lattice_state(300, 0.5, seed=19, lattice="fit")
    o = oracle.Oracle(st, damping=1, gamma=g, **QUIET)
    e0 = o.observables()["E"]
    code, tr = o.run(15000)
    assert code == 0
    T = o.t
    lib = oracle.lib()
    lib.orc_sync_all(o.h, T)
    e1 = o.observables()["E"]
    assert abs(e1 / (e0 * np.exp(-2 * g * T)) - 1) < 1e-10, (e1, e0, T)


# ---- the field with damping (freeFlyNormal takes the field's branch first)

def test_field_with_damping_closed_form(oracle):
    """With both on, an x crossing comes at the damped time
    -ln(1 - g D / vx) / g (logTime), but the flight to it is the field's:
    x advances by vx t undamped, y falls by f t^2 / 2, vx keeps its value
    and vy gains f t (EDMD.c:4957-4975).  Taking the damped flight instead
    would leave the disk exactly on the cell edge with vx e^{-g t}."""
    g, f = 0.05, -0.02
    st = disks([(4.5, 5.0), (15.0, 15.0)], [(0.8, 0.0), (0.0, 0.0)])  # cell [4, 6): 1.5 ahead
    o = oracle.Oracle(st, damping=1, gamma=g, add_field=1, field=f, **QUIET)
    t1 = -np.log(1 - g * 1.5 / 0.8) / g
    tc, d = o.predict_crossing(0)
    assert d == 2 and tc == pytest.approx(t1, rel=1e-13)
    code, tr = o.run(1)
    assert code == 0 and tr["type"][0] == F.CELLCROSS and tr["i"][0] == 0
    oracle.lib().orc_sync_all(o.h, o.t)
    s = o.state()
    assert s["x"][0] == pytest.approx(4.5 + 0.8 * t1, rel=1e-13)
    assert s["x"][0] > 6.0 + 0.05
    assert s["vx"][0] == 0.8
    assert s["y"][0] == pytest.approx(5.0 + 0.5 * f * t1 * t1, rel=1e-13)
    assert s["vy"][0] == pytest.approx(f * t1, rel=1e-13)
    # the resting disk falls under the field, undamped
    assert s["vy"][1] == pytest.approx(f * t1, rel=1e-13)


# ---- hard walls under a growth start (collisionEventGrow / doTheWallGrow)

def test_walls_under_growth_closed_form(oracle):
    """A growing disk's surface closes on a wall at |v -/+ vr| (EDMD.c:
    3118-3153): from x = 1.2 at vx = -0.3 with vr = 0.1 (target 0.5 at rate
    0.2) the left wall comes at 1.2 / 0.4 = 3; the top wall from y = 18.8
    at vy = 0.2 at 1.2 / 0.3 = 4; the bottom wall from y = 0.6 at vy = -0.1
    at 0.6 / 0.2 = 3.  A hit (EDMD.c:3333-3382) counts as a collision and
    turns the normal velocity away with 2 vr added: vx = 0.3 + 0.2; at the
    top (y > Ly - 3) vy = -0.2 - 0.2; at the bottom vy = 0.1 + 0.2.  No
    momentum goes to the walls during growth."""
    st = disks([(1.2, 9.0), (9.0, 18.8), (15.0, 0.6)],
               [(-0.3, 0.0), (0.0, 0.2), (0.0, -0.1)])
    o = oracle.Oracle(st, wall_x=1, wall_y=1, vr=0.2, **QUIET)
    for i, (tw, xy) in enumerate([(3.0, 0), (4.0, 1), (3.0, 1)]):
        t, p = o.predict_collision(i)
        assert p == -5 - xy and t == pytest.approx(tw, rel=1e-14), (i, t, p)
    code, tr = o.run(3)
    assert code == 0
    assert list(tr["type"]) == [WALL] * 3
    assert sorted(tr["i"][:2]) == [0, 2] and tr["i"][2] == 1
    oracle.lib().orc_sync_all(o.h, o.t)
    s = o.state()
    assert s["vx"][0] == pytest.approx(0.5, rel=1e-15) and s["vy"][0] == 0.0
    assert s["vy"][1] == pytest.approx(-0.4, rel=1e-15)
    assert s["vy"][2] == pytest.approx(0.3, rel=1e-15)
    ob = o.observables()
    assert ob["ncol"] == 3
    # the radii grew with the flights: 0.1 per unit time up to t = 4
    np.testing.assert_allclose(o.radii(), 0.4, rtol=1e-14)


# ---- the circular wall (addCircularWall, EDMD.c:2938-2954, 3158-3172,
# 3358-3376, 3441-3454, 5723-5764)

def test_own_trig_is_accurate(oracle):
    """The oracle's atan2 / sin / cos (the steps the GPU library also runs)
    agree with the C library to 2 ulp over the angles the circular wall
    meets; atan2 keeps IEEE's quadrants and signed zeros."""
    L = oracle.lib()
    rng = np.random.default_rng(7)
    s, c = C.c_double(), C.c_double()
    for th in np.concatenate([rng.uniform(-np.pi, np.pi, 20000), [0.0, np.pi, -np.pi, np.pi / 2]]):
        L.orc_sincos(float(th), C.byref(s), C.byref(c))
        assert abs(s.value - np.sin(th)) <= 2 * np.spacing(max(abs(np.sin(th)), 1e-300)), th
        assert abs(c.value - np.cos(th)) <= 2 * np.spacing(max(abs(np.cos(th)), 1e-300)), th
    ys = rng.normal(size=20000) * np.exp(rng.uniform(-5, 5, 20000))
    xs = rng.normal(size=20000) * np.exp(rng.uniform(-5, 5, 20000))
    for y, x in zip(ys, xs):
        r = np.arctan2(y, x)
        assert abs(L.orc_atan2(float(y), float(x)) - r) <= 2 * np.spacing(abs(r)), (y, x)
    for y, x in [(0.0, 1.0), (0.0, -1.0), (-0.0, -1.0), (1.0, 0.0), (-1.0, 0.0), (-0.0, 2.0)]:
        assert L.orc_atan2(y, x) == np.arctan2(y, x)
        assert np.signbit(L.orc_atan2(y, x)) == np.signbit(np.arctan2(y, x))


def test_circular_wall_closed_form(oracle):
    """In a circle of radius 10 about (10, 10) a disk of radius 0.5 at
    (11, 10) moving along +x at 1 meets the wall when its centre is 9.5 out,
    at t = 8.5, comes back along the diameter (the normal is exactly +x)
    and meets the far side 19 later; a hit at an oblique angle keeps the
    speed and reverses the normal component; the pressure row divides by
    the area pi 10^2 and the wall momentum by the perimeter pi 20."""
    st = disks([(11.0, 10.0), (10.0, 5.0)], [(1.0, 0.0), (0.0, 0.01)])
    o = oracle.Oracle(st, wall_circle=1, **QUIET)
    t, p = o.predict_collision(0)
    assert p == -5 - 2 and t == pytest.approx(8.5, rel=1e-15)
    o.run_until(9.0)
    oracle.lib().orc_sync_all(o.h, o.t)
    s = o.state()
    assert s["vx"][0] == -1.0 and s["vy"][0] == 0.0
    t, p = o.predict_collision(0)
    assert p == -7 and t == pytest.approx(27.5, rel=1e-14)
    # oblique: from (10, 12) at (1, 0.5)
    st = disks([(10.0, 12.0), (10.0, 5.0)], [(1.0, 0.5), (0.0, 0.01)])
    o = oracle.Oracle(st, wall_circle=1, **QUIET)
    t, p = o.predict_collision(0)
    assert p == -7
    code, tr = o.run(40)
    k = np.nonzero((tr["type"] == WALL) & (tr["i"] == 0))[0][0]
    o2 = oracle.Oracle(st, wall_circle=1, **QUIET)
    o2.run(int(k) + 1)
    oracle.lib().orc_sync_all(o2.h, o2.t)
    s = o2.state()
    x, y = s["x"][0] - 10, s["y"][0] - 10
    assert np.hypot(x, y) == pytest.approx(9.5, rel=1e-12)
    n = np.array([x, y]) / np.hypot(x, y)
    v1 = np.array([s["vx"][0], s["vy"][0]])
    assert np.hypot(*v1) == pytest.approx(np.hypot(1.0, 0.5), rel=1e-14)
    assert v1 @ n == pytest.approx(-(np.array([1.0, 0.5]) @ n), rel=1e-12)
    assert v1 @ n < 0


def test_circular_wall_under_growth_and_growstop(oracle):
    """A growing disk at the centre moving at v meets the circle when
    v t + vr t = R: t = 10 / 1.05 (target 0.5 at rate 0.1); the hit turns
    vx to v - 2 v - 2 vr = -1.1.  At GROWSTOP the circle's normalisation
    removes the angular momentum about the centre (EDMD.c:5730-5749) and
    scales to E/N = Einit."""
    st = disks([(10.0, 10.0), (10.0, 6.0)], [(1.0, 0.0), (0.3, -0.2)])
    o = oracle.Oracle(st, wall_circle=1, vr=0.1, **QUIET)
    t, p = o.predict_collision(0)
    assert p == -7 and t == pytest.approx(10 / 1.05, rel=1e-14)
    o.run_until(t + 0.01)
    oracle.lib().orc_sync_all(o.h, o.t)
    assert o.state()["vx"][0] == pytest.approx(-1.1, rel=1e-14)
    st = F.growth_state(200, 0.4, seed=41)
    c = st["Lx"] / 2
    keep = (st["x"] - c) ** 2 + (st["y"] - c) ** 2 < (c - 1.0) ** 2
    st = {k: (v[keep] if isinstance(v, np.ndarray) else v) for k, v in st.items()}
    o = oracle.Oracle(st, wall_circle=1, vr=0.1, **QUIET)
    code, tr = o.run(20000)
    assert code == 0 and np.sum(tr["type"] == F.GROWSTOP) == 1
    s = o.state()
    k = np.nonzero(tr["type"] == F.GROWSTOP)[0][0]
    o2 = oracle.Oracle(st, wall_circle=1, vr=0.1, **QUIET)
    o2.run(int(k) + 1)
    s = o2.state()
    Lz = np.sum(st["m"] * ((s["x"] - c) * s["vy"] - (s["y"] - c) * s["vx"]))
    E = 0.5 * np.sum(st["m"] * (s["vx"] ** 2 + s["vy"] ** 2))
    assert abs(Lz) < 1e-10 * np.sum(st["m"] * np.hypot(s["vx"], s["vy"]) * c)
    assert E / len(s["x"]) == pytest.approx(1.0, rel=1e-12)


# ---- the mid wall (addMidWall, EDMD.c:2897-2915, 3175-3195, 3393-3409,
# 5453-5477)

def test_mid_wall_closed_form(oracle):
    """x = Lx/2 = 10 stops the disks of type 1 met within 4 of it: from
    x = 8 moving at +1 a type-1 disk (radius 0.5) hits at t = 1.5 and comes
    back; from x = 12 moving at -1, also at 1.5; a type-0 disk on the same
    path passes; growing type-1 disks close on it at v + vr."""
    st = disks([(8.0, 5.0), (12.0, 15.0), (8.0, 10.0), (3.0, 18.0)],
               [(1.0, 0.0), (-1.0, 0.0), (1.0, 0.0), (0.0, 0.01)])
    st["type"] = np.array([1, 1, 0, 0])
    o = oracle.Oracle(st, mid_wall=1, **QUIET)
    for i in (0, 1):
        t, p = o.predict_collision(i)
        assert p == -5 and t == pytest.approx(1.5, rel=1e-15), i
    t, p = o.predict_collision(2)
    assert p >= 0 or t > 1e20
    code, tr = o.run(12)
    assert code == 0 and o.t > 3.0
    walls = tr["i"][tr["type"] == WALL]
    assert sorted(walls[:2]) == [0, 1] and 2 not in walls
    oracle.lib().orc_sync_all(o.h, o.t)
    s = o.state()
    assert s["vx"][0] == -1.0 and s["vx"][1] == 1.0 and s["vx"][2] == 1.0
    assert s["x"][2] == pytest.approx(8.0 + o.t, rel=1e-14)
    # growth: surface speed v + vr (vr = 0.1 * 0.5)
    o = oracle.Oracle(st, mid_wall=1, vr=0.1, **QUIET)
    t, p = o.predict_collision(0)
    assert p == -5 and t == pytest.approx(2.0 / 1.05, rel=1e-14)


def test_mid_wall_thermo_columns(oracle):
    """A channel with walls at x = 0, Lx and the mid wall: the solute
    (type 1, big disks) starts left and stays there; every x-wall hit goes
    to dp, and to dpLeft (within 3 of x = 0), dpMid or dpRight; each thermo
    row carries pmid = dpMid / (dt Ly), -(dpRight - dpLeft) / (dt Ly), the
    solvent's packing pi n sr^2 / (Lx/2 Ly) on each side and the mean
    kinetic energies, taken from the stored disks at the row's time."""
    rng = np.random.default_rng(12)
    L, n = 40.0, 300
    xs, ys = [], []
    while len(xs) < n:
        x, y = rng.uniform(1.5, L - 1.5), rng.uniform(1.5, L - 1.5)
        if all((x - a) ** 2 + (y - b) ** 2 > 1.2 ** 2 for a, b in zip(xs, ys)) and abs(x - 20) > 1:
            xs.append(x)
            ys.append(y)
    xs, ys = np.array(xs), np.array(ys)
    typ = ((xs < 20) & (np.arange(n) % 3 == 0)).astype(np.int32)
    v = rng.standard_normal((n, 2))
    st = {"x": xs, "y": ys, "vx": v[:, 0].copy(), "vy": v[:, 1].copy(),
          "rad": np.full(n, 0.5), "m": np.ones(n), "Lx": L, "Ly": L, "type": typ}
    kw = dict(mid_wall=1, wall_x=1, wall_y=1, res_tan=1.0, sizeratio=0.5,
              first_screen=1e9, dtime=1e9, first_thermo=2.0, dtime_thermo=2.0)
    o = oracle.Oracle(st, **kw)
    o.run_until(10.5)
    rows = o.thermo_rows()
    assert len(rows) == 5
    assert all(r["pwall"] > 0 for r in rows) and sum(r["pmid"] > 0 for r in rows) >= 3
    s = o.state()
    assert np.all(s["x"][typ == 1] < 20)
    o2 = oracle.Oracle(st, **kw)
    o2.run_until(2.0)
    s = o2.state()
    r = o2.thermo_rows()[0]
    solv = typ == 0
    e = 0.5 * (s["vx"] ** 2 + s["vy"] ** 2)
    nl, nr = np.sum(solv & (s["x"] < 20)), np.sum(solv & (s["x"] > 20))
    ratio = np.pi / (20.0 * L)
    assert r["pack_left"] == pytest.approx(ratio * nl * 0.25, rel=1e-14)
    assert r["pack_right"] == pytest.approx(ratio * nr * 0.25, rel=1e-14)
    assert r["temp_left"] == pytest.approx(np.sum(e[solv & (s["x"] < 20)]) / nl, rel=1e-12)
    assert r["temp_right"] == pytest.approx(np.sum(e[solv & (s["x"] > 20)]) / nr, rel=1e-12)
    assert r["temp_solute"] == pytest.approx(np.mean(e[typ == 1]), rel=1e-12)


def test_circular_wall_under_field_closed_form(oracle):
    """Under the field the circle time is the smallest positive root of
    the quartic |x0 + v t + f t^2/2 e_y|^2 = (R - r)^2 (EDMD.c:2918-2935,
    quartic.c): a disk dropped from rest at the centre meets it at
    sqrt(2 (R - r) / |f|); a thrown disk where bisection on the trajectory
    finds the first contact.  The solver finds the smallest positive root
    of quartics with known roots to 1e-8."""
    L = oracle.lib()
    rng = np.random.default_rng(3)
    for _ in range(2000):
        roots = rng.uniform(-5, 10, 4)
        c = np.poly(roots) * rng.uniform(0.1, 2)
        pos = roots[roots > 1e-9]
        if len(pos):
            assert L.orc_smallest_root(*map(float, c)) == pytest.approx(pos.min(), rel=1e-8)
    f = -0.1
    st = disks([(10.0, 10.0), (10.0, 4.0)], [(0.0, 0.0), (0.3, 0.0)])
    o = oracle.Oracle(st, wall_circle=1, add_field=1, field=f, **QUIET)
    t, p = o.predict_collision(0)
    assert p == -7 and t == pytest.approx(np.sqrt(2 * 9.5 / 0.1), rel=1e-9)
    x0, y0, vx, vy = 12.0, 11.0, 0.7, 0.9
    st = disks([(x0, y0), (10.0, 4.0)], [(vx, vy), (0.3, 0.0)])
    o = oracle.Oracle(st, wall_circle=1, add_field=1, field=f, **QUIET)
    t, p = o.predict_collision(0)

    def gap(s):
        return np.hypot(x0 + vx * s - 10, y0 + vy * s + 0.5 * f * s * s - 10) - 9.5
    ts = np.linspace(1e-6, 50, 200001)
    k = np.nonzero(gap(ts) > 0)[0][0]
    lo, hi = ts[k - 1], ts[k]
    for _ in range(100):
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if gap(mid) < 0 else (lo, mid)
    assert p == -7 and t == pytest.approx(lo, rel=1e-9)


# ---- thermal and shear walls (thermoWall, addShearWall)

@pytest.mark.parametrize("seed", [5489, 0, 1, 2 ** 31 + 7])
def test_mt19937_is_the_published_stream(oracle, seed):
    """The oracle's MT19937 against numpy's legacy RandomState, which seeds
    by init_genrand and returns the raw 32-bit outputs for the full range;
    3000 draws cross four twists.  The first output for 5489 is the
    generator's published 3499211612."""
    got = oracle.mt_draws(seed, 3000)
    want = np.random.RandomState(seed).randint(0, 2 ** 32, size=3000, dtype=np.uint32)
    np.testing.assert_array_equal(got, want)
    if seed == 5489:
        assert got[0] == 3499211612


def _drand0(k, hi):
    return (np.float64(k) * (1.0 / 4294967296.0)) * (hi - 0.0) + 0.0


def test_thermal_wall_closed_form(oracle):
    """A disk of mass 2 meets the left x wall, then the right one: each hit
    sets vx to -sign(vx) drand(0, T/m), the next output of the stream,
    keeps vy, and still counts the momentum (resW + 1) m |vx| to the
    wall."""
    T, seed = 0.8, 77
    st = disks([(2.0, 7.0), (14.0, 14.0)], [(-1.0, 0.25), (0.0, 0.0)])
    st["m"] = np.array([2.0, 1.0])
    kw = dict(wall_x=1, thermal_wall=1, temperature=T, rng_seed=seed, **QUIET)
    k = np.random.RandomState(seed).randint(0, 2 ** 32, size=4, dtype=np.uint32)
    code, tr = oracle.Oracle(st, **kw).run(200)
    w = np.nonzero(tr["type"] == WALL)[0]
    assert len(w) >= 2 and tr["t"][w[0]] == 1.5 and np.all(tr["i"][w[:2]] == 0)
    o = oracle.Oracle(st, **kw)
    o.run(int(w[0]) + 1)
    s = o.state()
    assert s["vx"][0] == _drand0(k[0], T / 2.0) and s["vy"][0] == 0.25
    o = oracle.Oracle(st, **kw)
    o.run(int(w[1]) + 1)
    s = o.state()
    assert s["vx"][0] == -_drand0(k[1], T / 2.0) and s["vy"][0] == 0.25


def test_thermal_circle_and_shear_wall_closed_form(oracle):
    """The circle: a disk running along +x from the centre meets it at the
    angle 0, so v = (-drand(0, T), -0 * drand(0, T)), two draws.  Shear
    walls: a y-wall hit sets vx to +0.1 at the top and -0.1 at the bottom
    (after the normal reflection)."""
    T, seed = 1.3, 5
    st = disks([(10.0, 10.0), (10.0, 4.0)], [(1.0, 0.0), (0.0, 0.0)])
    o = oracle.Oracle(st, wall_circle=1, thermal_wall=1, temperature=T, rng_seed=seed, **QUIET)
    k = np.random.RandomState(seed).randint(0, 2 ** 32, size=2, dtype=np.uint32)
    code, tr = o.run(30)
    w = np.nonzero(tr["type"] == WALL)[0]
    assert len(w) >= 1
    o2 = oracle.Oracle(st, wall_circle=1, thermal_wall=1, temperature=T, rng_seed=seed, **QUIET)
    o2.run(int(w[0]) + 1)
    s = o2.state()
    assert s["x"][0] == pytest.approx(19.5, abs=1e-12)
    assert s["vx"][0] == -_drand0(k[0], T)
    assert s["vy"][0] == -0.0 * _drand0(k[1], T) and np.signbit(s["vy"][0])
    # shear walls
    st = disks([(5.0, 18.0), (12.0, 2.0)], [(0.3, 1.0), (-0.2, -1.0)])
    kw = dict(wall_y=1, shear_wall=1, res_tan=1.0, **QUIET)
    code, tr = oracle.Oracle(st, **kw).run(100)
    w = np.nonzero(tr["type"] == WALL)[0]
    assert len(w) >= 2 and sorted(tr["i"][w[:2]]) == [0, 1]
    o = oracle.Oracle(st, **kw)
    o.run(int(w[1]) + 1)
    s = o.state()
    assert s["vx"][0] == 0.1 and s["vy"][0] == -1.0
    assert s["vx"][1] == -0.1 and s["vy"][1] == 1.0


def test_thermal_walls_set_the_wall_temperature(oracle):
    """A dilute gas between thermal walls: after many hits each disk's
    speed is set by its last wall, and the mean kinetic energy per disk
    settles near the walls' E[(T/m) u]^2 / 2-weighted value; here only its
    finiteness, its change from the start and momentum bookkeeping are
    required, with the draws reproducible by seed."""
    st = F.walled_state(100, 0.10, 3, wall_x=True, wall_y=True)
    runs = []
    for rep in range(2):
        o = oracle.Oracle(st, wall_x=1, wall_y=1, thermal_wall=1, temperature=0.2, rng_seed=11,
                          first_screen=1e9, dtime=1e9, first_thermo=1.0, dtime_thermo=1.0)
        code, tr = o.run(20000)
        assert code == 0
        runs.append(o.state())
    for k in ("x", "y", "vx", "vy"):
        np.testing.assert_array_equal(runs[0][k], runs[1][k])
    e = 0.5 * float(np.sum(runs[0]["vx"] ** 2 + runs[0]["vy"] ** 2)) / 100
    assert np.isfinite(e) and e < 0.5
