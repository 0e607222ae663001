"""Hard walls on the GPU against the oracle, bit for bit.

Walls are the reference's addWallx / addWally (EDMD.c:2845-2895, 3321-3330,
3384-3430, 5428-5451, 5896-5913), which its command-line build compiles out;
the oracle's restatement is pinned by tests/test_oracle_walls.py.  Traces
are compared in raw order (kind, disks or wall, validity, time bits), states
and thermo rows (with the wall pressure) bitwise."""
import numpy as np
import pytest

import edmd_fixtures as F
from conftest import bits

pytestmark = pytest.mark.gpu

WALL = 5
QUIET = dict(first_screen=1e9, dtime=1e9, first_thermo=1e9)


def walled(n, phi, seed, walls, temperature=1.0):
    """A fitted-lattice start pressed affinely into [r + 0.01, L - r - 0.01]
    along each walled direction (the lattice's outer rows may sit closer
    than a radius to the box edge; the squeeze leaves every pair apart)."""
    st = F.lattice_state(n, phi, seed=seed, lattice="fit", temperature=temperature)
    r, L = 0.5, st["Lx"]
    lo, span = r + 0.01, L - 2 * (r + 0.01)
    for key, flag in (("x", "wall_x"), ("y", "wall_y")):
        if walls.get(flag):
            st[key] = lo + st[key] / L * span
    dx = st["x"][:, None] - st["x"][None, :]
    dy = st["y"][:, None] - st["y"][None, :]
    if not walls.get("wall_x"):
        dx -= L * np.round(dx / L)
    if not walls.get("wall_y"):
        dy -= L * np.round(dy / L)
    d2 = dx * dx + dy * dy
    np.fill_diagonal(d2, 9.0)
    assert d2.min() > 1.0
    return st


def assert_trace_equal(got, ref):
    assert len(got) == len(ref)
    for k in ("type", "i", "j", "valid"):
        bad = np.nonzero(got[k] != ref[k])[0]
        assert len(bad) == 0, "%s differs first at event %d" % (k, bad[0])
    bad = np.nonzero(bits(got["t"]) != bits(ref["t"]))[0]
    assert len(bad) == 0, "time differs first at event %d" % bad[0]


def assert_state_equal(got, ref):
    for k in ("x", "y", "vx", "vy", "t"):
        np.testing.assert_array_equal(bits(got[k]), bits(ref[k]), err_msg=k)
    np.testing.assert_array_equal(got["coll"].astype(np.uint64), ref["coll"].astype(np.uint64))
    np.testing.assert_array_equal(got["cx"], ref["cx"])
    np.testing.assert_array_equal(got["cy"], ref["cy"])


CASES = [
    # (n, phi, seed, walls and restitution)
    (400, 0.5, 21, dict(wall_y=1, res_tan=1.0)),              # specular channel
    (400, 0.6, 22, dict(wall_x=1, wall_y=1)),                  # closed box, resW 1, resT 0
    (600, 0.45, 23, dict(wall_x=1, res_wall=0.9, res_tan=0.5)),
    (1500, 0.7, 24, dict(wall_y=1, res_tan=1.0)),              # gather-free staged scan, deeper calendar
]


@pytest.mark.parametrize("n,phi,seed,walls", CASES)
def test_wall_traces_match_oracle(gpu, oracle, n, phi, seed, walls):
    st = walled(n, phi, seed, walls)
    nev = 20000
    s = gpu.System(st, trace_capacity=nev, **walls, **QUIET)
    ev0 = s.events()
    o = oracle.Oracle(st, **walls, **QUIET)
    oe = o.events()
    # initial sweep: wall slots report partner -5 - xy in both
    for k in ("t_coll", "t_cross"):
        np.testing.assert_array_equal(bits(ev0[k]), bits(oe[k]), err_msg=k)
    np.testing.assert_array_equal(ev0["partner"], oe["partner"])
    np.testing.assert_array_equal(ev0["dir"], oe["dir"])
    s.run(max_events=nev)
    code, tr = o.run(nev)
    assert code == 0
    g = s.trace()
    assert np.sum(g["type"] == WALL) > 50
    assert_trace_equal(g, tr)
    assert_state_equal(s.state(), o.state())
    assert s.observables()["ncol"] == o.observables()["ncol"]


def test_wall_pressure_rows_and_thermostat_match_oracle(gpu, oracle):
    """Thermo rows with the wall pressure, and the velocity-rescale
    thermostat (full re-predictions with walls) in a closed box."""
    st = walled(500, 0.55, 31, dict(wall_x=1, wall_y=1), temperature=2.0)
    kw = dict(wall_x=1, wall_y=1, res_tan=1.0, first_thermo=0.5, dtime_thermo=0.5,
              noise=2, dtnoise=0.3, temperature=1.5, first_screen=1e9, dtime=1e9)
    nev = 30000
    s = gpu.System(st, trace_capacity=nev, thermo_capacity=512, **kw)
    s.run(max_events=nev)
    o = oracle.Oracle(st, **kw)
    code, tr = o.run(nev)
    assert code == 0
    assert_trace_equal(s.trace(), tr)
    rows, orows = s.thermo(), o.thermo_rows()
    assert len(rows) == len(orows) > 10
    for r, q in zip(rows, orows):
        for k in ("t", "e_per_n", "pressure", "pxx", "pyy", "pxy", "pyx", "pwall"):
            assert bits(np.float64(r[k])) == bits(np.float64(q[k])), k
        assert r["pwall"] > 0


def test_wall_batch_mixed_residencies_and_shapes(gpu, oracle):
    """A batch mixing wall and periodic systems in one launch (the loop's
    wall variant serves both), from HBM and from LDS, every sweep shape:
    each system evolves as the oracle evolves it alone."""
    specs = [(300, 0.5, 41, dict(wall_y=1, res_tan=1.0)), (350, 0.6, 42, {}),
             (280, 0.55, 43, dict(wall_x=1, wall_y=1)), (320, 0.65, 44, dict(wall_x=1))]
    states = [walled(n, phi, sd, w) for n, phi, sd, w in specs]
    per = {k: [w.get(k, d) for *_, w in specs]
           for k, d in (("wall_x", 0), ("wall_y", 0), ("res_tan", 0.0))}
    nev = 12000
    refs = []
    for (n, phi, sd, w), st in zip(specs, states):
        o = oracle.Oracle(st, **w, **QUIET)
        code, tr = o.run(nev)
        assert code == 0
        refs.append((tr, o.state()))
    for mode, shape in ((gpu.RES_HBM, 0), (gpu.RES_LDS, 3), (gpu.RES_LDS, 7)):
        b = gpu.Batch(states, per_system=per, trace_capacity=nev, **QUIET)
        b.set_residency(mode)
        b.set_sweep_shape(shape)
        b.sweep_predict()  # the shape under test predicts the starting slots again
        b2 = gpu.Batch(states, per_system=per, trace_capacity=nev, **QUIET)
        for s in range(len(states)):
            for k, v in b.events(s).items():
                np.testing.assert_array_equal(v, b2.events(s)[k], err_msg=k)
        b.run(max_events=nev)
        for s, (tr, so) in enumerate(refs):
            assert_trace_equal(b.trace(s), tr)
            assert_state_equal(b.state(s), so)


def test_wall_arguments_are_checked(gpu):
    st = F.lattice_state(100, 0.4, seed=5, lattice="fit")
    with pytest.raises(gpu.EdmdError):
        gpu.System(st, wall_y=2)
    # flat walls are taped (tests/test_gpu_deriv.py), the circle is not
    b = gpu.Batch([st], wall_y=1)
    b.deriv_begin(checkpoint_every=256)
    b.deriv_end()
    c = in_circle(st)
    b = gpu.Batch([c], wall_circle=1)
    with pytest.raises(gpu.EdmdError):
        b.deriv_begin(checkpoint_every=256)


# ---- the uniform field along y (addField)

FIELD_CASES = [
    (400, 0.4, 51, dict(wall_y=1, res_tan=1.0, add_field=1, field=-0.05)),   # gravity, channel
    (350, 0.5, 52, dict(wall_x=1, wall_y=1, add_field=1, field=-0.1)),       # closed box, resT 0
    (300, 0.5, 53, dict(add_field=1, field=-0.02)),                          # periodic, free fall
]


@pytest.mark.parametrize("n,phi,seed,kw", FIELD_CASES)
def test_field_traces_match_oracle(gpu, oracle, n, phi, seed, kw):
    """Parabolic flights, the neighbours flown (and stored) whenever their
    collision time is taken, quadratic crossing and wall times: traces and
    states bit for bit against the oracle, from HBM and from LDS."""
    st = walled(n, phi, seed, kw)
    nev = 20000
    o = oracle.Oracle(st, **kw, **QUIET)
    code, tr = o.run(nev)
    assert code == 0
    so = o.state()
    for mode in (gpu.RES_HBM, gpu.RES_LDS):
        b = gpu.Batch([st], trace_capacity=nev, **kw, **QUIET)
        b.set_residency(mode)
        b.run(max_events=nev)
        assert_trace_equal(b.trace(0), tr)
        assert_state_equal(b.state(0), so)


def test_field_with_thermostat_and_screenshots(gpu, oracle):
    """Screenshots (everyone flown), thermo rows and velocity rescales under
    the field, in a channel with walls."""
    kw = dict(wall_y=1, res_tan=1.0, add_field=1, field=-0.04, first_screen=0.7, dtime=0.7,
              first_thermo=0.5, dtime_thermo=0.5, noise=2, dtnoise=0.9, temperature=1.2)
    st = walled(450, 0.45, 54, kw)
    nev = 25000
    s = gpu.System(st, trace_capacity=nev, thermo_capacity=512, **kw)
    s.run(max_events=nev)
    o = oracle.Oracle(st, **kw)
    code, tr = o.run(nev)
    assert code == 0
    assert_trace_equal(s.trace(), tr)
    assert_state_equal(s.state(), o.state())
    rows, orows = s.thermo(), o.thermo_rows()
    assert len(rows) == len(orows) > 5
    for r, q in zip(rows, orows):
        for k in ("t", "e_per_n", "pressure", "pwall"):
            assert bits(np.float64(r[k])) == bits(np.float64(q[k])), k


def test_field_arguments_are_checked(gpu):
    st = F.lattice_state(100, 0.4, seed=5, lattice="fit")
    with pytest.raises(gpu.EdmdError):
        gpu.System(st, add_field=1, field=0.1)
    # the field is taped (tests/test_gpu_deriv.py), damping with it is not
    b = gpu.Batch([st], add_field=1)
    b.deriv_begin(checkpoint_every=256)
    b.deriv_end()
    b = gpu.Batch([st], add_field=1, damping=1, gamma=0.01)
    with pytest.raises(gpu.EdmdError):
        b.deriv_begin(checkpoint_every=256)


# ---- damping (damping / gamm)

DAMP_CASES = [
    (400, 0.5, 61, dict(damping=1, gamma=0.02)),                           # periodic
    (350, 0.55, 62, dict(damping=1, gamma=0.05, wall_y=1, res_tan=1.0)),  # channel
    # the field with damping: the field's flight, damped x and collision times
    (380, 0.5, 63, dict(damping=1, gamma=0.03, add_field=1, field=-0.05, wall_y=1, res_tan=1.0)),
    (300, 0.45, 64, dict(damping=1, gamma=0.01, add_field=1, field=-0.02, wall_x=1, wall_y=1)),
]


@pytest.mark.parametrize("n,phi,seed,kw", DAMP_CASES)
def test_damping_traces_match_oracle(gpu, oracle, n, phi, seed, kw):
    """Exponentially slowed flights (the library's own exp), logTime of
    every predicted time (its own log), neighbours flown and stored when
    scanned: traces and states bit for bit against the oracle, from HBM and
    from LDS, every sweep shape giving the same starting slots."""
    st = walled(n, phi, seed, kw)
    nev = 15000
    o = oracle.Oracle(st, **kw, **QUIET)
    oe = o.events()
    code, tr = o.run(nev)
    assert code == 0
    so = o.state()
    # the field with damping pairs damped times with undamped flights, which
    # can end in overlaps and NaN states; these runs stay finite
    assert all(np.all(np.isfinite(so[k])) for k in ("x", "y", "vx", "vy"))
    for mode, shape in ((gpu.RES_HBM, 1), (gpu.RES_LDS, 6)):
        b = gpu.Batch([st], trace_capacity=nev, **kw, **QUIET)
        ev = b.events(0)
        for k in ("t_coll", "t_cross"):
            np.testing.assert_array_equal(bits(ev[k]), bits(oe[k]), err_msg=k)
        b.set_residency(mode)
        b.set_sweep_shape(shape)
        b.run(max_events=nev)
        assert_trace_equal(b.trace(0), tr)
        assert_state_equal(b.state(0), so)


def test_damping_arguments_are_checked(gpu):
    st = F.lattice_state(100, 0.4, seed=5, lattice="fit")
    with pytest.raises(gpu.EdmdError):
        gpu.System(st, damping=1, gamma=0.0)
    # damping is taped (tests/test_gpu_deriv.py), not together with the field
    b = gpu.Batch([st], damping=1)
    b.deriv_begin(checkpoint_every=256)
    b.deriv_end()
    b = gpu.Batch([st], damping=1, add_field=1)
    with pytest.raises(gpu.EdmdError):
        b.deriv_begin(checkpoint_every=256)


# ---- hard walls under a growth start (collisionEventGrow, doTheWallGrow)

GROW_WALL_CASES = [
    (300, 0.55, 71, 0.0, 0.1, dict(wall_x=1)),
    (350, 0.6, 72, 0.1, 0.05, dict(wall_y=1)),
    (400, 0.65, 73, 0.0, 0.1, dict(wall_x=1, wall_y=1, res_tan=1.0)),
    # the field and damping wait for GROWSTOP (the Grow dispatch ignores them)
    (350, 0.55, 74, 0.0, 0.1, dict(wall_y=1, res_tan=1.0, add_field=1, field=-0.05)),
    (300, 0.6, 75, 0.1, 0.1, dict(wall_x=1, wall_y=1, damping=1, gamma=0.02)),
    (320, 0.5, 76, 0.0, 0.1, dict(wall_x=1, wall_y=1, res_tan=1.0, add_field=1, field=-0.03)),
]


@pytest.mark.parametrize("n,phi,seed,poly,vr,kw", GROW_WALL_CASES)
def test_walls_under_growth_match_oracle(gpu, oracle, n, phi, seed, poly, vr, kw):
    """Disks growing from radius 0 in a box with walls: wall slots with the
    surface closing at v -/+ vr, wall hits that add 2 vr, then GROWSTOP and
    the Normal walls (resW, resT, momentum to the walls): traces (mirrored
    pairs unordered, as for every growth start) and states, radii included,
    bit for bit against the oracle, from HBM and from LDS."""
    st = F.growth_state(n, phi, seed, poly=poly)
    nev = 40000
    o = oracle.Oracle(st, vr=vr, **kw, **QUIET)
    code, tr = o.run(nev)
    assert code == 0
    assert np.sum(tr["type"] == WALL) > 20
    assert np.sum(tr["type"] == F.GROWSTOP) == 1
    grown = np.nonzero(tr["type"] == F.GROWSTOP)[0][0]
    assert np.sum(tr["type"][:grown] == WALL) > 5 and np.sum(tr["type"][grown:] == WALL) > 5
    so = o.state()
    for mode in (gpu.RES_HBM, gpu.RES_LDS):
        b = gpu.Batch([st], trace_capacity=nev, vr=vr, **kw, **QUIET)
        b.set_residency(mode)
        b.run(max_events=nev)
        a, c = b.trace(0), tr
        for k in ("type", "i", "j", "valid"):
            np.testing.assert_array_equal(a[k], c[k], err_msg=k)
        np.testing.assert_array_equal(bits(a["t"]), bits(c["t"]))
        assert_state_equal(b.state(0), so)
        np.testing.assert_array_equal(bits(b.radii(0)), bits(o.radii()))
        assert b.observables(0)["ncol"] == o.observables()["ncol"]


# ---- the circular wall (addCircularWall)

def in_circle(st, margin=1.0):
    """The disks of st whose centres lie more than `margin` inside the
    circle of radius Lx/2 about (Lx/2, Lx/2)."""
    c = st["Lx"] / 2
    keep = (st["x"] - c) ** 2 + (st["y"] - c) ** 2 < (c - margin) ** 2
    return {k: (v[keep] if isinstance(v, np.ndarray) else v) for k, v in st.items()}


CIRCLE_CASES = [
    (500, 0.55, 81, 0.0, 0.0, dict()),
    (450, 0.5, 82, 0.0, 0.0, dict(res_wall=0.9)),
    (400, 0.5, 83, 0.0, 0.0, dict(damping=1, gamma=0.01)),
    (450, 0.5, 84, 0.1, 0.1, dict()),                     # growth start through GROWSTOP
    (400, 0.45, 85, 0.0, 0.0, dict(add_field=1, field=-0.05)),    # gravity: the quartic
    (380, 0.45, 86, 0.0, 0.0, dict(add_field=1, field=-0.02, damping=1, gamma=0.01)),
]


@pytest.mark.parametrize("n,phi,seed,poly,vr,kw", CIRCLE_CASES)
def test_circular_wall_matches_oracle(gpu, oracle, n, phi, seed, poly, vr, kw):
    """Disks inside a circular wall: every disk's circle time, hits that
    reflect about the normal at the library's own atan2 / sin / cos (the
    oracle runs the same steps), no minimum image, thermo rows over the
    disc's area with the wall pressure over pi Lx, growth starts through
    GROWSTOP's angular-momentum removal: traces, states and thermo rows bit
    for bit against the oracle, from HBM and from LDS."""
    kw = dict(kw, wall_circle=1, first_screen=1e9, dtime=1e9, first_thermo=0.5,
              dtime_thermo=0.5)
    if vr > 0:
        st = in_circle(F.growth_state(n, phi, seed, poly=poly))
    else:
        st = in_circle(F.lattice_state(n, phi, seed, lattice="fit"))
    nev = 20000
    o = oracle.Oracle(st, vr=vr, **kw)
    code, tr = o.run(nev)
    assert code == 0
    assert np.sum(tr["type"] == WALL) > 100
    if vr > 0:
        assert np.sum(tr["type"] == F.GROWSTOP) == 1
    so, orows = o.state(), o.thermo_rows()
    assert len(orows) > 3 and any(r["pwall"] > 0 for r in orows)
    for mode in (gpu.RES_HBM, gpu.RES_LDS):
        b = gpu.Batch([st], trace_capacity=nev, thermo_capacity=256, vr=vr, **kw)
        b.set_residency(mode)
        b.run(max_events=nev)
        g, c = b.trace(0), tr
        for k in ("type", "i", "j", "valid"):
            np.testing.assert_array_equal(g[k], c[k], err_msg=k)
        np.testing.assert_array_equal(bits(g["t"]), bits(c["t"]))
        assert_state_equal(b.state(0), so)
        rows = b.thermo(0)
        assert len(rows) == len(orows)
        for r, q in zip(rows, orows):
            for k in ("t", "e_per_n", "pressure", "pwall"):
                assert bits(np.float64(r[k])) == bits(np.float64(q[k])), k


def test_circular_wall_arguments_are_checked(gpu):
    st = in_circle(F.lattice_state(200, 0.5, seed=5, lattice="fit"))
    with pytest.raises(gpu.EdmdError):
        gpu.System(st, wall_circle=2)
    with pytest.raises(gpu.EdmdError):
        gpu.System(st, mid_wall=3)


# ---- the mid wall (addMidWall)

def osmosis(n, L, seed, poly=False):
    """Solvent (type 0) everywhere and solute (type 1, a third of the disks
    left of x = L/2) in a closed box, non-overlapping at radius 0.5."""
    rng = np.random.default_rng(seed)
    xs, ys = [], []
    while len(xs) < n:
        x, y = rng.uniform(1.5, L - 1.5), rng.uniform(1.5, L - 1.5)
        if abs(x - L / 2) > 1 and all((x - a) ** 2 + (y - b) ** 2 > 1.2 ** 2
                                      for a, b in zip(xs, ys)):
            xs.append(x)
            ys.append(y)
    xs, ys = np.array(xs), np.array(ys)
    typ = ((xs < L / 2) & (np.arange(n) % 3 == 0)).astype(np.int32)
    v = rng.standard_normal((n, 2))
    return {"x": xs, "y": ys, "vx": v[:, 0].copy(), "vy": v[:, 1].copy(),
            "rad": np.full(n, 0.5), "m": np.ones(n), "Lx": L, "Ly": L, "type": typ}


MID_CASES = [
    (300, 40.0, 91, 0.0, dict(wall_x=1, wall_y=1, res_tan=1.0)),
    (260, 36.0, 92, 0.0, dict(wall_x=1, wall_y=1, res_wall=0.95, damping=1, gamma=0.01)),
    (280, 40.0, 93, 0.1, dict(wall_x=1, wall_y=1, res_tan=1.0)),   # growth start
]


@pytest.mark.parametrize("n,L,seed,vr,kw", MID_CASES)
def test_mid_wall_matches_oracle(gpu, oracle, n, L, seed, vr, kw):
    """The solute stopped at x = Lx/2, the solvent passing, the hits' split
    into dpLeft / dpMid / dpRight and the mid-wall thermo columns: traces,
    states and every thermo column bit for bit against the oracle, from HBM
    and from LDS, with growth through GROWSTOP."""
    st = osmosis(n, L, seed)
    kw = dict(kw, mid_wall=1, sizeratio=0.5, first_screen=1e9, dtime=1e9, first_thermo=1.0,
              dtime_thermo=1.0)
    nev = 20000
    o = oracle.Oracle(st, vr=vr, **kw)
    code, tr = o.run(nev)
    assert code == 0
    so, orows = o.state(), o.thermo_rows()
    assert np.all(so["x"][st["type"] == 1] < L / 2)
    assert len(orows) > 3 and sum(r["pmid"] > 0 for r in orows) >= 2
    for mode in (gpu.RES_HBM, gpu.RES_LDS):
        b = gpu.Batch([st], trace_capacity=nev, thermo_capacity=1024, vr=vr, **kw)
        b.set_residency(mode)
        b.run(max_events=nev)
        # raw order: slow damped disks file collisions beyond the
        # reference's bucket ring, whose overflow list sets the orientation
        # of mirrored pair events when the ring wraps
        g, c = b.trace(0), tr
        for k in ("type", "i", "j", "valid"):
            np.testing.assert_array_equal(g[k], c[k], err_msg=k)
        np.testing.assert_array_equal(bits(g["t"]), bits(c["t"]))
        assert_state_equal(b.state(0), so)
        rows = b.thermo(0)
        assert len(rows) == len(orows)
        for r, q in zip(rows, orows):
            for k in ("t", "e_per_n", "pressure", "pwall", "pmid", "dp_lr", "pack_left",
                      "pack_right", "temp_left", "temp_right", "temp_solute"):
                assert bits(np.float64(r[k])) == bits(np.float64(q[k])), k


# ---- thermal and shear walls (thermoWall, addShearWall)

def _thermal_state(kind, seed):
    if kind == "box":
        return walled(400, 0.45, seed, dict(wall_x=1, wall_y=1))
    if kind == "channel":
        return walled(400, 0.45, seed, dict(wall_y=1))
    if kind == "circle":
        return in_circle(F.lattice_state(450, 0.5, seed, lattice="fit"))
    if kind == "mid":
        return osmosis(280, 40.0, seed)
    if kind == "grow":
        return F.growth_state(350, 0.55, seed, poly=0.0)
    raise ValueError(kind)


THERMAL_CASES = [
    ("box", 101, 0.0, dict(wall_x=1, wall_y=1, temperature=0.6, rng_seed=7)),
    ("channel", 102, 0.0, dict(wall_y=1, shear_wall=1, thermal_wall=0, res_tan=1.0)),
    ("channel", 103, 0.0, dict(wall_y=1, shear_wall=1, temperature=1.5, rng_seed=8)),
    ("circle", 104, 0.0, dict(wall_circle=1, temperature=0.9, rng_seed=9)),
    ("mid", 105, 0.0, dict(wall_x=1, wall_y=1, mid_wall=1, sizeratio=0.5, temperature=1.0,
                            rng_seed=10)),
    ("grow", 106, 0.1, dict(wall_x=1, wall_y=1, temperature=0.8, rng_seed=11)),
]


@pytest.mark.parametrize("kind,seed,vr,kw", THERMAL_CASES)
def test_thermal_and_shear_walls_match_oracle(gpu, oracle, kind, seed, vr, kw):
    """Wall hits that draw the new normal speed from the system's own
    MT19937 stream (twisted wave-parallel on the GPU, sequentially in the
    oracle), the circle's two draws, shear walls driving +-0.1 along x,
    the mid wall's solute and a growth start whose walls turn thermal at
    GROWSTOP: traces, states and thermo rows bit for bit against the
    oracle, from HBM and LDS; then the same after the autotuner's trials,
    which must restore the streams as well."""
    kw = dict(kw, first_screen=1e9, dtime=1e9, first_thermo=1.0, dtime_thermo=1.0)
    kw.setdefault("thermal_wall", 1)
    st = _thermal_state(kind, seed)
    nev = 40000 if vr > 0 else 20000
    o = oracle.Oracle(st, vr=vr, **kw)
    code, tr = o.run(nev)
    assert code == 0
    if vr > 0:
        assert np.sum(tr["type"] == F.GROWSTOP) == 1
    assert np.sum(tr["type"] == WALL) > (60 if kw["thermal_wall"] else 20)
    so, orows = o.state(), o.thermo_rows()
    for mode in (gpu.RES_HBM, gpu.RES_LDS, "tuned"):
        b = gpu.Batch([st], trace_capacity=nev, thermo_capacity=256, vr=vr, **kw)
        if mode == "tuned":
            b.autotune(sweep_reps=1, trial_events=3000)
        else:
            b.set_residency(mode)
        b.run(max_events=nev)
        g, c = b.trace(0), tr
        for k in ("type", "i", "j", "valid"):
            np.testing.assert_array_equal(g[k], c[k], err_msg="%s %s" % (mode, k))
        np.testing.assert_array_equal(bits(g["t"]), bits(c["t"]), err_msg=str(mode))
        assert_state_equal(b.state(0), so)
        rows = b.thermo(0)
        assert len(rows) == len(orows)
        for r, q in zip(rows, orows):
            for k in ("t", "e_per_n", "pressure", "pwall"):
                assert bits(np.float64(r[k])) == bits(np.float64(q[k])), (mode, k)


def test_thermal_and_shear_arguments_are_checked(gpu):
    st = walled(200, 0.4, 5, dict(wall_y=1))
    with pytest.raises(gpu.EdmdError):
        gpu.System(st, thermal_wall=1)             # no walls
    with pytest.raises(gpu.EdmdError):
        gpu.System(st, wall_x=1, shear_wall=1)     # shear walls need wall_y
    with pytest.raises(gpu.EdmdError):
        gpu.System(st, wall_y=1, thermal_wall=2)
    b = gpu.Batch([st], wall_y=1, thermal_wall=1)
    with pytest.raises(gpu.EdmdError):
        b.deriv_begin()
