"""Long-run observables against the reference built with its shipped flags.

The square-lattice system of N = 2000, diameter 1, phi = 0.70 runs to t = 60, about
1.3 x 10^6 valid collisions, on the GPU and in the reference compiled with
-O3 -ffast-math (tests/golden/make_golden.py, case sq2000_phi070_longrun).  That
build's trajectory leaves the strict one after ~10^4 collisions, so the two
runs are independent samples of the same equilibrium beyond the first few
time units: the reduced pressure (thermo rows every time unit,
saveThermo, EDMD.c:5232-5333) and g(r) (pcf.c:16-75, averaged over the
states at t = 30, 40, 50, 60) must agree within their statistical errors.
"""
import numpy as np
import pytest

from conftest import golden_state, load_golden

pytestmark = pytest.mark.gpu


def synced(x, y, vx, vy, tl, T, L):
    """Positions free-flown to the common time T (freeFlyNormal,
    EDMD.c:4954-4990), folded into the box."""
    return np.mod(x + (T - tl) * vx, L), np.mod(y + (T - tl) * vy, L)


@pytest.fixture(scope="module")
def runs(gpu):
    g = load_golden("sq2000_phi070_longrun")
    st = golden_state(g)
    s = gpu.System(st, first_thermo=1.0, dtime_thermo=1.0, thermo_capacity=128,
                   first_screen=1e9, dtime=1e9)
    L = st["Lx"]
    gg, gr = [], []
    for k, tm in enumerate(g["snap_tmax"]):
        s.run(tmax=float(tm))
        sg = s.state()
        T = s.observables()["t"]
        x, y = synced(sg["x"], sg["y"], sg["vx"], sg["vy"], sg["t"], T, L)
        r, gk = gpu.pcf(x, y, L, L, 0.1, 5.0)
        gg.append(gk)
        x, y = synced(g["snap%d_x" % k], g["snap%d_y" % k], g["snap%d_vx" % k],
                      g["snap%d_vy" % k], g["snap%d_t" % k], float(g["snap%d_time" % k]), L)
        r, gk = gpu.pcf(x, y, L, L, 0.1, 5.0)
        gr.append(gk)
    return g, s, r, np.mean(gg, axis=0), np.mean(gr, axis=0), len(gg), st


def test_reduced_pressure_agrees(runs):
    g, s, r, _, _, _, st = runs
    ref = g["thermo_rows"]
    rows = s.thermo()
    assert len(rows) == len(ref) == 60
    # collision counts climb at the same rate; both conserve E/N = 1
    assert abs(rows["ncol"][-1] / ref[-1, 1] - 1) < 0.02
    np.testing.assert_allclose(rows["e_per_n"], 1.0, rtol=1e-12)
    # equilibrium rows (t > 10): means within 4 combined standard errors
    pg, pr = rows["pressure"][10:], ref[10:, 3]
    se = np.sqrt(pg.var(ddof=1) / len(pg) + pr.var(ddof=1) / len(pr))
    assert abs(pg.mean() - pr.mean()) < 4 * se, (pg.mean(), pr.mean(), se)


def test_pair_correlation_agrees(runs):
    g, s, r, gg, gr, K, st = runs
    n = len(st["x"])
    rho = n / (st["Lx"] * st["Ly"])
    dr = 0.1
    # expected pair count per unit g in the shell around the bin centre r
    # (calculate_pcf's normalisation: g = 2 count / (2 pi r dr rho N))
    mu = np.pi * r * dr * rho * n
    sel = (r >= 0.9) & ((gg > 0) | (gr > 0))
    var = (gg[sel] + gr[sel]) / (K * mu[sel])
    chi2 = np.sum((gg[sel] - gr[sel]) ** 2 / var)
    assert chi2 < 2.5 * sel.sum(), (chi2, sel.sum())
    # no pair inside contact (bins ending below a diameter), and the contact
    # peak in the same bin
    inside = r + dr / 2 <= 1.0 - 1e-9
    assert np.all(gg[inside] == 0) and np.all(gr[inside] == 0)
    assert np.argmax(gg) == np.argmax(gr)
