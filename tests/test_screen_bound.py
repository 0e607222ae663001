"""The sweep's root screen is conservative: whenever it skips a marked
candidate, the root the reference's formula computes for that pair,
(-b - sqrt(b*b - v2*d)) / v2 in FP64 without contraction (EDMD.c:2721),
exceeds the best time already found.

The skip test (the sweep kernel's, for q = -b > 0 and a best time dt > 0)
is d*v2 - 2^-19 q^2 > 2 q v2 dt (1 + 2^-40).  It decreases in dt, so it is
enough that it never holds at dt = the pair's own computed root.  The
samples stress the rounding: grazing pairs (det a few ulps above 0),
pairs a hair from contact (d tiny against q^2/v2), slow relative motion,
and wide ranges of scale."""
import numpy as np


def skip(b, v2, d, dt):
    q = -b
    with np.errstate(over="ignore", invalid="ignore"):
        return (dt > 0) & (dt < 1e26) & (d * v2 - 2.0 ** -19 * (q * q) >
                                         (2 * q * v2 * dt) * (1 + 2.0 ** -40))


def root(b, v2, d):
    with np.errstate(invalid="ignore", divide="ignore"):
        return (-b - np.sqrt(b * b - v2 * d)) / v2


def marked(b, v2, d):
    det = b * b - v2 * d
    return ~(b > 0) & ~(det < 0)


def test_skip_never_holds_at_the_pairs_own_root():
    rng = np.random.default_rng(2024)
    n = 400000
    cases = []
    # generic approaching pairs at cell scale
    dx = rng.uniform(-3, 3, (n, 2))
    dv = rng.standard_normal((n, 2)) * rng.choice([1e-6, 1e-3, 1, 10], (n, 1))
    b = (dx * dv).sum(1)
    v2 = (dv * dv).sum(1)
    d = (dx * dx).sum(1) - 1.0
    cases.append((b, v2, d))
    # grazing: det = b^2 - v2 d within a few ulps of 0
    q = rng.uniform(1e-3, 3, n)
    v2g = rng.uniform(1e-4, 10, n)
    dg = q * q / v2g * (1 - rng.uniform(0, 1e-14, n))
    cases.append((-q, v2g, dg))
    # a hair from contact: d tiny against q^2 / v2
    dh = q * q / v2g * 10.0 ** rng.uniform(-16, -8, n)
    cases.append((-q, v2g, dh))
    checked = 0
    for b, v2, d in cases:
        m = marked(b, v2, d) & (v2 > 0)
        r = root(b[m], v2[m], d[m])
        ok = np.isfinite(r)
        assert not np.any(skip(b[m][ok], v2[m][ok], d[m][ok], r[ok]))
        checked += int(ok.sum())
    assert checked > 600000
    # without the 2^-19 q^2 margin the cancellation in the formula defeats
    # the bound: pairs a hair from contact compute a root below d/(2q)
    b, v2, d = cases[2]
    m = marked(b, v2, d)
    r = root(b[m], v2[m], d[m])
    q = -b[m]
    unmargined = d[m] * v2[m] > 2 * q * v2[m] * r
    assert np.sum(unmargined) > 1000


def test_skip_rejects_most_far_pairs():
    """The screen is useful: against a near partner's root, most of a
    disk's farther marked candidates are skipped."""
    rng = np.random.default_rng(7)
    n = 100000
    dx = rng.uniform(-3, 3, (n, 2))
    dv = rng.standard_normal((n, 2))
    b, v2 = (dx * dv).sum(1), (dv * dv).sum(1)
    d = (dx * dx).sum(1) - 1.0
    m = marked(b, v2, d) & (d > 0)
    r = root(b[m], v2[m], d[m])
    best = np.quantile(r, 0.05)
    s = skip(b[m], v2[m], d[m], np.full(m.sum(), best))
    assert np.all(r[s] > best)
    assert s.mean() > 0.5
