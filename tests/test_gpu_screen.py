"""The sweep's root screen on the GPU: the same predictions with it on as
off, over every sweep shape, monodisperse and polydisperse systems, tiles
that wrap and tiles that do not, walls; no misses when the skipped roots
are solved after all; and the fraction of roots it saves."""
import numpy as np
import pytest

import edmd_fixtures as F
from conftest import bits, golden_state, load_golden

pytestmark = pytest.mark.gpu

EV_FIELDS = ("t_coll", "t_cross", "partner", "dir")


def _events(b, k):
    e = b.events(k)
    return {f: e[f].copy() for f in EV_FIELDS}


def _same(a, c, label):
    for f in EV_FIELDS:
        x, y = a[f], c[f]
        if x.dtype.kind == "f":
            np.testing.assert_array_equal(bits(x), bits(y), err_msg="%s %s" % (label, f))
        else:
            np.testing.assert_array_equal(x, y, err_msg="%s %s" % (label, f))


def test_screen_leaves_every_prediction_unchanged(gpu):
    states = [F.lattice_state(4096, 0.80, seed=3, lattice="tri"),
              golden_state(load_golden("poly500_phi050")),
              F.lattice_state(2000, 0.70, seed=4, lattice="fit"),
              F.lattice_state(400, 0.60, seed=5, lattice="fit")]
    b = gpu.Batch(states)
    b.run(max_events=20000)  # away from the lattice
    nshape = 8
    for shape in range(nshape):
        b.set_sweep_shape(shape)
        b.set_screen(0)
        b.sweep_predict()
        ref = [_events(b, k) for k in range(len(states))]
        b.set_screen(2)
        b.sweep_predict()
        marks, solved, misses = b.screen_stats()
        assert misses == 0 and 0 < solved < marks, (shape, marks, solved)
        for k in range(len(states)):
            _same(_events(b, k), ref[k], "shape %d system %d" % (shape, k))
        b.set_screen(1)
        b.sweep_predict()
        for k in range(len(states)):
            _same(_events(b, k), ref[k], "shape %d system %d" % (shape, k))


def test_screen_on_the_large_system_and_with_walls(gpu, oracle):
    """N = 2^20 at phi = 0.80: identical predictions, the fraction of roots
    left to solve; a walled system continues exactly as the oracle after
    a screened sweep."""
    big = gpu.Batch([F.lattice_state(1 << 20, 0.80, seed=1, lattice="tri")])
    big.run(max_events=20000)
    big.sweep_predict()
    ref = _events(big, 0)
    big.set_screen(2)
    big.sweep_predict()
    marks, solved, misses = big.screen_stats()
    assert misses == 0
    assert solved < 0.6 * marks, (marks, solved)
    _same(_events(big, 0), ref, "N=2^20")
    st = F.walled_state(400, 0.55, 9, wall_x=True, wall_y=True)
    kw = dict(wall_x=1, wall_y=1, res_wall=0.9)
    b = gpu.Batch([st], trace_capacity=20000, **kw)
    b.set_screen(1)
    b.sweep_predict()
    b.run(max_events=20000)
    o = oracle.Oracle(st, **kw)
    code, tr = o.run(20000)
    g = b.trace(0)
    for f in ("type", "i", "j", "valid"):
        np.testing.assert_array_equal(g[f], tr[f])
    np.testing.assert_array_equal(bits(g["t"]), bits(tr["t"]))
