"""The reference's own starts run through the compatibility layer on the
GPU: constantInit / particlesInit / cellListInit / eventListInit from the
golden command lines (random growth starts through GROWSTOP, the text load
path), then the events, against the reference's own traces and final states
(tests/golden/cli_*.npz).  Growth traces are compared with mirrored pair
events as unordered pairs and equal-time runs as sets, as the other growth
vectors; final states and radii bit for bit."""
import ctypes as C

import numpy as np
import pytest

from conftest import bits, load_golden
from test_start_reference import CLI, cli_args, start

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", CLI)
def test_reference_start_runs_like_reference(gpu, name, tmp_path):
    L = gpu.lib()
    g = load_golden(name)
    nev = len(g["trace_t"])
    args = cli_args(g, tmp_path) + ["--no-output", "--trace-capacity", str(nev)]
    assert start(L, args) == 0
    L.cellListInit()
    L.eventListInit()
    assert L.edmd_compat_status() == 0
    L.edmd_compat_context.restype = C.c_void_p
    ctx = C.c_void_p(L.edmd_compat_context())
    gpu.check(L.edmd_run(ctx, nev, 2 ** 64 - 1, 1e300))
    n = len(g["init_x"])
    tr = np.zeros(nev, dtype=gpu.TRACE_DT)
    got_n = C.c_int64()
    gpu.check(L.edmd_get_trace(ctx, tr.ctypes.data, nev, C.byref(got_n)))
    assert got_n.value == nev
    ref = np.zeros(nev, dtype=gpu.TRACE_DT)
    for k in ("type", "i", "j", "valid", "t"):
        ref[k] = g["trace_" + k]
    for k in ("type", "i", "j", "valid"):
        np.testing.assert_array_equal(tr[k], ref[k], err_msg=k)
    np.testing.assert_array_equal(bits(tr["t"]), bits(ref["t"]))
    st = {k: np.zeros(n) for k in ("x", "y", "vx", "vy", "t")}
    gpu.check(L.edmd_get_state(ctx, *[st[k].ctypes.data for k in ("x", "y", "vx", "vy", "t")],
                               None, None, None))
    for k in ("x", "y", "vx", "vy", "t"):
        np.testing.assert_array_equal(bits(st[k]), bits(g["state_" + k]), err_msg=k)
    rad = np.zeros(n)
    gpu.check(L.edmd_get_radii(ctx, rad.ctypes.data))
    np.testing.assert_array_equal(bits(rad), bits(g["state_rad"]))
    L.freeArrays()
