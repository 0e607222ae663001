"""CPU-side checks of libedmd_hip.so: the C ABI exports every entry point its
headers declare, failures are reported as status codes (never a fallback),
and the host-only pieces (lattice starts, LAMMPS dump I/O, option parsing)
behave as specified.  No kernel runs here."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import edmd_hip as E

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(REPO, "graphical-edmd_amd", "bin", "edmd")


def declared_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set()
    for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text, re.M):
        line = m.group(0)
        if line.startswith(("typedef", "#", "extern void (*")):
            continue
        names.add(m.group(1))
    return names


def test_library_exports_every_declared_entry_point():
    L = E.lib()
    missing = []
    for h in E.HEADER_PATHS:
        names = declared_functions(h)
        assert len(names) > 10
        for n in sorted(names):
            if not hasattr(L, n):
                missing.append(n)
    assert not missing, missing


def test_compat_globals_exported():
    L = E.lib()
    assert C.c_int.in_dll(L, "N").value == 500  # the reference's default, EDMD.c:80
    assert C.c_double.in_dll(L, "phi").value == 0.38
    assert C.c_double.in_dll(L, "tmax").value == 4000
    for name in ("Lx", "Ly", "t", "E", "px", "py", "pressure", "res"):
        C.c_double.in_dll(L, name)
    for name in ("ncol", "ncross"):
        C.c_ulong.in_dll(L, name)
    # eventList (a node** set by eventListInit) and the dispatch pointers
    assert not C.c_void_p.in_dll(L, "eventList").value
    for name in ("freeFly", "collisionEvent", "crossingEvent", "doTheCollision"):
        assert C.c_void_p.in_dll(L, name).value


@pytest.mark.skipif(E.device_count() > 0, reason="checks the no-device path")
def test_no_device_is_an_error_not_a_fallback():
    st = E.init_lattice(64, 0.5)
    with pytest.raises(E.EdmdError) as ei:
        E.System(st)
    assert ei.value.code == 19
    assert "no HIP device" in str(ei.value)


@pytest.mark.parametrize("lattice,n", [(0, 1000), (0, 4096), (1, 1024), (1, 4096), (2, 257),
                                       (2, 2048), (2, 1111)])
def test_lattice_start(lattice, n):
    phi, rad, temp = (0.75 if lattice == 2 else 0.7), 0.5, 1.3
    st = E.init_lattice(n, phi, rad=rad, lattice=lattice, seed=9, temperature=temp)
    vx, vy = st["vx"], st["vy"]
    assert abs(vx.sum()) < 1e-12 * n and abs(vy.sum()) < 1e-12 * n
    assert abs(0.5 * np.sum(vx * vx + vy * vy) / n - temp) < 1e-12
    area = st["Lx"] * st["Ly"]
    assert abs(n * np.pi * rad * rad / area - phi) < 1e-12
    x, y = st["x"], st["y"]
    assert x.min() >= 0 and x.max() < st["Lx"] and y.min() >= 0 and y.max() < st["Ly"]
    # nearest-neighbour distance through the periodic box
    from scipy.spatial import cKDTree
    tree = cKDTree(np.c_[x, y], boxsize=[st["Lx"], st["Ly"]])
    d, _ = tree.query(np.c_[x, y], k=2)
    assert d[:, 1].min() > 2 * rad
    st2 = E.init_lattice(n, phi, rad=rad, lattice=lattice, seed=9, temperature=temp)
    assert np.array_equal(st2["vx"], vx)  # seeded
    st3 = E.init_lattice(n, phi, rad=rad, lattice=lattice, seed=10, temperature=temp)
    assert not np.array_equal(st3["vx"], vx)


@pytest.mark.parametrize("kw", [dict(poly=0.1), dict(rad=1.0, sizeratio=0.4, fsmall=0.3),
                                dict(poly=0.0)])
def test_random_growth_start(kw):
    """Target radii fill the box to phi exactly; truncated Gaussian
    diameters or the binary mixture's species counts; centres inside the
    box; velocity components of variance e_per / m; seeded."""
    n, phi = 4000, 0.8
    st = E.init_random(n, phi, seed=5, e_per=0.05, **kw)
    L = st["Lx"]
    assert st["Ly"] == L
    assert abs(np.pi * np.sum(st["rad"] ** 2) / (L * L) - phi) < 1e-12
    assert st["x"].min() >= 0 and st["x"].max() < L and st["y"].min() >= 0 and st["y"].max() < L
    rad = kw.get("rad", 0.5)
    if kw.get("poly", 0) > 0:
        d = st["rad"] / rad
        assert d.min() >= 0.7 and d.max() <= 1.3
        assert abs(d.mean() - 1) < 0.01 and abs(d.std() - 0.1) < 0.01
        np.testing.assert_allclose(st["m"], d ** 3)
    elif kw.get("fsmall", 0) > 0:
        nsmall = int(n * kw["fsmall"])
        assert np.all(st["rad"][: n - nsmall] == rad)
        assert np.all(st["rad"][n - nsmall:] == rad * kw["sizeratio"])
        assert np.all(st["m"] == 1)
    # <m v_x^2> = e_per per component
    for c in ("vx", "vy"):
        assert abs(np.mean(st["m"] * st[c] ** 2) / 0.05 - 1) < 0.06
    st2 = E.init_random(n, phi, seed=5, e_per=0.05, **kw)
    assert np.array_equal(st2["x"], st["x"]) and np.array_equal(st2["vy"], st["vy"])
    st3 = E.init_random(n, phi, seed=6, e_per=0.05, **kw)
    assert not np.array_equal(st3["x"], st["x"])


def test_growth_rate_follows_reference_rule():
    """optimizeGrowConstant (EDMD.c:5766-5800) for a periodic box."""
    assert E.growth_rate(0.5) == 0.1
    assert E.growth_rate(0.7) == 0.1
    assert E.growth_rate(0.75) == 0.1 / 8
    assert E.growth_rate(0.8) == 0.1 / 8
    assert E.growth_rate(0.84) == 0.1 * (0.8 / 0.84) ** 30 / 8


def test_lattice_rejects_impossible_packing():
    with pytest.raises(E.EdmdError) as ei:
        E.init_lattice(100, 0.85, rad=0.5, lattice=0)  # square lattice overlaps above pi/4
    assert ei.value.code == 22


def test_dump_round_trip(tmp_path):
    L = E.lib()
    st = E.init_lattice(300, 0.6, rad=0.5, seed=4)
    st["rad"] = np.linspace(0.4, 0.6, 300)
    p = str(tmp_path / "a.dump").encode()
    args = [st[k].ctypes.data for k in ("x", "y", "vx", "vy", "rad", "m")]
    E.check(L.edmd_dump_write(p, 0, 1.5, 300, st["Lx"], st["Ly"], *args, 1))
    n, lx, ly = C.c_int(), C.c_double(), C.c_double()
    E.check(L.edmd_dump_info(p, C.byref(n), C.byref(lx), C.byref(ly)))
    assert (n.value, lx.value, ly.value) == (300, st["Lx"], st["Ly"])
    back = {k: np.zeros(300) for k in ("x", "y", "vx", "vy", "rad", "m")}
    E.check(L.edmd_dump_read(p, 300, *[back[k].ctypes.data for k in
                                       ("x", "y", "vx", "vy", "rad", "m")]))
    for k in back:
        assert np.array_equal(back[k], st[k]), k
    # the reference's own layout (x printed with 3 decimals, EDMD.c:5124),
    # appended as a second frame: the reader takes the last frame
    E.check(L.edmd_dump_write(p, 1, 2.5, 300, st["Lx"], st["Ly"], *args, 0))
    E.check(L.edmd_dump_read(p, 300, *[back[k].ctypes.data for k in
                                       ("x", "y", "vx", "vy", "rad", "m")]))
    assert np.max(np.abs(back["x"] - st["x"])) <= 5e-4
    assert np.max(np.abs(back["vx"] - st["vx"])) <= 5e-7


def test_dump_reader_matches_fixture_writer(tmp_path):
    """Dumps in the layout the reference loads (tests/golden/edmd_fixtures)."""
    import edmd_fixtures as F
    st = F.lattice_state(500, 0.65, seed=3)
    path = str(tmp_path / "init.dump")
    F.write_dump(path, st)
    n, lx, ly = C.c_int(), C.c_double(), C.c_double()
    E.check(E.lib().edmd_dump_info(path.encode(), C.byref(n), C.byref(lx), C.byref(ly)))
    assert n.value == 500 and lx.value == st["Lx"]
    back = {k: np.zeros(500) for k in ("x", "y", "vx", "vy", "rad", "m")}
    E.check(E.lib().edmd_dump_read(path.encode(), 500, *[back[k].ctypes.data for k in
                                                          ("x", "y", "vx", "vy", "rad", "m")]))
    for k in back:
        assert np.array_equal(back[k], st[k])


def test_dump_errors(tmp_path):
    L = E.lib()
    n = C.c_int()
    assert L.edmd_dump_info(str(tmp_path / "missing.dump").encode(), C.byref(n), None, None) == 5
    bad = tmp_path / "bad.dump"
    bad.write_text("not a dump\n")
    assert L.edmd_dump_info(str(bad).encode(), C.byref(n), None, None) == 5


def test_ctx_from_dump_reports_io_errors(tmp_path):
    """edmd_ctx_create_from_dump reads the file before touching the GPU."""
    with pytest.raises(E.EdmdError) as ei:
        E.System.from_dump(tmp_path / "missing.dump")
    assert ei.value.code == 5
    L = E.lib()
    assert L.edmd_ctx_create_from_dump(str(tmp_path / "missing.dump").encode(), None, None) == 22
    assert L.edmd_run_traced(None, 10, 10, 1.0, None, None) == 22
    assert L.edmd_load_dump(str(tmp_path / "missing.dump").encode(), None) == 22
    h = C.c_void_p()
    assert L.edmd_load_dump(str(tmp_path / "missing.dump").encode(), C.byref(h)) == 5
    assert not h.value
    assert L.edmd_adjoint_begin(None, 0, 0) == 22
    assert L.edmd_adjoint_run(None, 10, 10, 1.0) == 22
    assert L.edmd_adjoint_tangent(None, None, None, None, None) == 22
    assert L.edmd_adjoint_gradient(None, 1.0, 0.0, None, None) == 22
    assert L.edmd_adjoint_values(None, None, None, None) == 22
    assert L.edmd_adjoint_end(None) == 22


@pytest.mark.skipif(E.device_count() > 0, reason="checks the no-device path")
def test_cli_reports_missing_device():
    r = subprocess.run([CLI, "-N", "100", "-p", "0.5", "-x", "0", "-t", "10", "--no-output"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 19
    assert "no HIP device" in r.stderr




def test_python_mirrors_match_the_c_header(tmp_path):
    """The ctypes mirrors of the public structs (descriptor, thermo row,
    observables, tune report) have the C header's sizes and field offsets:
    compiled with the host C compiler, no GPU needed."""
    import ctypes as C
    fields = {
        "edmd_system_desc": (E.SystemDesc, ["n", "Lx", "x", "res", "noise", "cell_capacity",
                                            "trace_capacity", "vr", "einit", "wall_x", "wall_y",
                                            "res_wall", "res_tan", "add_field", "field",
                                            "damping", "gamma"]),
        "edmd_observables": (E.Observables, ["t", "ncol", "E", "py", "status", "err_j"]),
        "edmd_tune_report": (E.TuneReport, ["sweep_ms", "sweep_shape", "in_lds", "loop_ms_hbm",
                                            "loop_ms_lds"]),
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "edmd.h"', 'int main(void) {']
    for name, (_, fl) in fields.items():
        lines.append('printf("%%s %%zu\\n", "%s", sizeof(%s));' % (name, name))
        for f in fl:
            lines.append('printf("%%s.%%s %%zu\\n", "%s", "%s", offsetof(%s, %s));' % (name, f, name, f))
    lines.append('printf("thermo %zu pwall %zu\\n", sizeof(edmd_thermo_row), '
                 'offsetof(edmd_thermo_row, pwall));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(REPO, "include"), "-o", str(exe),
                    str(src)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = dict(l.rsplit(" ", 1) for l in out if l and not l.startswith("thermo"))
    for name, (cls, fl) in fields.items():
        assert int(got[name]) == C.sizeof(cls), name
        for f in fl:
            assert int(got["%s.%s" % (name, f)]) == getattr(cls, f).offset, (name, f)
    th = [l for l in out if l.startswith("thermo")][0].split()
    assert int(th[1]) == E.THERMO_DT.itemsize and int(th[3]) == E.THERMO_DT.fields["pwall"][1]
