"""Fixture helpers shared by the golden-vector generator and the tests.

Inputs are plain numpy arrays (full FP64 precision).  The LAMMPS dump text
written here is the input format the reference's load path reads
(EDMD.c:799-807, 987-995, 1387-1451; parser.c:360-395, 444-479, 900-917):
the box length is row 0 column 1 of the BOX section and atom columns are
found by name prefix, so ``id type x y vx vy radius m coll`` is read back
exactly when the values are printed with 17 significant digits.
"""
from __future__ import annotations

import math
import os
import struct
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF_DIR = os.path.join(REPO, "oracle", "_ref")

CELLCROSS, COLLISION, SCREENSHOT, THERMO, ADDINGNOISE = 0, 1, 2, 3, 4
GROWSTOP = 9  # enum event, EDMD.h:21-34


def box_length(n, phi, rad=0.5, aspect=1.0):
    """Square-box side for n monodisperse disks of radius ``rad`` at area
    fraction ``phi`` (the reference's sizing, EDMD.c:1150, written for a
    general radius)."""
    return math.sqrt(math.pi * n * rad * rad / (phi * aspect))


def lattice_state(n, phi, seed, rad=0.5, lattice="square", temperature=1.0):
    """Seeded initial state: lattice positions, Gaussian velocities with zero
    total momentum scaled to E/N = temperature (the reference's
    normalizePhysicalQ convention, EDMD.c:5723-5764, where E = sum m v^2 / 2).
    Returns a dict of numpy arrays plus the box length."""
    rng = np.random.default_rng(seed)
    if lattice == "square":
        L = box_length(n, phi, rad)
        side = int(math.ceil(math.sqrt(n)))
        a = L / side
        k = np.arange(n)
        x = (k % side + 0.5) * a
        y = (k // side + 0.5) * a
        Lx = Ly = L
    elif lattice == "tri":
        # Triangular lattice as in the reference's Hex start (EDMD.c:1024-1049,
        # 1582-1637): Ny rows (even), Nx columns, alternate rows shifted.
        ny = int(math.sqrt(n))
        if ny % 2 == 1:
            ny += 1
        nx = n // ny
        if nx * ny != n:
            raise ValueError("triangular lattice needs n = nx*ny")
        Ly = math.sqrt(math.sqrt(3) / 2 * ny / nx * math.pi * n * rad * rad / phi)
        Lx = 2 / math.sqrt(3) * nx / ny * Ly
        dx = Lx / nx
        dy = Ly / ny
        i = np.repeat(np.arange(nx), ny)
        j = np.tile(np.arange(ny), nx)
        x = i * dx + (j % 2) * (dx / 2) + dx / 4
        y = j * dy + dy / 2
        L = None
    elif lattice == "fit":
        # triangular stacking fitted to a square box for any n: even row count
        # chosen for the largest nearest-neighbour distance
        L = box_length(n, phi, rad)
        best = None
        for ny in range(2, 2 * int(math.ceil(math.sqrt(n))) + 5, 2):
            nx = -(-n // ny)
            dx, dy = L / nx, L / ny
            d = min(dx, math.sqrt(dx * dx / 4 + dy * dy), 2 * dy)
            if best is None or d > best[0]:
                best = (d, nx, ny)
        d, nx, ny = best
        if d < 2 * rad:
            raise ValueError("fitted lattice overlaps")
        dx, dy = L / nx, L / ny
        k = np.arange(n)
        i, j = k % nx, k // nx
        x = (i + 0.25 + 0.5 * (j % 2)) * dx
        y = (j + 0.5) * dy
        Lx = Ly = L
    else:
        raise ValueError(lattice)
    if lattice == "square" and L / int(math.ceil(math.sqrt(n))) < 2 * rad:
        raise ValueError("square lattice overlaps at phi=%g" % phi)
    v = rng.standard_normal((n, 2))
    v -= v.mean(axis=0)
    e_per = 0.5 * (v * v).sum() / n
    v *= math.sqrt(temperature / e_per)
    st = {
        "x": np.ascontiguousarray(x, dtype=np.float64),
        "y": np.ascontiguousarray(y, dtype=np.float64),
        "vx": np.ascontiguousarray(v[:, 0]),
        "vy": np.ascontiguousarray(v[:, 1]),
        "rad": np.full(n, rad, dtype=np.float64),
        "m": np.ones(n, dtype=np.float64),
        "Lx": float(Lx if L is None else L),
        "Ly": float(Ly if L is None else L),
    }
    return st


def walled_state(n, phi, seed, wall_x=True, wall_y=False, temperature=1.0):
    """A fitted-lattice start pressed affinely into [r + 0.01, L - r - 0.01]
    along each walled direction (the lattice's outer rows may sit closer
    than a radius to the box edge; the squeeze leaves every pair apart)."""
    st = lattice_state(n, phi, seed=seed, lattice="fit", temperature=temperature)
    r, L = 0.5, st["Lx"]
    lo, span = r + 0.01, L - 2 * (r + 0.01)
    if wall_x:
        st["x"] = lo + st["x"] / L * span
    if wall_y:
        st["y"] = lo + st["y"] / L * span
    return st


def growth_state(n, phi, seed, poly=0.0, rad=0.5, e_per=0.05):
    """Seeded growth start (the reference's default start, EDMD.c:1691-1810,
    with this build's generator): uniform random centres in a square box
    sized for packing fraction phi at the target radii, Gaussian velocities
    with E/N = e_per (EinitGrow = 0.05), masses 1.  poly > 0 draws target
    diameters 2 rad * N(1, poly) truncated to [1 - 3 poly, 1 + 3 poly] and
    masses (d / 2 rad)^3.  The radii in the returned state are the targets."""
    rng = np.random.default_rng(seed)
    if poly > 0:
        d = np.empty(0)
        while len(d) < n:
            c = rng.normal(1.0, poly, 2 * n)
            d = np.concatenate([d, c[(c >= 1 - 3 * poly) & (c <= 1 + 3 * poly)]])
        d = d[:n]
        r = rad * d
        m = d ** 3
    else:
        r = np.full(n, rad)
        m = np.ones(n)
    L = math.sqrt(math.pi * float(np.sum(r * r)) / phi)
    x = rng.uniform(0, L, n)
    y = rng.uniform(0, L, n)
    v = rng.standard_normal((n, 2)) * math.sqrt(e_per)
    return {"x": x, "y": y, "vx": np.ascontiguousarray(v[:, 0]),
            "vy": np.ascontiguousarray(v[:, 1]), "rad": r, "m": m, "Lx": L, "Ly": L}


def scale_state(st, f):
    """Same system with all lengths multiplied by f (velocities unchanged):
    collision times scale by f, so event order is unchanged."""
    out = dict(st)
    for k in ("x", "y", "rad"):
        out[k] = st[k] * f
    out["Lx"] = st["Lx"] * f
    out["Ly"] = st["Ly"] * f
    return out


def write_dump(path, st):
    n = len(st["x"])
    lines = [
        "ITEM: TIMESTEP", "0", "ITEM: NUMBER OF ATOMS", str(n),
        "ITEM: BOX BOUNDS pp pp pp",
        "0 %.17g" % st["Lx"], "0 %.17g" % st["Ly"], "0 0",
        "ITEM: ATOMS id type x y vx vy radius m coll",
    ]
    for i in range(n):
        lines.append("%d 1 %.17g %.17g %.17g %.17g %.17g %.17g 0" % (
            i, st["x"][i], st["y"][i], st["vx"][i], st["vy"][i],
            st["rad"][i], st["m"][i]))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_dump(path):
    """Reads the last frame of a dump written by write_dump or by the
    library's writer (columns located by name, as parser.c:900-917)."""
    with open(path) as f:
        text = f.read().split("\n")
    starts = [k for k, l in enumerate(text) if l.startswith("ITEM: TIMESTEP")]
    k = starts[-1]
    n = int(text[k + 3])
    Lx = float(text[k + 5].split()[1])
    Ly = float(text[k + 6].split()[1])
    cols = text[k + 8].split()[2:]
    rows = [text[k + 9 + i].split() for i in range(n)]

    def col(name):
        idx = next(c for c, nm in enumerate(cols) if nm.startswith(name))
        return np.array([float(r[idx]) for r in rows])
    return {"x": col("x"), "y": col("y"), "vx": col("vx"), "vy": col("vy"),
            "rad": col("radius"), "m": col("m"), "Lx": Lx, "Ly": Ly}


# ---------------------------------------------------------------- harness I/O

def ref_binary(kind="strict"):
    return os.path.join(REF_DIR, "ref_" + kind)


def run_ref(args, kind="strict", timeout=600):
    return subprocess.run([ref_binary(kind)] + [str(a) for a in args],
                          check=True, capture_output=True, text=True,
                          timeout=timeout)


SWEEP_REC = np.dtype([("t_cross", "<f8"), ("t_coll", "<f8"), ("dir", "<i4"),
                      ("partner", "<i4"), ("coll_actual", "<u8"),
                      ("cx", "<i4"), ("cy", "<i4")])
TRACE_REC = np.dtype([("type", "<i4"), ("i", "<i4"), ("j", "<i4"),
                      ("valid", "<i4"), ("t", "<f8")])
STATE_REC = np.dtype([("x", "<f8"), ("y", "<f8"), ("vx", "<f8"), ("vy", "<f8"),
                      ("t", "<f8"), ("coll", "<u8"), ("cx", "<i4"),
                      ("cy", "<i4")])


def parse_sweep(path):
    with open(path, "rb") as f:
        raw = f.read()
    n, = struct.unpack_from("<i", raw, 0)
    Lx, Ly = struct.unpack_from("<dd", raw, 4)
    ncx, ncy = struct.unpack_from("<ii", raw, 20)
    rec = np.frombuffer(raw, dtype=SWEEP_REC, count=n, offset=28)
    return {"N": n, "Lx": Lx, "Ly": Ly, "ncx": ncx, "ncy": ncy, "rec": rec.copy()}


def parse_trace(path):
    with open(path, "rb") as f:
        raw = f.read()
    nev, = struct.unpack_from("<q", raw, 0)
    off = 8
    tr = np.frombuffer(raw, dtype=TRACE_REC, count=nev, offset=off).copy()
    off += nev * TRACE_REC.itemsize
    n, = struct.unpack_from("<i", raw, off)
    off += 4
    state = np.frombuffer(raw, dtype=STATE_REC, count=n, offset=off).copy()
    off += n * STATE_REC.itemsize
    ncol, ncross = struct.unpack_from("<QQ", raw, off)
    off += 16
    g = struct.unpack_from("<8d", raw, off)
    glob = dict(zip(["t", "virXX", "virYY", "virXY", "virYX", "E", "px", "py"], g))
    glob["ncol"] = ncol
    glob["ncross"] = ncross
    return {"trace": tr, "state": state, "glob": glob}


def parse_grow(path):
    """parse_trace plus the final radius of every disk."""
    out = parse_trace(path)
    with open(path, "rb") as f:
        raw = f.read()
    n = len(out["state"])
    out["rad"] = np.frombuffer(raw, dtype="<f8", count=n, offset=len(raw) - 8 * n).copy()
    return out


def parse_noise(path):
    """parse_grow plus E right after every velocity rescale."""
    out = parse_trace(path)
    with open(path, "rb") as f:
        raw = f.read()
    n = len(out["state"])
    off = 8 + len(out["trace"]) * TRACE_REC.itemsize + 4 + n * STATE_REC.itemsize + 16 + 64
    out["rad"] = np.frombuffer(raw, dtype="<f8", count=n, offset=off).copy()
    off += 8 * n
    nr, = struct.unpack_from("<q", raw, off)
    out["rescale_E"] = np.frombuffer(raw, dtype="<f8", count=nr, offset=off + 8).copy()
    return out


def parse_cli(path):
    """The cli mode's output: the box, the growth rate and the disks right
    after particlesInit, then what parse_grow reads."""
    with open(path, "rb") as f:
        raw = f.read()
    n, = struct.unpack_from("<i", raw, 0)
    Lx, Ly, vr = struct.unpack_from("<3d", raw, 4)
    init = np.frombuffer(raw, dtype="<f8", count=7 * n, offset=28).reshape(n, 7).copy()
    off = 28 + 56 * n
    tmp = path + ".rest"
    with open(tmp, "wb") as f:
        f.write(raw[off:])
    out = parse_grow(tmp)
    os.remove(tmp)
    out.update(Lx=Lx, Ly=Ly, vr=vr,
               init={k: init[:, c].copy() for c, k in
                     enumerate(("x", "y", "vx", "vy", "rad", "m", "vr"))})
    return out


def write_text_state(path, st, types):
    """The reference's text start format (EDMD.c:980-986, 1366-1386): a line
    "N Lx Ly", then "type x y rad m" per disk, values at 17 digits."""
    with open(path, "w") as f:
        f.write("%d %.17g %.17g\n" % (len(st["x"]), st["Lx"], st["Ly"]))
        for k in range(len(st["x"])):
            f.write("%d %.17g %.17g %.17g %.17g\n" % (types[k], st["x"][k], st["y"][k],
                                                     st["rad"][k], st["m"][k]))


def parse_pcf(path):
    with open(path, "rb") as f:
        raw = f.read()
    nb, = struct.unpack_from("<i", raw, 0)
    r = np.frombuffer(raw, dtype="<f8", count=nb, offset=4).copy()
    g = np.frombuffer(raw, dtype="<f8", count=nb, offset=4 + 8 * nb).copy()
    return r, g


def bits_equal(a, b):
    return np.float64(a).view(np.uint64) == np.float64(b).view(np.uint64)
