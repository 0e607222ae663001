"""The host-side calendar consistency test used by the GPU tests: it accepts a
calendar built by the rule it states and flags planted mistakes."""
import numpy as np

from calendar_check import check_calendar, level_sizes


def build(ev, slot_rank):
    n = len(ev["t_coll"])
    tk, tc = ev["t_coll"], ev["t_cross"]
    rk, rx = slot_rank[0::2], slot_rank[1::2]
    coll = (tk < tc) | ((tk == tc) & (rk < rx))
    key = list(np.where(coll, tk, tc))
    idx = list(range(n))
    aux = list(np.where(coll, ev["partner"], -ev["dir"]))
    lr = np.where(coll, rk, rx)
    sizes = level_sizes(n)
    base = 0
    for l in range(1, len(sizes)):
        for g in range(sizes[l]):
            ch = range(base + 64 * g, base + min(64 * g + 64, sizes[l - 1]))
            best = min(ch, key=lambda c: (key[c], lr[idx[c]] if key[c] < np.inf else 0, c))
            key.append(key[best]); idx.append(idx[best]); aux.append(aux[best])
        base += sizes[l - 1]
    return {"key": np.array(key), "idx": np.array(idx, dtype=np.int32),
            "aux": np.array(aux, dtype=np.int32), "slot_rank": slot_rank, "leaf_rank": lr}


def events(n, seed):
    rng = np.random.default_rng(seed)
    # coarse times so that equal keys (ties decided by rank) are frequent
    ev = {"t_coll": rng.integers(0, 50, n) / 8.0, "t_cross": rng.integers(0, 50, n) / 8.0,
          "partner": rng.integers(0, n, n).astype(np.int32),
          "dir": rng.integers(1, 5, n).astype(np.int32)}
    ev["t_coll"][rng.random(n) < 0.05] = np.inf
    slot_rank = rng.permutation(2 * n).astype(np.int64)
    return ev, slot_rank


def test_accepts_consistent_calendar():
    for n, seed in ((50, 1), (2000, 2), (5000, 3)):
        ev, sr = events(n, seed)
        assert check_calendar(build(ev, sr), ev) == []


def test_flags_mixed_node_and_wrong_tie():
    ev, sr = events(2000, 4)
    cal = build(ev, sr)
    n = 2000
    # a node that takes the key and leaf of one child and the event of another
    bad = dict(cal, aux=cal["aux"].copy())
    bad["aux"][n + 5] = cal["aux"][n + 5] + 1
    assert any("level 1 node 5" in s for s in check_calendar(bad, ev))
    # a leaf whose event is the later one on an exact tie in time
    p = int(np.nonzero(ev["t_coll"] == ev["t_cross"])[0][0])
    sr2 = sr.copy()
    sr2[2 * p], sr2[2 * p + 1] = sr[2 * p + 1], sr[2 * p]
    assert check_calendar(dict(cal, slot_rank=sr2), ev)
