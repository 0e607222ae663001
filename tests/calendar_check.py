"""Consistency of the GPU event calendar, checked on the host.

Every leaf must hold the earlier of its disk's two pending events (the
collision slot wins an exact tie in time when its insertion rank is lower),
and every node of every level the minimum of its up to 64 children by
(time, rank of the leaf it came from), the first child among +inf keys.
"""
import numpy as np


def level_sizes(n):
    sizes = [n]
    while sizes[-1] > 64:
        sizes.append((sizes[-1] + 63) // 64)
    return sizes


def check_calendar(cal, ev):
    """Returns a list of human-readable inconsistencies (empty when none).
    cal: Batch.calendar(s); ev: Batch.events(s) of the same system."""
    n = len(ev["t_coll"])
    bad = []
    key, idx, aux = cal["key"], cal["idx"], cal["aux"]
    sr, lr = cal["slot_rank"], cal["leaf_rank"]
    tk, tc = ev["t_coll"], ev["t_cross"]
    rk, rx = sr[0::2], sr[1::2]
    coll = (tk < tc) | ((tk == tc) & (rk < rx))
    want_key = np.where(coll, tk, tc)
    want_key = np.where(np.isnan(want_key), np.inf, want_key)
    want_aux = np.where(coll, ev["partner"], -ev["dir"])
    want_rank = np.where(coll, rk, rx)
    for name, got, want in (("leaf key", key[:n], want_key), ("leaf idx", idx[:n], np.arange(n)),
                            ("leaf aux", aux[:n], want_aux), ("leaf rank", lr, want_rank)):
        w = np.nonzero(got != want)[0]
        if len(w):
            p = w[0]
            bad.append("%s differs at %d leaves, first %d: %r != %r" % (name, len(w), p, got[p],
                                                                       want[p]))
    sizes = level_sizes(n)
    off = 0
    for l in range(1, len(sizes)):
        base = off
        off += sizes[l - 1]
        m, ng = sizes[l - 1], sizes[l]
        # children of node g: base + 64 g .. base + min(64 g + 64, m), padded
        pos = 64 * np.arange(ng)[:, None] + np.arange(64)[None, :]
        real = pos < m
        c = base + np.minimum(pos, m - 1)
        ck = np.where(real, key[c], np.inf)
        kmin = ck.min(axis=1)
        tied = real & (ck == kmin[:, None])
        # equal finite keys: the lower leaf rank; +inf keys: the first child
        rank = np.where(np.isfinite(kmin)[:, None], lr[idx[c]], 0)
        score = np.where(tied, rank, np.iinfo(np.int64).max)
        tied &= score == score.min(axis=1)[:, None]
        best = c[np.arange(ng), np.argmax(tied, axis=1)]
        k = off + np.arange(ng)
        wrong = (key[k] != key[best]) | (idx[k] != idx[best]) | (aux[k] != aux[best])
        for g in np.nonzero(wrong)[0]:
            b = best[g]
            bad.append("level %d node %d: (%r, %d, %d) expected (%r, %d, %d)"
                       % (l, g, key[k[g]], idx[k[g]], aux[k[g]], key[b], idx[b], aux[b]))
    return bad
