"""Straight-line pure-Python transcription of coupe's RCB for tiny inputs.

Written independently of oracle/rcb_oracle.c, from the algorithm's statement
(coupe/src/algorithms/recursive_bisection.rs:456-705): side tests compare
``c < t`` directly (the C oracle uses ``c - t < 0``), candidate sets are
built as Python lists, and the canonical pivot is the smallest ``c >= t``.
Used to cross-check the oracle on seeded small cases.

``mutant`` plants one mistake in the transcription (tests/test_mutants.py
checks that the golden vectors and the oracle comparison catch each one):
  "max_lt"        ``vmax < t + d`` instead of ``<=`` (:556)
  "side_le"       ``c <= t`` goes left instead of ``c < t`` (:487)
  "weight_le"     ``wl <= S - wl`` moves right instead of ``<`` (:566)
  "imb_lt"        ``imb < tol`` instead of ``<=`` (:554)
  "no_prev_stop"  the ``count == prev`` stop is dropped (:552)
  "none_min"      the none-branch raises vmin instead of lowering vmax (:540-544)
  "axis_rev"      the axis cycle runs backwards (:624, :636)
"""
import numpy as np

f32 = np.float32


def _wrap(kind, v):
    if kind == "int":
        return ((int(v) + (1 << 31)) % (1 << 32)) - (1 << 31)
    if kind == "int64":
        return ((int(v) + (1 << 63)) % (1 << 64)) - (1 << 63)
    return float(v)


MUTANTS = ("max_lt", "side_le", "weight_le", "imb_lt", "no_prev_stop", "none_min", "axis_rev")


def rcb(points, weights, iter_count, tolerance, kind, mutant=None):
    assert mutant is None or mutant in MUTANTS, mutant
    n = len(points)
    dim = len(points[0]) if n else 0
    c = [[f32(p[k]) for p in points] for k in range(dim)]
    w = [_wrap(kind, x) if kind != "double" else float(x) for x in weights]
    ids = [None] * n
    if n == 0:
        return []
    total = 0 if kind != "double" else 0.0
    for x in w:
        total = _wrap(kind, total + x)
    lo = [min(p[k] for p in points) for k in range(dim)]
    hi = [max(p[k] for p in points) for k in range(dim)]

    def to_f64(x):
        return np.float64(x)

    def split(P, a, vmin, vmax, S):
        prev = None
        # f32 bisection ends within a few hundred candidates; a mutant may not
        for _ in range(4096):
            t = f32(f32(vmin + vmax) / f32(2.0))
            if mutant == "side_le":
                left = [i for i in P if c[a][i] <= t]
                right = [i for i in P if not (c[a][i] <= t)]
            else:
                left = [i for i in P if c[a][i] < t]
                right = [i for i in P if not (c[a][i] < t)]
            wl = 0 if kind != "double" else 0.0
            for i in left:
                wl = _wrap(kind, wl + w[i])
            if not right:
                if prev == len(left):
                    return P, [], S, vmax
                if mutant == "none_min":
                    vmin = t
                else:
                    vmax = t
                prev = len(left)
                continue
            m = min(c[a][i] for i in right)
            d = f32(m - t)
            with np.errstate(divide="ignore", invalid="ignore"):
                ideal = to_f64(S) / np.float64(2.0)
                imb = abs((to_f64(wl) - ideal) / ideal)
            stop_prev = len(left) == prev and mutant != "no_prev_stop"
            stop_max = vmax < f32(t + d) if mutant == "max_lt" else vmax <= f32(t + d)
            stop_imb = imb < tolerance if mutant == "imb_lt" else imb <= tolerance
            if stop_prev or stop_max or stop_imb:
                return left, right, wl, t
            prev = len(left)
            go_right = wl <= _wrap(kind, S - wl) if mutant == "weight_le" else wl < _wrap(kind, S - wl)
            if go_right:
                vmin = t
            else:
                vmax = t
        raise RuntimeError("the split loop did not end")

    def recurse(P, level, hid, a, S, blo, bhi):
        if not P:
            return
        if level == 0:
            for i in P:
                ids[i] = hid
            return
        L, R, wl, s = split(P, a, f32(blo[a]), f32(bhi[a]), S)
        bl = list(bhi)
        bl[a] = float(s)
        br = list(blo)
        br[a] = float(s)
        nxt = (a - 1) % dim if mutant == "axis_rev" else (a + 1) % dim
        recurse(L, level - 1, (2 * hid + 1) % (1 << 64), nxt, wl, blo, bl)
        recurse(R, level - 1, (2 * hid + 2) % (1 << 64), nxt, _wrap(kind, S - wl), br, bhi)

    recurse(list(range(n)), iter_count, 0, 0, total, lo, hi)
    off = min(ids)
    return [x - off for x in ids]
