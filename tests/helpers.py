"""Comparison helpers for parity tests."""
import numpy as np

COORD_NAMES = ("x", "xp", "y", "yp", "s", "delta")


def bit_equal(a, b):
    """element-wise equality treating NaN == NaN (any payload)"""
    a = np.asarray(a)
    b = np.asarray(b)
    return (a == b) | (np.isnan(a) & np.isnan(b))


def assert_rows_identical(gpu, ref, n_gpu, n_ref, what=""):
    """full (n, 10) arrays in reference row order, bit-identical"""
    assert n_gpu == n_ref, "%s survivors differ: gpu %d oracle %d" % (what, n_gpu, n_ref)
    eq = bit_equal(gpu, ref)
    if not eq.all():
        bad = np.argwhere(~eq)
        r, c = bad[0]
        raise AssertionError("%s: %d mismatching entries; first row %d col %d gpu=%r ref=%r (ids %r/%r)" % (
            what, len(bad), r, c, gpu[r, c], ref[r, c], gpu[r, 6], ref[r, 6]))


def by_id(rows, n):
    """dict id -> row for survivors"""
    return {int(r[6]): r for r in rows[:n]}


def deviation_stats(a, na, b, nb, floors=None):
    """per-coordinate max and RMS of |a-b|/max(|b|, floor) over survivors common to both,
    matched by particle ID; floors default to the RMS of each coordinate of b"""
    A = by_id(a, na)
    B = by_id(b, nb)
    ids = sorted(set(A) & set(B))
    if not ids:
        return {}, 0
    X = np.array([A[i][:6] for i in ids])
    Y = np.array([B[i][:6] for i in ids])
    if floors is None:
        floors = np.sqrt(np.mean(Y ** 2, axis=0))
        floors = np.where(floors > 0, floors, 1.0)
    rel = np.abs(X - Y) / np.maximum(np.abs(Y), floors)
    out = {}
    for k, name in enumerate(COORD_NAMES):
        out[name] = (float(np.max(rel[:, k])), float(np.sqrt(np.mean(rel[:, k] ** 2))))
    return out, len(ids)


def loss_mask(rows, n):
    """set of (id, loss pass, loss s) of lost particles"""
    return {(int(r[6]), int(r[7]), float(r[4])) for r in rows[n:]}
