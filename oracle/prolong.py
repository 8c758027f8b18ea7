"""Test infrastructure, not product code: an f64 numpy restatement of the reference's prolongation of a coarse
equilibrium to the next grid (e7_prolong.rs:9-47) -- the checker of the library's prolongation kernel. Only
tests may import it.

Pinned by tests/test_prolong.py: the reference's own unit test of `prolong` (e7_prolong.rs:72-83, five
known vertices of a 3x3 -> 5x5 prolongation), closed forms of the projection (a vertex inside the sphere or
below the floor lands at d = dhat), and the r/L of the prolonged n17 equilibrium on the n33 mesh against the
first line of the reference's experiments/drape7/drape-n33/base/prolong.log (4.845e2).
"""
import numpy as np


def prolong(coarse, coarse_columns):
    """e7_prolong.rs:9-23: fine vertex (i, j) of the (2n - 1)^2 grid = the mean of the coarse corners
    (i/2, j/2), (ceil(i/2), j/2), (i/2, ceil(j/2)), (ceil(i/2), ceil(j/2)), summed in that order from 0."""
    c = np.asarray(coarse, np.float64).reshape(-1, 3)
    n = coarse_columns
    f = 2 * (n - 1) + 1
    v = np.arange(f * f)
    i, j = v % f, v // f
    i0, i1, j0, j1 = i // 2, (i + 1) // 2, j // 2, (j + 1) // 2
    acc = np.zeros((f * f, 3))
    for a, b in ((i0, j0), (i1, j0), (i0, j1), (i1, j1)):
        acc = acc + c[b * n + a]
    return acc / 4.0


def colliders_of(packed):
    """[(kind, a, b, c, w)] from the packed collider header [count, dhat, kappa, flag, 8 per collider]."""
    p = np.asarray(packed, np.float64)
    return [tuple(p[4 + 8 * k:4 + 8 * k + 5]) for k in range(int(p[0]))], float(p[1])


def distance(col, x):
    """contact.rs:14-19 (dot summed left to right)."""
    kind, a, b, c, w = col
    if kind == 1.0:
        return a * x[0] + b * x[1] + c * x[2] - w
    q = (x[0] - a, x[1] - b, x[2] - c)
    return np.sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]) - w


def normal(col, x):
    """contact.rs:22-30."""
    kind, a, b, c, w = col
    if kind == 1.0:
        return (a, b, c)
    q = (x[0] - a, x[1] - b, x[2] - c)
    length = np.sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2])
    return (q[0] / length, q[1] / length, q[2] / length)


def project_out(positions, packed):
    """e7_prolong.rs:26-41: every vertex closer than dhat to a collider moves along its normal to d = dhat, one
    pass over the colliders in order. Returns (positions, moved count)."""
    cols, dhat = colliders_of(packed)
    x = np.array(positions, np.float64).reshape(-1, 3)
    moved = 0
    for v in range(x.shape[0]):
        p = [float(x[v, 0]), float(x[v, 1]), float(x[v, 2])]
        before = list(p)
        for col in cols:
            d = distance(col, p)
            if d < dhat:
                n = normal(col, p)
                p = [p[k] + (dhat - d) * n[k] for k in range(3)]
        x[v] = p
        moved += int(p != before)
    return x, moved


def infeasible(positions, packed):
    """e7_prolong.rs:44-46: vertices at or inside a collider."""
    cols, _ = colliders_of(packed)
    x = np.asarray(positions, np.float64).reshape(-1, 3)
    return [v for v in range(x.shape[0]) if cols and min(distance(c, x[v]) for c in cols) <= 0.0]
