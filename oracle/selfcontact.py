"""Oracle for the sheet's self-contact (no reference counterpart: the reference's contact is vertex-collider
only): numpy f64 point-triangle and edge-edge distances, the brute-force candidate sets, the barrier energy
and its gradient and Gauss-Newton coupling, a segment-triangle crossing test. Test infrastructure only: only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only as the checker.

The distances are computed differently from the device code on purpose (which follows Voronoi-region
classification): the point-triangle distance is the smaller of the in-plane projection distance (when the
projection falls inside) and the three point-segment distances; the edge-edge distance is the unconstrained
closest pair of the two lines when it lies inside both segments, else the smallest of the four
endpoint-segment distances. Pins (tests/test_self_contact_oracle.py): both against dense sampling of the
primitives, the barrier gradient against central differences of the energy, and the closed forms (a point
above a triangle's interior, crossing edges).
"""
from __future__ import annotations

import numpy as np


def unique_edges(tri: np.ndarray) -> np.ndarray:
    tri = np.asarray(tri, np.int64).reshape(-1, 3)
    e = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])
    e = np.sort(e, axis=1)
    return np.unique(e, axis=0)


def point_segment(p, a, b):
    """(distance, t) of the closest point a + t (b - a) to p, t in [0, 1]; p, a, b [..., 3]."""
    ab = b - a
    den = np.einsum("...i,...i", ab, ab)
    t = np.where(den > 0, np.einsum("...i,...i", p - a, ab) / np.where(den > 0, den, 1.0), 0.0)
    t = np.clip(t, 0.0, 1.0)
    q = a + t[..., None] * ab
    return np.linalg.norm(p - q, axis=-1), t


def point_triangle(p, a, b, c):
    """(distance, barycentric weights [..., 3] of the closest point) of p to triangle (a, b, c)."""
    p, a, b, c = (np.asarray(v, np.float64) for v in (p, a, b, c))
    ab, ac = b - a, c - a
    n = np.cross(ab, ac)
    nn = np.einsum("...i,...i", n, n)
    # projection onto the plane and its barycentric coordinates
    ap = p - a
    d00, d01, d11 = (np.einsum("...i,...i", ab, ab), np.einsum("...i,...i", ab, ac),
                     np.einsum("...i,...i", ac, ac))
    d20, d21 = np.einsum("...i,...i", ap, ab), np.einsum("...i,...i", ap, ac)
    den = d00 * d11 - d01 * d01
    v = (d11 * d20 - d01 * d21) / den
    w = (d00 * d21 - d01 * d20) / den
    u = 1.0 - v - w
    inside = (u >= 0) & (v >= 0) & (w >= 0) & (nn > 0)
    plane = np.abs(np.einsum("...i,...i", ap, n)) / np.sqrt(np.where(nn > 0, nn, 1.0))
    dab, tab = point_segment(p, a, b)
    dbc, tbc = point_segment(p, b, c)
    dca, tca = point_segment(p, c, a)
    cand = np.stack([np.where(inside, plane, np.inf), dab, dbc, dca], axis=-1)
    k = np.argmin(cand, axis=-1)
    d = np.take_along_axis(cand, k[..., None], axis=-1)[..., 0]
    bary = np.stack([
        np.stack([u, v, w], axis=-1),
        np.stack([1 - tab, tab, 0 * tab], axis=-1),
        np.stack([0 * tbc, 1 - tbc, tbc], axis=-1),
        np.stack([tca, 0 * tca, 1 - tca], axis=-1),
    ], axis=-2)
    wts = np.take_along_axis(bary, k[..., None, None].repeat(3, axis=-1), axis=-2)[..., 0, :]
    return d, wts


def segment_segment(p1, q1, p2, q2):
    """(distance, s, t) of the closest points p1 + s (q1 - p1), p2 + t (q2 - p2)."""
    p1, q1, p2, q2 = (np.asarray(v, np.float64) for v in (p1, q1, p2, q2))
    d1, d2, r = q1 - p1, q2 - p2, p1 - p2
    a, e = np.einsum("...i,...i", d1, d1), np.einsum("...i,...i", d2, d2)
    b, c, f = np.einsum("...i,...i", d1, d2), np.einsum("...i,...i", d1, r), np.einsum("...i,...i", d2, r)
    den = a * e - b * b
    ok = den > 1e-14 * a * e
    s0 = np.where(ok, (b * f - c * e) / np.where(ok, den, 1.0), -1.0)
    t0 = np.where(ok, (a * f - b * c) / np.where(ok, den, 1.0), -1.0)
    inner = ok & (s0 >= 0) & (s0 <= 1) & (t0 >= 0) & (t0 <= 1)
    dl = np.linalg.norm((p1 + s0[..., None] * d1) - (p2 + t0[..., None] * d2), axis=-1)
    e1, t1 = point_segment(p1, p2, q2)
    e2, t2 = point_segment(q1, p2, q2)
    e3, s3 = point_segment(p2, p1, q1)
    e4, s4 = point_segment(q2, p1, q1)
    cand = np.stack([np.where(inner, dl, np.inf), e1, e2, e3, e4], axis=-1)
    k = np.argmin(cand, axis=-1)
    d = np.take_along_axis(cand, k[..., None], axis=-1)[..., 0]
    S = np.stack([s0, 0 * t1, 1 + 0 * t2, s3, s4], axis=-1)
    Tt = np.stack([t0, t1, t2, 0 * s3, 1 + 0 * s4], axis=-1)
    s = np.take_along_axis(S, k[..., None], axis=-1)[..., 0]
    t = np.take_along_axis(Tt, k[..., None], axis=-1)[..., 0]
    return d, s, t


def candidates(x, tri, within: float):
    """Brute force: point-triangle pairs (v, t) with v not in t and edge-edge pairs (e, f), e < f, without a
    shared vertex, closer than `within`; as vertex quads (p, a, b, c) / (p1, q1, p2, q2) with distances."""
    x = np.asarray(x, np.float64).reshape(-1, 3)
    tri = np.asarray(tri, np.int64).reshape(-1, 3)
    V = len(x)
    quads, dists, kinds = [], [], []
    for t, (a, b, c) in enumerate(tri):
        v = np.arange(V)
        keep = (v != a) & (v != b) & (v != c)
        d, _ = point_triangle(x[keep], x[a], x[b], x[c])
        sel = v[keep][d < within]
        for vv, dd in zip(sel, d[d < within]):
            quads.append((vv, a, b, c))
            dists.append(dd)
            kinds.append(0)
    E = unique_edges(tri)
    for i, (p, q) in enumerate(E):
        other = E[i + 1:]
        keep = (other[:, 0] != p) & (other[:, 0] != q) & (other[:, 1] != p) & (other[:, 1] != q)
        o = other[keep]
        d, _, _ = segment_segment(x[p], x[q], x[o[:, 0]], x[o[:, 1]])
        for (r, s), dd in zip(o[d < within], d[d < within]):
            quads.append((p, q, r, s))
            dists.append(dd)
            kinds.append(1)
    return np.array(quads, np.int64).reshape(-1, 4), np.array(dists), np.array(kinds, np.int64)


def barrier(d, dhat, kappa):
    """b(d) = -kappa (d - dhat)^2 ln(d / dhat) for d < dhat, else 0 (contact.rs:81-108), with b' and b''."""
    d = np.asarray(d, np.float64)
    act = d < dhat
    dd = np.where(act, d, dhat)
    t = dd - dhat
    ln = np.log(dd / dhat)
    b = np.where(act, -kappa * t * t * ln, 0.0)
    b1 = np.where(act, -kappa * (2 * t * ln + t * t / dd), 0.0)
    b2 = np.where(act, -kappa * (2 * ln + 4 * t / dd - t * t / (dd * dd)), 0.0)
    return b, b1, b2


def pair_geometry(x, quad, kind):
    """(distance, unit direction n, coefficients c [4]) of one pair: grad of d w.r.t. vertex k is c_k n."""
    x = np.asarray(x, np.float64).reshape(-1, 3)
    p = x[list(quad)]
    if kind == 0:
        d, w = point_triangle(p[0], p[1], p[2], p[3])
        q = w[0] * p[1] + w[1] * p[2] + w[2] * p[3]
        diff = p[0] - q
        c = np.array([1.0, -w[0], -w[1], -w[2]])
    else:
        d, s, t = segment_segment(p[0], p[1], p[2], p[3])
        diff = (p[0] + s * (p[1] - p[0])) - (p[2] + t * (p[3] - p[2]))
        c = np.array([1.0 - s, s, -(1.0 - t), -t])
    return float(d), diff / d, c


def energy(x, quads, kinds, dhat, kappa):
    """Sum of b(d) over the given pairs at positions x [V, 3]."""
    x = np.asarray(x, np.float64).reshape(-1, 3)
    total = 0.0
    for quad, kind in zip(quads, kinds):
        d = pair_geometry(x, quad, kind)[0]
        total += float(barrier(d, dhat, kappa)[0])
    return total


def gradient(x, quads, kinds, dhat, kappa):
    """Energy gradient [V, 3]: b'(d) c_k n per vertex of every pair closer than dhat."""
    x = np.asarray(x, np.float64).reshape(-1, 3)
    g = np.zeros_like(x)
    for quad, kind in zip(quads, kinds):
        d, n, c = pair_geometry(x, quad, kind)
        _, b1, _ = barrier(d, dhat, kappa)
        for k, v in enumerate(quad):
            g[v] += b1 * c[k] * n
    return g


def gauss_newton(x, quads, kinds, dhat, kappa, V):
    """Dense Gauss-Newton Hessian [3V, 3V]: b''(d) (c_k n)(c_l n)^T in block (v_k, v_l)."""
    x = np.asarray(x, np.float64).reshape(-1, 3)
    H = np.zeros((3 * V, 3 * V))
    for quad, kind in zip(quads, kinds):
        d, n, c = pair_geometry(x, quad, kind)
        _, _, b2 = barrier(d, dhat, kappa)
        nn = np.outer(n, n)
        for k, vk in enumerate(quad):
            for l, vl in enumerate(quad):
                H[3 * vk:3 * vk + 3, 3 * vl:3 * vl + 3] += b2 * c[k] * c[l] * nn
    return H


def segment_crosses_triangle(p, q, a, b, c) -> bool:
    """Whether segment pq crosses the interior of triangle abc (Moller-Trumbore on the segment)."""
    p, q, a, b, c = (np.asarray(v, np.float64) for v in (p, q, a, b, c))
    d = q - p
    e1, e2 = b - a, c - a
    h = np.cross(d, e2)
    det = e1 @ h
    if abs(det) < 1e-300:
        return False
    inv = 1.0 / det
    s = p - a
    u = inv * (s @ h)
    if u < 0 or u > 1:
        return False
    qv = np.cross(s, e1)
    v = inv * (d @ qv)
    if v < 0 or u + v > 1:
        return False
    t = inv * (e2 @ qv)
    return 0.0 <= t <= 1.0


def crossings(x, tri) -> int:
    """Number of (edge, triangle) pairs without a shared vertex where the edge crosses the triangle."""
    x = np.asarray(x, np.float64).reshape(-1, 3)
    tri = np.asarray(tri, np.int64).reshape(-1, 3)
    E = unique_edges(tri)
    lo, hi = x[tri].min(axis=1), x[tri].max(axis=1)
    count = 0
    for p, q in E:
        elo, ehi = np.minimum(x[p], x[q]), np.maximum(x[p], x[q])
        near = np.all((lo <= ehi) & (hi >= elo), axis=1)
        for t in np.nonzero(near)[0]:
            a, b, c = tri[t]
            if p in (a, b, c) or q in (a, b, c):
                continue
            if segment_crosses_triangle(x[p], x[q], x[a], x[b], x[c]):
                count += 1
    return count
