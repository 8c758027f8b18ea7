"""Pins of the self-contact oracle (oracle/selfcontact.py) by mathematics: distances against dense sampling of
the primitives and closed forms, the barrier gradient against central differences of the energy, the
Gauss-Newton coupling against the outer product of that gradient's direction, and the crossing test on
constructed cases."""
import numpy as np
import pytest

from oracle import selfcontact as S


def test_point_triangle_matches_dense_sampling():
    rng = np.random.default_rng(1)
    g = np.linspace(0.0, 1.0, 301)
    u, v = np.meshgrid(g, g, indexing="ij")
    keep = u + v <= 1.0
    u, v = u[keep], v[keep]
    for _ in range(40):
        a, b, c, p = rng.standard_normal((4, 3))
        d, w = S.point_triangle(p, a, b, c)
        pts = a + u[:, None] * (b - a) + v[:, None] * (c - a)
        sampled = np.linalg.norm(pts - p, axis=1).min()
        assert d <= sampled + 1e-12
        assert sampled - d <= 5e-3 * max(1.0, np.linalg.norm(b - a) + np.linalg.norm(c - a))
        q = w[0] * a + w[1] * b + w[2] * c
        assert abs(np.linalg.norm(p - q) - d) < 1e-12 and abs(w.sum() - 1) < 1e-12 and w.min() >= -1e-12


def test_segment_segment_matches_dense_sampling():
    rng = np.random.default_rng(2)
    g = np.linspace(0.0, 1.0, 401)
    for k in range(60):
        p1, q1, p2, q2 = rng.standard_normal((4, 3))
        if k % 10 == 0:
            q2 = p2 + 0.7 * (q1 - p1)  # parallel
        d, s, t = S.segment_segment(p1, q1, p2, q2)
        A = p1 + g[:, None] * (q1 - p1)
        B = p2 + g[:, None] * (q2 - p2)
        sampled = np.linalg.norm(A[:, None, :] - B[None, :, :], axis=2).min()
        assert d <= sampled + 1e-12
        assert sampled - d <= 5e-3 * (np.linalg.norm(q1 - p1) + np.linalg.norm(q2 - p2))
        x = (p1 + s * (q1 - p1)) - (p2 + t * (q2 - p2))
        assert abs(np.linalg.norm(x) - d) < 1e-12


def test_closed_forms():
    a, b, c = np.array([0.0, 0, 0]), np.array([1.0, 0, 0]), np.array([0.0, 1, 0])
    d, w = S.point_triangle(np.array([0.2, 0.3, 0.7]), a, b, c)
    assert abs(d - 0.7) < 1e-15 and np.allclose(w, [0.5, 0.2, 0.3])
    d, w = S.point_triangle(np.array([2.0, 0.0, 0.0]), a, b, c)  # beyond vertex b
    assert abs(d - 1.0) < 1e-15 and np.allclose(w, [0, 1, 0])
    d, s, t = S.segment_segment(np.array([-1.0, 0, 0]), np.array([1.0, 0, 0]), np.array([0.0, -1, 0.25]),
                                np.array([0.0, 1, 0.25]))
    assert abs(d - 0.25) < 1e-15 and abs(s - 0.5) < 1e-15 and abs(t - 0.5) < 1e-15


def random_cluster(rng):
    """Two small triangles and a point, close to each other (pairs closer than 0.3)."""
    x = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0.2, 0.3, 0.1], [0.6, 0.2, 0.15], [0.3, 0.7, 0.12]], float)
    return x + 0.02 * rng.standard_normal(x.shape)


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(3)
    tri = np.array([[0, 1, 2], [3, 4, 5]])
    dhat, kappa = 0.3, 2.0
    for _ in range(5):
        x = random_cluster(rng)
        quads, dists, kinds = S.candidates(x, tri, dhat)
        assert len(quads) > 0
        g = S.gradient(x, quads, kinds, dhat, kappa)
        h = 1e-7
        fd = np.zeros_like(x)
        for v in range(x.shape[0]):
            for r in range(3):
                xp, xm = x.copy(), x.copy()
                xp[v, r] += h
                xm[v, r] -= h
                fd[v, r] = (S.energy(xp, quads, kinds, dhat, kappa) - S.energy(xm, quads, kinds, dhat, kappa)) / (2 * h)
        assert np.abs(g - fd).max() <= 1e-6 * max(1.0, np.abs(fd).max())


def test_gauss_newton_is_outer_product_of_the_gradient_directions():
    rng = np.random.default_rng(4)
    tri = np.array([[0, 1, 2], [3, 4, 5]])
    x = random_cluster(rng)
    quads, _, kinds = S.candidates(x, tri, 0.3)
    H = S.gauss_newton(x, quads, kinds, 0.3, 2.0, 6)
    assert np.allclose(H, H.T)
    assert np.linalg.eigvalsh(H).min() >= -1e-9 * np.abs(H).max()
    # one pair: b'' g g^T with g the distance gradient
    H1 = S.gauss_newton(x, quads[:1], kinds[:1], 0.3, 2.0, 6)
    d, n, c = S.pair_geometry(x, quads[0], kinds[0])
    gvec = np.zeros(18)
    for k, v in enumerate(quads[0]):
        gvec[3 * v:3 * v + 3] += c[k] * n
    assert np.allclose(H1, S.barrier(d, 0.3, 2.0)[2] * np.outer(gvec, gvec))


def test_crossing_test():
    a, b, c = np.array([0.0, 0, 0]), np.array([1.0, 0, 0]), np.array([0.0, 1, 0])
    assert S.segment_crosses_triangle(np.array([0.2, 0.2, -1]), np.array([0.2, 0.2, 1]), a, b, c)
    assert not S.segment_crosses_triangle(np.array([0.2, 0.2, 0.1]), np.array([0.2, 0.2, 1]), a, b, c)
    assert not S.segment_crosses_triangle(np.array([0.8, 0.8, -1]), np.array([0.8, 0.8, 1]), a, b, c)
