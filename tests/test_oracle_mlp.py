"""Pins of the T2-guess network oracle (oracle/mlp.py): finite-difference gradients, the
linear-regime closed form, the saturation bound and the (ij)(ab) symmetry of the features."""
import numpy as np

from oracle import mlp as M


def _net(h, seed, scale=0.5):
    rng = np.random.default_rng(seed)
    return scale * rng.standard_normal(M.nparams(h))


def test_gradient_matches_finite_differences():
    h = 8
    p = _net(h, 0)
    rng = np.random.default_rng(1)
    X = rng.standard_normal((13, 6))
    T = rng.standard_normal(13)
    L, g = M.loss_grad(p, h, X, T)
    assert abs(L - M.loss(p, h, X, T)) < 1e-15
    eps = 1e-6
    fd = np.zeros_like(p)
    for k in range(p.size):
        q = p.copy(); q[k] += eps
        r = p.copy(); r[k] -= eps
        fd[k] = (M.loss(q, h, X, T) - M.loss(r, h, X, T)) / (2 * eps)
    assert np.max(np.abs(fd - g)) < 1e-8 * max(1.0, np.max(np.abs(g)))


def test_linear_regime_closed_form():
    # tiny weights keep tanh in its linear range: y ~= w4.(W3 (W2 (W1 x + b1) + b2) + b3) + b4
    h = 16
    p = _net(h, 2, scale=1e-4)
    P = M.unpack(p, h)
    X = np.random.default_rng(3).standard_normal((7, 6))
    y, _ = M.forward(p, h, X)
    lin = (((X @ P["W1"].T + P["b1"]) @ P["W2"].T + P["b2"]) @ P["W3"].T + P["b3"]) @ P["w4"] + P["b4"]
    # tanh(z) = z - z^3/3 + ...: with |z| < 1e-3 the relative deviation of each layer is < 1e-6
    assert np.max(np.abs(y - lin)) < 1e-6 * np.max(np.abs(lin - P["b4"]))


def test_saturation_bound_and_layout():
    h = 4
    p = _net(h, 4, scale=30.0)
    P = M.unpack(p, h)
    X = np.random.default_rng(5).standard_normal((50, 6)) * 10
    y, _ = M.forward(p, h, X)
    assert np.all(np.abs(y - P["b4"]) <= np.sum(np.abs(P["w4"])) + 1e-12)
    # column-major [out, in]: W1[r, f] sits at p[f*h + r]
    o = M.offsets(h)
    assert P["W1"][2, 5] == p[o["W1"] + 5 * h + 2]


def test_features_pair_symmetry():
    rng = np.random.default_rng(6)
    o, v = 3, 4
    T2 = rng.standard_normal((o, o, v, v))
    T2 = T2 + np.transpose(T2, (1, 0, 3, 2))
    eo, ev = rng.standard_normal(o), rng.standard_normal(v)
    d = np.abs(rng.standard_normal((o, o)))
    d = d + d.T
    F = M.features(T2, eo, ev, d)
    Fr = F.reshape(v, v, o, o, 6)                 # [b, a, j, i, feature]
    swapped = np.transpose(Fr, (1, 0, 3, 2, 4))   # element (j,i,b,a)
    # features of (j,i,b,a) = features of (i,j,a,b) with eps_i<->eps_j and eps_a<->eps_b exchanged
    assert np.array_equal(swapped[..., 0], Fr[..., 0])
    assert np.array_equal(swapped[..., 1], Fr[..., 2]) and np.array_equal(swapped[..., 3], Fr[..., 4])
    assert np.array_equal(swapped[..., 5], Fr[..., 5])
    # element order is column-major: row e = i + o*(j + o*(a + v*b))
    i, j, a, b = 1, 2, 3, 0
    e = i + o * (j + o * (a + v * b))
    assert np.allclose(F[e], [T2[i, j, a, b], eo[i], eo[j], ev[a], ev[b], d[i, j]])
