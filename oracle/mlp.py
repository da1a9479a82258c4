"""CPU oracle (test infrastructure only): FP64 restatement of the learned T2-guess network.

Only ``tests/`` and ``__graft_entry__.smoke()`` may import this module, and only as the
checker of the GPU network (``fh_mlp_*`` in libfermihip).  The network has no counterpart in
Fermi.jl: it is this project's learned initial guess, so its pins are mathematical.

Network: x (6 features) -> tanh(W1 x + b1) -> tanh(W2 . + b2) -> tanh(W3 . + b3) -> w4 . + b4.
Flat parameter layout (shared with the GPU): W1[h,6] b1[h] W2[h,h] b2[h] W3[h,h] b3[h] w4[h] b4,
each W column-major [out, in] (so ``W = p[off:off+out*in].reshape(in, out).T``).
Loss of a batch: L = 1/(2n) sum_s (y_s - t_s)^2.

Pins (``tests/test_oracle_mlp.py``):
  * ``loss_grad``: every gradient component against central finite differences of ``loss``;
  * ``forward`` of a network whose hidden layers are driven in the linear range of tanh equals
    the product of the affine maps (closed form), and tanh saturation bounds the output by
    |b4| + sum |w4|;
  * ``features``: the (ij)(ab) swap maps the features of T2[i,j,a,b] onto those of T2[j,i,b,a].
"""
import numpy as np

__all__ = ["nparams", "offsets", "unpack", "forward", "loss", "loss_grad", "features"]


def nparams(h):
    return h * 6 + h + 2 * (h * h + h) + h + 1


def offsets(h):
    o = {}
    o["W1"] = 0
    o["b1"] = o["W1"] + h * 6
    o["W2"] = o["b1"] + h
    o["b2"] = o["W2"] + h * h
    o["W3"] = o["b2"] + h
    o["b3"] = o["W3"] + h * h
    o["w4"] = o["b3"] + h
    o["b4"] = o["w4"] + h
    return o


def unpack(p, h):
    p = np.asarray(p, dtype=np.float64)
    o = offsets(h)
    W1 = p[o["W1"]:o["W1"] + 6 * h].reshape(6, h).T
    W2 = p[o["W2"]:o["W2"] + h * h].reshape(h, h).T
    W3 = p[o["W3"]:o["W3"] + h * h].reshape(h, h).T
    return {"W1": W1, "b1": p[o["b1"]:o["b1"] + h], "W2": W2, "b2": p[o["b2"]:o["b2"] + h], "W3": W3,
            "b3": p[o["b3"]:o["b3"] + h], "w4": p[o["w4"]:o["w4"] + h], "b4": p[o["b4"]]}


def forward(p, h, X):
    """X[n, 6] -> (y[n], activations)."""
    P = unpack(p, h)
    X = np.asarray(X, dtype=np.float64)
    A1 = np.tanh(X @ P["W1"].T + P["b1"])
    A2 = np.tanh(A1 @ P["W2"].T + P["b2"])
    A3 = np.tanh(A2 @ P["W3"].T + P["b3"])
    y = A3 @ P["w4"] + P["b4"]
    return y, (X, A1, A2, A3)


def loss(p, h, X, T):
    y, _ = forward(p, h, X)
    r = y - np.asarray(T, dtype=np.float64)
    return 0.5 * float(np.mean(r * r))


def loss_grad(p, h, X, T):
    """(L, dL/dp) by reverse-mode differentiation of the three tanh layers."""
    P = unpack(p, h)
    y, (X, A1, A2, A3) = forward(p, h, X)
    n = X.shape[0]
    dy = (y - np.asarray(T, dtype=np.float64)) / n
    g = np.zeros(nparams(h))
    o = offsets(h)
    g[o["w4"]:o["w4"] + h] = A3.T @ dy
    g[o["b4"]] = dy.sum()
    dZ3 = np.outer(dy, P["w4"]) * (1.0 - A3 * A3)
    g[o["W3"]:o["W3"] + h * h] = (dZ3.T @ A2).T.ravel()
    g[o["b3"]:o["b3"] + h] = dZ3.sum(0)
    dZ2 = (dZ3 @ P["W3"]) * (1.0 - A2 * A2)
    g[o["W2"]:o["W2"] + h * h] = (dZ2.T @ A1).T.ravel()
    g[o["b2"]:o["b2"] + h] = dZ2.sum(0)
    dZ1 = (dZ2 @ P["W2"]) * (1.0 - A1 * A1)
    g[o["W1"]:o["W1"] + 6 * h] = (dZ1.T @ X).T.ravel()
    g[o["b1"]:o["b1"] + h] = dZ1.sum(0)
    r = y - np.asarray(T, dtype=np.float64)
    return 0.5 * float(np.mean(r * r)), g


def features(T2mp2, eps_o, eps_v, dij, x_mean=None, x_scale=None):
    """Feature rows of every amplitude of T2[o,o,v,v] in column-major element order:
    (T2_MP2, eps_i, eps_j, eps_a, eps_b, d_ij), standardised as (x - x_mean) * x_scale."""
    o, _, v, _ = T2mp2.shape
    i, j, a, b = np.meshgrid(np.arange(o), np.arange(o), np.arange(v), np.arange(v), indexing="ij")
    F = np.stack([T2mp2, eps_o[i], eps_o[j], eps_v[a], eps_v[b], dij[i, j]], axis=-1)
    F = np.transpose(F, (3, 2, 1, 0, 4)).reshape(-1, 6)     # column-major element order
    if x_mean is not None:
        F = (F - np.asarray(x_mean)) * np.asarray(x_scale)
    return F
