"""Pairwise tensor contraction over BLAS for the CPU oracle (test infrastructure only).

TensorOperations.jl evaluates every ``@tensor`` expression by pairwise
transpose-transpose-GEMM-transpose steps.  ``ein`` does the same with
``numpy.tensordot`` (BLAS GEMM), choosing the pairwise order of 3+ operands with
``numpy.einsum_path``; it returns exactly what ``numpy.einsum`` returns for the same
subscripts (the oracle tests compare against brute-force spin-orbital results, which
pins this helper as well).
"""
import numpy as np

__all__ = ["ein"]


def _reduce_private(sub, A, keep):
    """Sum axes of A whose letters appear nowhere else (not in ``keep``)."""
    drop = [k for k, ch in enumerate(sub) if ch not in keep]
    if not drop:
        return sub, A
    A = A.sum(axis=tuple(drop))
    return "".join(ch for ch in sub if ch in keep), A


def _perm_cost(sub, order, size):
    """Relative cost of bringing an operand (C-order letters ``sub``) into ``order``."""
    if sub == order:
        return 0.0
    return size * (1.0 if sub[-1] == order[-1] else 4.0)


def _pair(ia, ib, out, A, B):
    ia, A = _reduce_private(ia, A, ib + out)
    ib, B = _reduce_private(ib, B, ia + out)
    if any(ch in ib and ch in out for ch in ia):
        return np.einsum(f"{ia},{ib}->{out}", A, B)          # batch index: not a GEMM
    con_a = [ch for ch in ia if ch in ib]
    con_b = [ch for ch in ib if ch in ia]
    # tensordot(X, Y) lays X out as (free, contracted) and Y as (contracted, free); pick the
    # operand order and contracted-index order that permute the least data
    best = None
    for X, ix, Y, iy in ((A, ia, B, ib), (B, ib, A, ia)):
        for con in (con_a, con_b):
            fx = "".join(ch for ch in ix if ch not in con)
            fy = "".join(ch for ch in iy if ch not in con)
            cost = _perm_cost(ix, fx + "".join(con), X.size) + _perm_cost(iy, "".join(con) + fy, Y.size)
            if best is None or cost < best[0]:
                best = (cost, X, ix, Y, iy, con, fx + fy)
    _, X, ix, Y, iy, con, rl = best
    fx = [ch for ch in ix if ch not in con]
    fy = [ch for ch in iy if ch not in con]
    dims = dict(zip(ix, X.shape))
    dims.update(zip(iy, Y.shape))
    nfx = int(np.prod([dims[c] for c in fx])) if fx else 1
    nfy = int(np.prod([dims[c] for c in fy])) if fy else 1
    ncon = int(np.prod([dims[c] for c in con])) if con else 1
    # transpose + reshape (numpy's reshape copy is the fast path for permuted operands)
    Xm = np.transpose(X, [ix.index(c) for c in fx + con]).reshape(nfx, ncon)
    Ym = np.transpose(Y, [iy.index(c) for c in con + fy]).reshape(ncon, nfy)
    R = (Xm @ Ym).reshape([dims[c] for c in fx + fy])
    return np.transpose(R, [rl.index(ch) for ch in out]) if rl != out else R


def ein(sub, *ops):
    inputs, out = sub.split("->")
    ins = inputs.split(",")
    if len(ops) == 1:
        return np.einsum(sub, ops[0])
    if len(ops) == 2:
        return _pair(ins[0], ins[1], out, ops[0], ops[1])
    path, _ = np.einsum_path(sub, *ops, optimize="optimal")
    ops, ins = list(ops), list(ins)
    for step in path[1:]:
        idx = sorted(step, reverse=True)
        pick_ops = [ops.pop(p) for p in idx]
        pick_ins = [ins.pop(p) for p in idx]
        rest = "".join(ins) + out
        if len(pick_ops) == 1:
            s, X = _reduce_private(pick_ins[0], pick_ops[0], rest)
        else:
            keep = "".join(ch for ch in dict.fromkeys(pick_ins[0] + pick_ins[1]) if ch in rest)
            X = _pair(pick_ins[0], pick_ins[1], keep, pick_ops[0], pick_ops[1])
            s = keep
        ops.append(X)
        ins.append(s)
    s, X = ins[0], ops[0]
    s, X = _reduce_private(s, X, out)
    return np.transpose(X, [s.index(ch) for ch in out]) if s != out else X
