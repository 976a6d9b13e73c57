"""Shared test helpers: build a problem, run the oracle on it, compare."""
import numpy as np

import labels as L


def packed_to_dense(p, n):
    A = np.zeros((n, n))
    k = 0
    for i in range(n):
        A[i, :i + 1] = p[k:k + i + 1]
        k += i + 1
    return A + np.tril(A, -1).T


def dense_to_packed(A):
    n = A.shape[0]
    return np.concatenate([A[i, :i + 1] for i in range(n)])


def oracle_pass(orc, rec_ptr, inder, glder, labels, var, par, nvar, **kw):
    return orc.loopn(rec_ptr, inder, glder.astype(np.float64), labels, var, par, nvar, **kw)


def maps(rec_ptr, inder, glder, **kw):
    return L.label_maps(rec_ptr, inder, glder, **kw)


def rel_err(a, b):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    s = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() / s
