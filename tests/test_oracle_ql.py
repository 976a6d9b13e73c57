"""Pin the oracle's constraint elimination (QL decomposition of the constraint matrix,
src/mpqldec.f90) against the reference's own routines compiled by flang into oracle/_ref, and
against closed forms.  CPU only."""
import os
import zlib

import numpy as np
import pytest

import oracle


def _ref():
    if not os.path.exists(oracle.REF):
        pytest.skip("reference numeric library (oracle/_ref) not built")
    # the QL module allocates its arrays once per library instance: a private copy per case
    return oracle.Reference(fresh=True)


def _case(name):
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    if name == "layers":          # 3 layers x 2 interleaved sums: disjoint groups, one block per layer
        n, cons = 40, []
        for L in range(3):
            b = 1 + 12 * L
            cons.append((list(range(b, b + 12, 2)), list(rng.normal(size=6))))
            cons.append((list(range(b + 1, b + 12, 2)), list(rng.normal(size=6))))
        return n, cons
    if name == "coupled":         # plus one constraint over three layers: a single block
        n, cons = _case("layers")
        return n, cons + [([5, 20, 30], [1.0, -0.5, 0.25])]
    if name == "free_tail":       # constrained blocks first, 15 unconstrained parameters last
        n = 45
        cons = [(list(range(1, 11)), list(rng.normal(size=10))), (list(range(3, 14)), list(rng.normal(size=11))),
                (list(range(16, 30)), list(rng.normal(size=14)))]
        return n, cons
    if name == "random":          # random subsets, overlapping in arbitrary ways
        n, cons = 30, []
        for _ in range(5):
            ix = sorted(rng.choice(np.arange(1, n + 1), size=int(rng.integers(3, 12)), replace=False).tolist())
            cons.append((ix, list(rng.normal(size=len(ix)))))
        return n, cons
    if name == "short_block":     # one block ending at parameter 28 of 30: overlaps positions 26..30
        n, cons = 30, [([1, 9, 17, 28], list(rng.normal(size=4)))]
        for _ in range(4):
            ix = sorted(rng.choice(np.arange(2, 28), size=int(rng.integers(3, 10)), replace=False).tolist())
            cons.append((ix, list(rng.normal(size=len(ix)))))
        return n, cons
    if name == "single":
        return 12, [(list(range(1, 13)), list(rng.normal(size=12)))]
    raise KeyError(name)


CASES = ["layers", "coupled", "free_tail", "random", "short_block", "single"]


def _dense_c(n, cons, order):
    C = np.zeros((len(cons), n))
    for k, j in enumerate(order):
        for v, f in zip(*cons[j]):
            C[k, v - 1] += f
    return C


def _band(M, mb, nrows=None):
    n = M.shape[0] if nrows is None else nrows
    return np.concatenate([M[i, max(0, i - mb + 1):i + 1] for i in range(n)])


def _reflectors(q, n):
    m = q["ncon"]
    V = np.zeros((n, m))
    for k in range(m):
        lo, hi = q["irng"][2 * k], q["irng"][2 * k + 1]
        if hi >= lo:
            V[lo - 1:hi, k] = q["matV"][q["ioffRow"][k]:q["ioffRow"][k] + hi - lo + 1]
        V[n - m + k, k] = q["vecVk"][k]
    return V


def test_layout_follows_pede_rules():
    """grpcon/prpcon (src/pede.f90:840-1003, 1077-1371) on a hand-worked example: four
    constraints without shared parameters are four groups; overlapping ranges merge groups
    1-2 and 3-4 into two blocks; storage starts at the chain's first parameter and ends at
    the block's last."""
    cons = [([7, 9], [1.0, 1.0]), ([1, 3, 5], [1.0, 2.0, 3.0]), ([8, 10], [1.0, -1.0]), ([2, 4], [0.5, 0.5])]
    lay = oracle.ql_layout(10, cons)
    assert lay["order"] == [1, 3, 0, 2]
    assert lay["rcon"].reshape(-1, 4).tolist() == [[1, 5, 1, 5], [2, 4, 2, 5], [7, 9, 7, 10], [8, 10, 8, 10]]
    assert lay["bcon"].reshape(-1, 3).tolist() == [[1, 1, 5], [3, 7, 10], [5, 0, 0]]
    assert lay["nblk"] == 2


@pytest.mark.parametrize("name", CASES)
def test_qldecb_bitwise(orc, name):
    n, cons = _case(name)
    lay = oracle.ql_layout(n, cons)
    R = _ref()
    emin, emax = R.ql_decompose(n, lay)
    q = orc.qldecb(n, lay)
    I = np.eye(n)
    for t in (False, True):
        assert np.array_equal(orc.qlmlq(q, I, t), R.qlmlq(I, t))
    d = np.random.default_rng(1).normal(size=len(cons))
    assert np.array_equal(orc.qlbsub(q, d), R.qlbsub(d, len(cons)))
    dg = np.abs(np.diag(q["matL"].reshape(len(cons), -1)))
    assert (abs(emin), abs(emax)) == (dg.min(), dg.max())


@pytest.mark.parametrize("name", CASES)
def test_ql_closed_forms(orc, name):
    """Q orthogonal, C^T = Q [0; L] with L lower triangular, L^T y = d, and the solution by
    elimination equal to the solution of the KKT system of a random SPD matrix (with the
    aligned start of L; the reference's own L, see test_reference_l_start, coincides with it
    unless a constraint block's storage ends before the last parameter while overlapping the
    eliminated positions)."""
    n, cons = _case(name)
    lay = oracle.ql_layout(n, cons)
    q = orc.qldecb(n, lay, aligned=True)
    q0 = orc.qldecb(n, lay)
    m = len(cons)
    assert np.array_equal(q["matV"], q0["matV"]) and np.array_equal(q["vecVk"], q0["vecVk"])
    last = lay["bcon"].reshape(-1, 3)[lay["nblk"] - 1, 2]
    if last == n or last < n - m + 1:
        assert np.array_equal(q["matL"], q0["matL"])
    Q = orc.qlmlq(q, np.eye(n), False)
    assert np.abs(Q.T @ Q - np.eye(n)).max() < 1e-14
    L = q["matL"].reshape(m, m).T
    assert np.array_equal(L, np.tril(L))
    C = _dense_c(n, cons, lay["order"])
    Z = np.zeros((n, m))
    Z[n - m:] = L
    assert np.abs(Q @ Z - C.T).max() < 1e-13 * max(1.0, np.abs(C).max())
    rng = np.random.default_rng(2)
    d = rng.normal(size=m)
    y = orc.qlbsub(q, d)
    assert np.abs(L.T @ y - d).max() < 1e-12
    X = rng.normal(size=(n, n))
    A = X @ X.T + np.eye(n)
    b = rng.normal(size=n)
    nf = n - m
    e = np.zeros(n)
    e[nf:] = y
    M = Q.T @ A @ Q
    bred = (Q.T @ (b - A @ (Q @ e)))[:nf]
    x = Q @ np.concatenate([np.linalg.solve(M[:nf, :nf], bred), y])
    K = np.block([[A, C.T], [C, np.zeros((m, m))]])
    xk = np.linalg.solve(K, np.concatenate([b, d]))[:n]
    assert np.abs(x - xk).max() < 1e-10 * max(1.0, np.abs(xk).max())


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("mb", [1, 3, 6])
def test_qlpssq_bitwise(orc, name, mb):
    """pede's band of Q^T A Q: the oracle's restatement against the reference's routine, and
    the summed form the GPU path evaluates (band of A + sum_k D_k v_k v_k^T - 2 Y_k v_k^T
    - 2 v_k Y_k^T, Y_k = A v_k + corrections, D_k = 4 (v^T A v + v^T A_k v)) against both."""
    n, cons = _case(name)
    lay = oracle.ql_layout(n, cons)
    R = _ref()
    R.ql_decompose(n, lay)
    q = orc.qldecb(n, lay)
    rng = np.random.default_rng(3)
    X = rng.normal(size=(n, n))
    A = X @ X.T + n * np.eye(n)
    band = _band(A, mb)
    br = R.qlpssq(A, band, mb, True)
    bo = orc.qlpssq(q, A, band, mb, True)
    assert np.array_equal(bo, br)
    # summed form (the coefficient recurrence of src/mpqldec.f90:878-935 on X = V^T A V)
    m = len(cons)
    V = _reflectors(q, n)
    G = V.T @ V
    vt = np.where(np.abs(G) > 16 * np.finfo(float).eps, G, 0.0)
    np.fill_diagonal(vt, 1.0)
    XA = V.T @ A @ V
    Cf, D = np.zeros((m, m)), np.zeros(m)
    lo, hi = [m - 1] * m, [0] * m
    for jj in range(m):
        k = m - 1 - jj
        vtAv = XA[k, k] + sum(Cf[k, i] * vt[k, i] for i in range(lo[k], hi[k] + 1))
        D[k] = 4 * XA[k, k] + 4 * vtAv
        for jj2 in range(jj + 1, m):
            k2 = m - 1 - jj2
            sv, m1, m2 = vt[k, k2], lo[k2], hi[k2]
            vta = XA[k, k2] + (sum(Cf[k2, i] * vt[k, i] for i in range(m1, m2 + 1)) if m1 <= m2 else 0.0)
            m1, m2 = min(m1, k), max(m2, k)
            Cf[k2, k] += 2 * (2 * vtAv * sv - vta)
            if sv != 0:
                m1, m2 = min(m1, lo[k]), max(m2, hi[k])
                for i in range(m1, m2 + 1):
                    Cf[k2, i] -= 2 * Cf[k, i] * sv
            lo[k2], hi[k2] = m1, m2
    Y = A @ V + V @ Cf.T
    Ms = A + V @ np.diag(D) @ V.T - 2 * Y @ V.T - 2 * V @ Y.T
    assert np.abs(_band(Ms, mb) - br).max() < 1e-13 * np.abs(br).max()


def test_reference_l_start(orc):
    """The reference copies the stored part of a column at or below its eliminated position
    into L starting at row iclast - ncol (src/mpqldec.f90:300-305); when the constraint block
    ends at parameter 28 of 30 and overlaps the eliminated positions 26..30, that start is two
    rows low: Q^T C^T (the exact [0; L]) and the reference's L then differ in the first columns,
    while the aligned start reproduces Q^T C^T.  Q itself is the same either way, and so is the
    solution whenever the constraint residuals vanish (pede corrects the parameters onto the
    constraints first, src/pede.f90:1560-1680)."""
    n, cons = _case("short_block")
    lay = oracle.ql_layout(n, cons)
    assert lay["nblk"] == 1 and lay["bcon"].reshape(-1, 3)[0, 2] == 28
    m = len(cons)
    q0, q1 = orc.qldecb(n, lay), orc.qldecb(n, lay, aligned=True)
    Q = orc.qlmlq(q0, np.eye(n), False)
    C = _dense_c(n, cons, lay["order"])
    bottom = (Q.T @ C.T)[n - m:]
    assert np.abs((Q.T @ C.T)[:n - m]).max() < 1e-13
    assert np.abs(q1["matL"].reshape(m, m).T - bottom).max() < 1e-13
    assert np.abs(q0["matL"].reshape(m, m).T - bottom).max() > 0.1
