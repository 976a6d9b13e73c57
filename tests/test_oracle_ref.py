"""Pin the oracle's numeric routines against the reference's own Fortran routines
(src/mpnum.f90, src/minresModule.f90 compiled by flang into oracle/_ref) and against
closed forms.  CPU only."""
import numpy as np
import pytest

from helpers import dense_to_packed, packed_to_dense


def spd(rng, n, cond=1e3):
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    ev = np.geomspace(1.0, cond, n)
    return (Q * ev) @ Q.T


@pytest.mark.parametrize("n", [1, 2, 3, 4, 7, 16, 33, 64])
def test_sqminv_bitwise(orc, ref, n):
    rng = np.random.default_rng(n)
    A = spd(rng, n)
    v, b = dense_to_packed(A), rng.normal(size=n)
    a1, a2 = orc.sqminv(v, b), ref.sqminv(v, b)
    assert a1[2] == a2[2] == n
    assert np.array_equal(a1[0], a2[0]) and np.array_equal(a1[1], a2[1])
    # closed form: inverse and solution
    Ai = packed_to_dense(a1[0], n)
    assert np.allclose(Ai @ A, np.eye(n), atol=1e-9)
    assert np.allclose(A @ a1[1], b, rtol=1e-9, atol=1e-9)


def test_sqminv_rank_defect(orc, ref):
    rng = np.random.default_rng(7)
    X = rng.normal(size=(6, 4))
    A = X @ X.T                                # rank 4 of 6
    v, b = dense_to_packed(A), rng.normal(size=6)
    a1, a2 = orc.sqminv(v, b), ref.sqminv(v, b)
    assert a1[2] == a2[2] == 4
    assert np.array_equal(a1[0], a2[0]) and np.array_equal(a1[1], a2[1])


@pytest.mark.parametrize("n", [1, 5, 30, 100])
def test_sqminl_bitwise(orc, ref, n):
    rng = np.random.default_rng(100 + n)
    A = spd(rng, n, 1e6)
    v, b = dense_to_packed(A), rng.normal(size=n)
    a1, a2 = orc.sqminl(v, b), ref.sqminl(v, b)
    assert a1[2] == a2[2] == n
    assert np.array_equal(a1[0], a2[0]) and np.array_equal(a1[1], a2[1])
    assert np.allclose(a1[1], np.linalg.solve(A, b), rtol=1e-7)


def test_sqminl_kkt(orc, ref):
    """Lagrange multipliers: indefinite KKT matrix (zero diagonal in the constraint rows)."""
    rng = np.random.default_rng(3)
    n, m = 12, 3
    A = spd(rng, n)
    Cm = rng.normal(size=(m, n))
    K = np.block([[A, Cm.T], [Cm, np.zeros((m, m))]])
    v, b = dense_to_packed(K), rng.normal(size=n + m)
    a1, a2 = orc.sqminl(v, b), ref.sqminl(v, b)
    assert a1[2] == a2[2] == n + m
    assert np.array_equal(a1[1], a2[1])
    assert np.allclose(a1[1], np.linalg.solve(K, b), rtol=1e-8)


@pytest.mark.parametrize("n", [1, 4, 20, 90])
def test_chdec2_bitwise(orc, ref, n):
    rng = np.random.default_rng(200 + n)
    A = spd(rng, n)
    v, b = dense_to_packed(A), rng.normal(size=n)
    c1, c2 = orc.chdec2(v, n), ref.chdec2(v, n)
    assert c1[1] == c2[1] == n
    assert np.array_equal(c1[0], c2[0])
    x1, x2 = orc.chslv2(c1[0], b), ref.chslv2(c2[0], b)
    assert np.array_equal(x1, x2)
    assert np.allclose(x1, np.linalg.solve(A, b), rtol=1e-8)


def test_chindl_table(orc, ref):
    for m in (1, 2, 3):
        for nd in list(range(0, 40)) + [50, 99, 1000]:
            assert orc.chindl(m, nd) == ref.chindl(m, nd)
    # table anchors from the CI log (CI/refFiles/millepede.log:67-101): Chi^2/Ndf(3) at ndf 1, 10, 100
    assert abs(orc.chindl(3, 1) - 9.0) < 1e-6
    assert abs(orc.chindl(3, 10) - 2.690) < 1e-3
    assert abs(orc.chindl(3, 100) - 1.426) < 1e-3


@pytest.mark.parametrize("n,m", [(1, 1), (2, 5), (4, 20), (7, 3)])
def test_dbavat_bitwise(orc, ref, n, m):
    rng = np.random.default_rng(n * 10 + m)
    V = spd(rng, n)
    v = dense_to_packed(V)
    a = rng.normal(size=n * m)
    w = rng.normal(size=m * (m + 1) // 2)
    w1, w2 = orc.dbavat(v, a, w, n, m), ref.dbavat(v, a, w, n, m)
    assert np.array_equal(w1, w2)
    G = a.reshape(m, n)
    assert np.allclose(packed_to_dense(w1 - w, m), G @ V @ G.T, rtol=1e-10, atol=1e-12)


def _band(n, mb, A):
    india = np.zeros(n, np.int32)
    for i in range(1, min(mb, n) + 1):
        india[i - 1] = (i * i + i) // 2
    for i in range(min(mb, n) + 1, n + 1):
        india[i - 1] = india[i - 2] + mb
    c = np.zeros(india[-1])
    for i in range(1, n + 1):
        for j in range(max(1, i - mb + 1), i + 1):
            c[india[i - 1] - i + j - 1] = A[i - 1, j - 1]
    return c, india


def test_band_preconditioner_bitwise(orc, ref):
    """lltdec/equdecs/equslvs with constraint blocks (src/mpnum.f90:2155-2678)."""
    import ctypes as C
    rng = np.random.default_rng(11)
    n, mb, m = 40, 4, 3
    A = spd(rng, n, 50.0)
    c, india = _band(n, mb, A)
    # constraint blocks: constraints 1-2 on params 5..14, constraint 3 on params 20..30
    l = np.array([1, 2, 4, 10 + mb - 1, 3, 3, 19, 11 + mb - 1], np.int32)
    kpart = []
    for (p1, p2, i0, nin) in l.reshape(-1, 4):
        for j in range(p1, p2 + 1):
            row = np.zeros(nin)
            row[: nin - (mb - 1)] = rng.normal(size=nin - (mb - 1))
            kpart.append(row)
    cc = np.concatenate([c, np.zeros(m * (m + 1) // 2), np.concatenate(kpart)])
    nm = len(np.concatenate(kpart))
    ind = np.concatenate([india, np.zeros(m, np.int32)])
    r_c, r_ind, r_n1, r_n2 = ref.equdecs(n, m, 2, nm, 0, cc, ind, l)
    o_c = cc.copy()
    o_ind = ind.copy()
    n1, n2 = C.c_int32(), C.c_int32()
    orc.lib.orc_equdecs(n, m, 2, 0, o_c.ctypes.data_as(C.POINTER(C.c_double)),
                        o_ind.ctypes.data_as(C.POINTER(C.c_int32)), l.ctypes.data_as(C.POINTER(C.c_int32)),
                        C.byref(n1), C.byref(n2))
    assert np.array_equal(o_c, r_c) and np.array_equal(o_ind, r_ind)
    assert (n1.value, n2.value) == (r_n1, r_n2)
    x = rng.normal(size=n + m)
    rx = ref.equslvs(n, m, 2, nm, r_c, r_ind, l, x)
    ox = x.copy()
    orc.lib.orc_equslvs(n, m, 2, o_c.ctypes.data_as(C.POINTER(C.c_double)),
                        o_ind.ctypes.data_as(C.POINTER(C.c_int32)), l.ctypes.data_as(C.POINTER(C.c_int32)),
                        ox.ctypes.data_as(C.POINTER(C.c_double)))
    assert np.array_equal(ox, rx)


def test_diag_preconditioner_bitwise(orc, ref):
    """precons/presols (src/mpnum.f90:2881-3069)."""
    import ctypes as C
    rng = np.random.default_rng(12)
    n, p = 30, 2
    d = rng.uniform(0.5, 2.0, size=n)
    l = np.array([1, 2, 3, 12], np.int32)
    a = rng.normal(size=2 * 12)
    nm = len(a)
    r_cu, r_a, r_s, r_nr = ref.precons(p, n, 1, nm, d, a, l)
    cu = np.zeros(n)
    oa = a.copy()
    s = np.zeros(3)
    nr = C.c_int32()
    P = lambda x, t=C.c_double: x.ctypes.data_as(C.POINTER(t))
    orc.lib.orc_precons(p, n, 1, P(d), P(cu), P(oa), P(l, C.c_int32), P(s), C.byref(nr))
    assert np.array_equal(cu, r_cu) and np.array_equal(oa, r_a) and np.array_equal(s, r_s)
    y = rng.normal(size=n + p)
    rx = ref.presols(p, n, 1, nm, r_cu, r_a, l, r_s, y)
    ox = np.zeros(n + p)
    orc.lib.orc_presols(p, n, 1, P(cu), P(oa), P(l, C.c_int32), P(s), P(ox), P(y))
    assert np.array_equal(ox, rx)


@pytest.mark.parametrize("precon", [False, True])
def test_minres_matches_reference(orc, ref, precon):
    rng = np.random.default_rng(5)
    n = 80
    A = spd(rng, n, 1e4)
    b = rng.normal(size=n)
    d = np.diag(A).copy()
    ms = (lambda x: x / d) if precon else None
    x1, s1, i1, o1 = orc.minres(n, lambda x: A @ x, ms, b, 1e-10, precon=precon)
    x2, s2, i2, o2 = ref.minres(n, lambda x: A @ x, ms, b, 1e-10, precon=precon)
    assert (s1, i1) == (s2, i2)
    assert np.array_equal(x1, x2)
    assert np.allclose(x1, np.linalg.solve(A, b), rtol=1e-6)


def test_minres_kkt_indefinite(orc, ref):
    rng = np.random.default_rng(6)
    n, m = 40, 4
    A = spd(rng, n, 100.0)
    Cm = rng.normal(size=(m, n))
    K = np.block([[A, Cm.T], [Cm, np.zeros((m, m))]])
    b = rng.normal(size=n + m)
    x1, s1, i1, _ = orc.minres(n + m, lambda x: K @ x, None, b, 1e-12, precon=False)
    x2, s2, i2, _ = ref.minres(n + m, lambda x: K @ x, None, b, 1e-12, precon=False)
    assert (s1, i1) == (s2, i2) and np.array_equal(x1, x2)
    assert np.allclose(x1, np.linalg.solve(K, b), rtol=1e-7, atol=1e-9)


def _qlp_case(kind, rng):
    if kind == "spd":
        n = 70
        A = spd(rng, n, 1e5)
    elif kind == "kkt":
        n0, m = 40, 5
        A0 = spd(rng, n0, 100.0)
        Cm = rng.normal(size=(m, n0))
        A = np.block([[A0, Cm.T], [Cm, np.zeros((m, m))]])
        n = n0 + m
    else:                                       # singular, consistent right-hand side
        n = 50
        X = rng.normal(size=(n, n - 3))
        A = X @ X.T
    b = A @ rng.normal(size=n) if kind == "singular" else rng.normal(size=n)
    return n, A, b


@pytest.mark.parametrize("kind", ["spd", "kkt", "singular"])
@pytest.mark.parametrize("trancond", [1.0e7, 1.0, 1.0e2])
@pytest.mark.parametrize("precon", [False, True])
def test_minresqlp_matches_reference(orc, ref, kind, trancond, precon):
    """MINRES-QLP (src/minresqlpModule.f90:411-1330) in the oracle against the reference's own
    routine: trancond 1e7 is pede's default (QR, QLP once ill-conditioned), 1 forces QLP from
    the first iteration (mresmode 2), 1e2 switches mid-run."""
    rng = np.random.default_rng(11 + len(kind))
    n, A, b = _qlp_case(kind, rng)
    d = np.abs(np.diag(A)) + 1.0
    ms = (lambda x: x / d) if precon else None
    x1, s1, i1, o1 = orc.minresqlp(n, lambda x: A @ x, ms, b, 1e-10, precon=precon, trancond=trancond)
    x2, s2, i2, o2 = ref.minresqlp(n, lambda x: A @ x, ms, b, 1e-10, precon=precon, trancond=trancond)
    assert (s1, i1) == (s2, i2)
    np.testing.assert_allclose(x1, x2, rtol=1e-13, atol=1e-13 * np.abs(x2).max())
    np.testing.assert_allclose(o1[:5], o2, rtol=1e-12)
    # stopping rule istop 4: ||r|| <= rtol (||A|| ||x|| + ||b||) for the recurrence residual
    assert s1 == 4
    scale = np.linalg.norm(A, 2) * np.linalg.norm(x1) + np.linalg.norm(b)
    assert np.linalg.norm(A @ x1 - b) <= 1e-8 * scale
    if kind != "singular":
        assert np.allclose(x1, np.linalg.solve(A, b), rtol=1e-4, atol=1e-5 * np.abs(x1).max())
    elif not precon:
        # minimum-length solution of a consistent singular system: x in range(A)
        xp = np.linalg.pinv(A) @ b
        assert np.allclose(x1, xp, rtol=1e-5, atol=1e-6 * np.abs(xp).max())


def test_minresqlp_early_exits(orc, ref):
    """b = 0 (istop 3) and an unsymmetric operator (istop 9) stop before the first iteration."""
    rng = np.random.default_rng(3)
    A = spd(rng, 10)
    x1, s1, i1, _ = orc.minresqlp(10, lambda x: A @ x, None, np.zeros(10), 1e-10, precon=False)
    x2, s2, i2, _ = ref.minresqlp(10, lambda x: A @ x, None, np.zeros(10), 1e-10, precon=False)
    assert (s1, i1) == (s2, i2) == (3, 0) and not x1.any() and not x2.any()
    U = A + np.triu(rng.normal(size=(10, 10)), 1)
    b = rng.normal(size=10)
    x1, s1, i1, _ = orc.minresqlp(10, lambda x: U @ x, None, b, 1e-10, precon=False)
    x2, s2, i2, _ = ref.minresqlp(10, lambda x: U @ x, None, b, 1e-10, precon=False)
    assert (s1, i1) == (s2, i2) == (9, 0)


def test_packed_matvec(orc):
    rng = np.random.default_rng(8)
    A = spd(rng, 25)
    x = rng.normal(size=25)
    assert np.allclose(orc.packed_matvec(dense_to_packed(A), x), A @ x, rtol=1e-12)


def bordered_band(rng, n, nbdr, nbnd, side):
    """SPD matrix with a band of half-width nbnd plus nbdr dense border rows/columns."""
    A = np.zeros((n, n))
    nmb = n - nbdr
    b0 = nbdr if side == 1 else 0
    for i in range(nmb):
        for j in range(max(0, i - nbnd), i + 1):
            A[b0 + i, b0 + j] = A[b0 + j, b0 + i] = rng.normal()
    border = list(range(nbdr)) if side == 1 else list(range(nmb, n))
    for i in border:
        A[i, :] = rng.normal(size=n)
        A[:, i] = A[i, :]
    return A @ A.T * 0 + (A + A.T) / 2 + n * np.eye(n)


@pytest.mark.parametrize("side", [1, 2])
@pytest.mark.parametrize("inv", [0, 1, 2])
@pytest.mark.parametrize("n,nbdr,nbnd", [(10, 0, 2), (24, 4, 5), (60, 4, 5), (13, 2, 1)])
def test_sqmibb_bitwise(orc, ref, side, inv, n, nbdr, nbnd):
    rng = np.random.default_rng(n * 100 + nbdr * 10 + nbnd + side)
    A = bordered_band(rng, n, nbdr, nbnd, side)
    v, b = dense_to_packed(A), rng.normal(size=n)
    o = orc.sqmibb(v, b, nbdr, nbnd, inv, side)
    r = ref.sqmibb(v, b, nbdr, nbnd, inv, side)
    assert o[2] == r[2] == n
    assert np.array_equal(o[1], r[1])
    assert (o[3], o[4]) == (r[3], r[4])
    if inv > 0:
        assert np.array_equal(o[0], r[0])
    assert np.allclose(o[1], np.linalg.solve(A, b), rtol=1e-9)
    if inv == 2:
        assert np.allclose(packed_to_dense(o[0], n), np.linalg.inv(A), rtol=1e-8, atol=1e-12)
