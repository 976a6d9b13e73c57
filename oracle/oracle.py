"""ctypes access to the CPU oracle (pede_oracle.c) and to the reference's own numeric
routines compiled from src/*.f90 (oracle/_ref/libpedenum.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product library.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF = os.path.join(HERE, "_ref", "libpedenum.so")
REF_GEN = os.path.join(HERE, "_ref", "mptest1_gen")

D = C.POINTER(C.c_double)
I32 = C.POINTER(C.c_int32)
I64 = C.POINTER(C.c_int64)


def build(reference_root="/root/reference"):
    """Compile the oracle (gcc) and, when the reference sources are present, the reference's
    numeric routines (flang) into oracle/_ref."""
    subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-o", LIB + ".tmp",
                    os.path.join(HERE, "pede_oracle.c"), "-lm"], check=True)
    os.replace(LIB + ".tmp", LIB)
    if os.path.isdir(os.path.join(reference_root, "src")):
        subprocess.run(["make", "-s", "-f", os.path.join(HERE, "Makefile.ref"), f"REF={reference_root}"],
                       check=True, cwd=HERE)


def _p(a, t):
    return a.ctypes.data_as(t)


def ql_layout(nvar, cons):
    """pede's constraint layout for the QL decomposition (src/pede.f90:840-1003 grpcon,
    1077-1371 prpcon, 1423-1460 feasma) for constraints over variable parameters.

    cons: list of (indices (1-based variable index), factors).  Constraints are sorted by their
    first and last parameter (ties in input order), split into groups sharing no parameter
    (numbered by first appearance), groups merged into blocks of overlapping ranges.  Returns
    the block-compressed matrix a, bcon (3 x nblk+1), rcon (4 x ncon), bpar (2 x 2) as flat
    column-major int32 arrays, and the input constraint of each QL column."""
    m = len(cons)
    first = [min(ix) for ix, _ in cons]
    last = [max(ix) for ix, _ in cons]
    srt = sorted(range(m), key=lambda j: (first[j], last[j], j))
    parent = list(range(m))

    def root(i):
        while parent[i] != i:
            parent[i] = parent[parent[i]]
            i = parent[i]
        return i

    owner = {}
    for j in range(m):
        for v in cons[j][0]:
            if v in owner:
                a, b = root(j), root(owner[v])
                if a != b:
                    parent[max(a, b)] = min(a, b)
            else:
                owner[v] = j
    gid, grp = {}, [0] * m
    for j in srt:
        r = root(j)
        if r not in gid:
            gid[r] = len(gid)
        grp[j] = gid[r]
    pos = {j: q for q, j in enumerate(srt)}
    order = sorted(range(m), key=lambda j: (grp[j], pos[j]))
    vf = [first[j] for j in order]
    vl = [last[j] for j in order]
    gstart, gf, gl = [], [], []
    for q in range(m):
        if q == 0 or grp[order[q]] != grp[order[q - 1]]:
            gstart.append(q)
            gf.append(vf[q])
            gl.append(vl[q])
        else:
            gf[-1] = min(gf[-1], vf[q])
            gl[-1] = max(gl[-1], vl[q])
    ng = len(gstart)
    gstart.append(m)
    rcon = np.zeros((m, 4), np.int32)
    bcon = []
    ilast, isblck = 0, 0
    for g in range(ng):
        ilast = max(ilast, gl[g])
        nxt = gf[g + 1] if g + 1 < ng else 1 << 30
        if nxt > ilast:
            bcon.append((gstart[isblck] + 1, gf[isblck], ilast))
            jfrst = gf[g]
            for i in range(g, isblck - 1, -1):
                jfrst = min(jfrst, gf[i])
                for q in range(gstart[i], gstart[i + 1]):
                    rcon[q] = (gf[i], gl[i], jfrst, ilast)
            isblck = g + 1
    nblk = len(bcon)
    bcon.append((m + 1, 0, 0))
    ioff = [0]
    for q in range(m):
        ioff.append(ioff[-1] + rcon[q, 3] - (rcon[q, 2] - 1))
    a = np.zeros(max(1, ioff[-1]))
    for q in range(m):
        for v, f in zip(*cons[order[q]]):
            a[ioff[q] + v - rcon[q, 2]] += f
    bpar = np.array([0, 0, nvar, nblk], np.int32)
    return {"ncon": m, "nblk": nblk, "a": a, "bcon": np.array(bcon, np.int32).ravel(), "rcon": rcon.ravel(),
            "bpar": bpar, "order": order, "matsize": ioff[-1]}


class _RecStruct(C.Structure):
    _fields_ = [("nrec", C.c_int64), ("rec_ptr", I64), ("inder", I32), ("glder", D),
                ("ntot", C.c_int32), ("labels", I32), ("var_index", I32), ("global_par", D),
                ("nvar", C.c_int32)]


class _PassOpts(C.Structure):
    _fields_ = [("icalcm", C.c_int32), ("nloopn", C.c_int32), ("lhuber", C.c_int32), ("lfitbb", C.c_int32),
                ("chicut", C.c_float), ("chhuge", C.c_float), ("dwcut", C.c_float), ("cndlmx", C.c_float),
                ("nthreads", C.c_int32)]


class _PassStats(C.Structure):
    _fields_ = [("chi2", C.c_double), ("ndf", C.c_int64), ("wndf", C.c_double), ("nrej", C.c_int64 * 6),
                ("naccept", C.c_int64), ("nbandrec", C.c_int64)]


OP = C.CFUNCTYPE(None, C.c_void_p, C.c_int, D, D)
APS = C.CFUNCTYPE(None, C.c_void_p, C.c_int, D, C.c_int, C.c_int, D)


class Oracle:
    def __init__(self):
        if not os.path.exists(LIB):
            build()
        self.lib = C.CDLL(LIB)
        L = self.lib
        L.orc_chindl.restype = C.c_float
        L.orc_chindl.argtypes = [C.c_int, C.c_int]
        L.orc_sqminv.argtypes = [D, D, C.c_int, I32]
        L.orc_sqminl.argtypes = [D, D, C.c_int, I32]
        L.orc_chdec2.argtypes = [D, C.c_int, I32, D, D]
        L.orc_chslv2.argtypes = [D, D, C.c_int]
        L.orc_dbavat.argtypes = [D, D, D, C.c_int, C.c_int]
        L.orc_lltdec.argtypes = [C.c_int, D, I32, I32, C.c_int]
        L.orc_lltfwd.argtypes = [C.c_int, D, I32, D]
        L.orc_lltbwd.argtypes = [C.c_int, D, I32, D]
        L.orc_equdecs.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, D, I32, I32, I32, I32]
        L.orc_equslvs.argtypes = [C.c_int, C.c_int, C.c_int, D, I32, I32, D]
        L.orc_precons.argtypes = [C.c_int, C.c_int, C.c_int, D, D, D, I32, D, I32]
        L.orc_presols.argtypes = [C.c_int, C.c_int, C.c_int, D, D, I32, D, D, D]
        L.orc_minres.argtypes = [C.c_int, OP, OP, C.c_void_p, D, C.c_double, C.c_int, D, C.c_int, C.c_double,
                                 I32, I32, D]
        L.orc_minresqlp.argtypes = [C.c_int, OP, OP, C.c_void_p, D, C.c_double, C.c_int, D, C.c_int, C.c_double,
                                    C.c_double, C.c_double, C.c_double, I32, I32, D]
        L.orc_packed_matvec.argtypes = [C.c_int, D, D, D]
        L.orc_loopn.argtypes = [C.POINTER(_RecStruct), C.POINTER(_PassOpts), D, D, D, I64, C.POINTER(_PassStats)]
        L.orc_loopn.restype = C.c_int
        L.orc_loopn_blocks.argtypes = [C.POINTER(_RecStruct), C.POINTER(_PassOpts), C.c_int, C.c_int32, I32, I64, I32,
                                        I64, C.c_int64, D, D, I64, C.POINTER(_PassStats)]
        L.orc_loopn_hash.argtypes = [C.POINTER(_RecStruct), C.POINTER(_PassOpts), C.POINTER(C.c_uint64), D,
                                     C.c_int64, D, I64, C.POINTER(_PassStats), I64]
        L.orc_loopn_hash.restype = C.c_int
        L.orc_sqmibb.argtypes = [D, D, C.c_int, C.c_int, C.c_int, C.c_int, I32, D, D, C.c_int]
        L.orc_read_mille.argtypes = [C.c_char_p, C.c_int, C.POINTER(I64), C.POINTER(I32), C.POINTER(D)]
        L.orc_read_mille.restype = C.c_int64
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_qldecb.argtypes = [C.c_int, C.c_int, C.c_int, D, I32, I32, D, I64, I32, I32, D, D, C.c_int]
        L.orc_qlmlq.argtypes = [C.c_int, C.c_int, D, I64, I32, D, D, C.c_int, C.c_int]
        L.orc_qlbsub.argtypes = [C.c_int, D, D, D]
        L.orc_qlpssq.argtypes = [C.c_int, C.c_int, D, I64, I32, D, APS, C.c_void_p, D, C.c_int, C.c_int]

    # ---- constraint elimination (QL) ------------------------------------------------------------
    def qldecb(self, npar, lay, aligned=False):
        """QL decomposition of a constraint layout from ql_layout (one parameter block); aligned
        selects the corrected start of L (see orc_qldecb)."""
        ncon = lay["ncon"]
        a = np.ascontiguousarray(lay["a"], dtype=np.float64)
        q = {"npar": npar, "ncon": ncon, "matV": np.zeros(max(1, len(a))), "ioffRow": np.zeros(ncon + 1, np.int64),
             "ioffPar": np.zeros(ncon, np.int32), "irng": np.zeros(2 * ncon, np.int32), "vecVk": np.zeros(ncon),
             "matL": np.zeros(ncon * ncon)}
        self.lib.orc_qldecb(npar, ncon, lay["nblk"], _p(a, D), _p(lay["bcon"], I32), _p(lay["rcon"], I32),
                            _p(q["matV"], D), _p(q["ioffRow"], I64), _p(q["ioffPar"], I32), _p(q["irng"], I32),
                            _p(q["vecVk"], D), _p(q["matL"], D), 1 if aligned else 0)
        return q

    def qlmlq(self, q, x, t):
        """Q x (t False) or Q^T x (t True); x is npar or npar x m (columns)."""
        x = np.array(x, dtype=np.float64, order="F", copy=True)
        m = 1 if x.ndim == 1 else x.shape[1]
        self.lib.orc_qlmlq(q["npar"], q["ncon"], _p(q["matV"], D), _p(q["ioffRow"], I64), _p(q["irng"], I32),
                           _p(q["vecVk"], D), _p(x, D), m, 1 if t else 0)
        return x

    def qlbsub(self, q, d):
        d = np.ascontiguousarray(d, dtype=np.float64)
        y = np.zeros(q["ncon"])
        self.lib.orc_qlbsub(q["ncon"], _p(q["matL"], D), _p(d, D), _p(y, D))
        return y

    def qlpssq(self, q, A, band, mb, t=True):
        """pede's band of Q^T A Q (see orc_qlpssq) for a dense symmetric A."""
        A = np.asarray(A, dtype=np.float64)

        def _ap(ctx, nn, xp, i0, i1, yp):
            xa = np.ctypeslib.as_array(xp, shape=(nn,))
            ya = np.ctypeslib.as_array(yp, shape=(nn,))
            ya[:] += A[:, i0 - 1:i1] @ xa[i0 - 1:i1]

        fa = APS(_ap)
        b = np.array(band, dtype=np.float64, copy=True)
        self.lib.orc_qlpssq(q["npar"], q["ncon"], _p(q["matV"], D), _p(q["ioffRow"], I64), _p(q["irng"], I32),
                            _p(q["vecVk"], D), fa, None, _p(b, D), mb, 1 if t else 0)
        return b

    # ---- dense numerics ------------------------------------------------------------------------
    def chindl(self, n, nd):
        return self.lib.orc_chindl(n, nd)

    def sqminv(self, v, b):
        v = np.ascontiguousarray(v, dtype=np.float64).copy()
        b = np.ascontiguousarray(b, dtype=np.float64).copy()
        r = C.c_int32()
        self.lib.orc_sqminv(_p(v, D), _p(b, D), len(b), C.byref(r))
        return v, b, r.value

    def sqminl(self, v, b):
        v = np.ascontiguousarray(v, dtype=np.float64).copy()
        b = np.ascontiguousarray(b, dtype=np.float64).copy()
        r = C.c_int32()
        self.lib.orc_sqminl(_p(v, D), _p(b, D), len(b), C.byref(r))
        return v, b, r.value

    def chdec2(self, g, n):
        g = np.ascontiguousarray(g, dtype=np.float64).copy()
        r = C.c_int32()
        e1, e2 = C.c_double(), C.c_double()
        self.lib.orc_chdec2(_p(g, D), n, C.byref(r), C.byref(e1), C.byref(e2))
        return g, r.value, e1.value, e2.value

    def chslv2(self, g, x):
        x = np.ascontiguousarray(x, dtype=np.float64).copy()
        g = np.ascontiguousarray(g, dtype=np.float64)
        self.lib.orc_chslv2(_p(g, D), _p(x, D), len(x))
        return x

    def dbavat(self, v, a, w, n, m):
        w = np.ascontiguousarray(w, dtype=np.float64).copy()
        v = np.ascontiguousarray(v, dtype=np.float64)
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.lib.orc_dbavat(_p(v, D), _p(a, D), _p(w, D), n, m)
        return w

    def sqmibb(self, v, b, nbdr, nbnd, inv, side=1):
        v = np.ascontiguousarray(v, dtype=np.float64).copy()
        b = np.ascontiguousarray(b, dtype=np.float64).copy()
        r = C.c_int32()
        e1, e2 = C.c_double(), C.c_double()
        self.lib.orc_sqmibb(_p(v, D), _p(b, D), len(b), nbdr, nbnd, inv, C.byref(r), C.byref(e1), C.byref(e2), side)
        return v, b, r.value, e1.value, e2.value

    def packed_matvec(self, a, x):
        a = np.ascontiguousarray(a, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        self.lib.orc_packed_matvec(len(x), _p(a, D), _p(x, D), _p(y, D))
        return y

    def minres(self, n, aprod, msolve, b, rtol, itnlim=2000, precon=True):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(n)

        def _ap(ctx, nn, xp, yp):
            xa = np.ctypeslib.as_array(xp, shape=(nn,))
            ya = np.ctypeslib.as_array(yp, shape=(nn,))
            ya[:] = aprod(xa)

        def _ms(ctx, nn, xp, yp):
            xa = np.ctypeslib.as_array(xp, shape=(nn,))
            ya = np.ctypeslib.as_array(yp, shape=(nn,))
            ya[:] = msolve(xa) if msolve else xa

        fa, fm = OP(_ap), OP(_ms)
        istop, itn = C.c_int32(), C.c_int32()
        out = np.zeros(5)
        self.lib.orc_minres(n, fa, fm, None, _p(b, D), 0.0, 1 if precon else 0, _p(x, D), itnlim, rtol,
                            C.byref(istop), C.byref(itn), _p(out, D))
        return x, istop.value, itn.value, out

    def minresqlp(self, n, aprod, msolve, b, rtol, itnlim=2000, precon=True, maxxnorm=None, trancond=1.0e7,
                  acondlim=0.0, shift=0.0):
        """MINRES-QLP as mminrsqlp calls it; maxxnorm defaults to n/sqrt(eps) (src/pede.f90:8365)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(n)
        if maxxnorm is None:
            maxxnorm = n / np.sqrt(np.finfo(float).eps)

        def _ap(ctx, nn, xp, yp):
            xa = np.ctypeslib.as_array(xp, shape=(nn,))
            ya = np.ctypeslib.as_array(yp, shape=(nn,))
            ya[:] = aprod(xa)

        def _ms(ctx, nn, xp, yp):
            xa = np.ctypeslib.as_array(xp, shape=(nn,))
            ya = np.ctypeslib.as_array(yp, shape=(nn,))
            ya[:] = msolve(xa) if msolve else xa

        fa, fm = OP(_ap), OP(_ms)
        istop, itn = C.c_int32(), C.c_int32()
        out = np.zeros(6)
        self.lib.orc_minresqlp(n, fa, fm, None, _p(b, D), shift, 1 if precon else 0, _p(x, D), itnlim, rtol, maxxnorm,
                               trancond, acondlim, C.byref(istop), C.byref(itn), _p(out, D))
        return x, istop.value, itn.value, out

    # ---- record loop ---------------------------------------------------------------------------
    def loopn(self, rec_ptr, inder, glder, labels, var_index, par, nvar, icalcm=1, nloopn=1, lhuber=0,
              chicut=0.0, chhuge=50.0, dwcut=0.0, cndlmx=100.0, lfitbb=1, pre_weight=None):
        rec_ptr = np.ascontiguousarray(rec_ptr, dtype=np.int64)
        inder = np.ascontiguousarray(inder, dtype=np.int32)
        glder = np.ascontiguousarray(glder, dtype=np.float64)
        labels = np.ascontiguousarray(labels, dtype=np.int32)
        var_index = np.ascontiguousarray(var_index, dtype=np.int32)
        par = np.ascontiguousarray(par, dtype=np.float64)
        R = _RecStruct(len(rec_ptr) - 1, _p(rec_ptr, I64), _p(inder, I32), _p(glder, D), len(labels),
                       _p(labels, I32), _p(var_index, I32), _p(par, D), nvar)
        O = _PassOpts(icalcm, nloopn, lhuber, lfitbb, chicut, chhuge, dwcut, cndlmx, 1)
        S = _PassStats()
        mat = np.zeros(nvar * (nvar + 1) // 2 if icalcm == 1 else 1)
        vec = np.zeros(nvar)
        cnt = np.zeros(nvar, dtype=np.int64)
        pw = None if pre_weight is None else np.ascontiguousarray(pre_weight, dtype=np.float64)
        rc = self.lib.orc_loopn(C.byref(R), C.byref(O), _p(pw, D) if pw is not None else None,
                                _p(mat, D) if icalcm == 1 else None, _p(vec, D), _p(cnt, I64), C.byref(S))
        if rc != 0:
            raise RuntimeError(f"oracle loopn failed with code {rc}")
        stats = dict(chi2=S.chi2, ndf=S.ndf, nrej=list(S.nrej), naccept=S.naccept, nband=S.nbandrec)
        return mat, vec, cnt, stats

    def loopn_hash(self, rec_ptr, inder, glder, labels, var_index, par, nvar, cap_log2=24, **kw):
        """Record pass accumulating the matrix into a hash table; returns (keys, vals, vec, cnt, stats)
        with keys = i<<32 | j (1-based, i >= j) of the used slots."""
        rec_ptr = np.ascontiguousarray(rec_ptr, dtype=np.int64)
        inder = np.ascontiguousarray(inder, dtype=np.int32)
        glder = np.ascontiguousarray(glder, dtype=np.float64)
        labels = np.ascontiguousarray(labels, dtype=np.int32)
        var_index = np.ascontiguousarray(var_index, dtype=np.int32)
        par = np.ascontiguousarray(par, dtype=np.float64)
        R = _RecStruct(len(rec_ptr) - 1, _p(rec_ptr, I64), _p(inder, I32), _p(glder, D), len(labels),
                       _p(labels, I32), _p(var_index, I32), _p(par, D), nvar)
        O = _PassOpts(kw.get("icalcm", 1), kw.get("nloopn", 1), kw.get("lhuber", 0), kw.get("lfitbb", 1),
                      kw.get("chicut", 0.0), kw.get("chhuge", 50.0), kw.get("dwcut", 0.0), kw.get("cndlmx", 100.0), 1)
        S = _PassStats()
        cap = 1 << cap_log2
        keys = np.zeros(cap, np.uint64)
        vals = np.zeros(cap)
        vec = np.zeros(nvar)
        cnt = np.zeros(nvar, np.int64)
        used = C.c_int64()
        rc = self.lib.orc_loopn_hash(C.byref(R), C.byref(O), keys.ctypes.data_as(C.POINTER(C.c_uint64)), _p(vals, D),
                                     cap, _p(vec, D), _p(cnt, I64), C.byref(S), C.byref(used))
        if rc != 0:
            raise RuntimeError(f"oracle loopn_hash failed with code {rc}")
        m = keys != 0
        stats = dict(chi2=S.chi2, ndf=S.ndf, nrej=list(S.nrej), naccept=S.naccept, nband=S.nbandrec, used=used.value)
        return keys[m], vals[m], vec, cnt, stats

    def loopn_blocks(self, rec_ptr, inder, glder, labels, var_index, par, nvar, blocks, nthreads, **kw):
        """Record pass on nthreads threads into the device's group-block layout (blocks: gfirst,
        brow, bcol, boff, nstored as Pede.block_structure returns them); returns (blk, vec, cnt,
        stats).  The CPU baseline of the timing."""
        rec_ptr = np.ascontiguousarray(rec_ptr, dtype=np.int64)
        inder = np.ascontiguousarray(inder, dtype=np.int32)
        glder = np.ascontiguousarray(glder, dtype=np.float64)
        labels = np.ascontiguousarray(labels, dtype=np.int32)
        var_index = np.ascontiguousarray(var_index, dtype=np.int32)
        par = np.ascontiguousarray(par, dtype=np.float64)
        R = _RecStruct(len(rec_ptr) - 1, _p(rec_ptr, I64), _p(inder, I32), _p(glder, D), len(labels),
                       _p(labels, I32), _p(var_index, I32), _p(par, D), nvar)
        O = _PassOpts(kw.get("icalcm", 1), kw.get("nloopn", 1), kw.get("lhuber", 0), kw.get("lfitbb", 1),
                      kw.get("chicut", 0.0), kw.get("chhuge", 50.0), kw.get("dwcut", 0.0), kw.get("cndlmx", 100.0),
                      nthreads)
        gfirst, brow, bcol, boff, nstored = blocks
        gfirst = np.ascontiguousarray(gfirst, dtype=np.int32)
        brow = np.ascontiguousarray(brow, dtype=np.int64)
        bcol = np.ascontiguousarray(bcol, dtype=np.int32)
        boff = np.ascontiguousarray(boff, dtype=np.int64)
        blk = np.empty(max(nstored, 1))
        vec = np.zeros(nvar)
        cnt = np.zeros(nvar, np.int64)
        S = _PassStats()
        rc = self.lib.orc_loopn_blocks(C.byref(R), C.byref(O), nthreads, len(gfirst) - 1, _p(gfirst, I32), _p(brow, I64),
                                       _p(bcol, I32), _p(boff, I64), nstored, _p(blk, D), _p(vec, D), _p(cnt, I64),
                                       C.byref(S))
        if rc != 0:
            raise RuntimeError(f"oracle loopn_blocks failed with code {rc}")
        stats = dict(chi2=S.chi2, ndf=S.ndf, nrej=list(S.nrej), naccept=S.naccept, nband=S.nbandrec)
        return blk, vec, cnt, stats

    def read_mille(self, path, fortran=False):
        rp, ii, gg = I64(), I32(), D()
        n = self.lib.orc_read_mille(path.encode(), 1 if fortran else 0, C.byref(rp), C.byref(ii), C.byref(gg))
        if n < 0:
            raise IOError(f"oracle read_mille failed ({n})")
        rec_ptr = np.ctypeslib.as_array(rp, shape=(n + 1,)).copy()
        nw = int(rec_ptr[-1])
        inder = np.ctypeslib.as_array(ii, shape=(max(nw, 1),))[:nw].copy()
        glder = np.ctypeslib.as_array(gg, shape=(max(nw, 1),))[:nw].copy()
        for p in (rp, ii, gg):
            self.lib.orc_free(C.cast(p, C.c_void_p))
        return rec_ptr, inder, glder


class Reference:
    """The reference's Fortran numeric routines (src/mpnum.f90, src/Dbandmatrix.f90,
    src/minresModule.f90) built with flang; all arguments by reference."""

    def __init__(self, fresh=False):
        """fresh: load a private copy of the library, so that module state (the QL decomposition
        of src/mpqldec.f90 allocates its arrays once per process) starts unallocated."""
        if not os.path.exists(REF):
            raise FileNotFoundError(REF)
        path = REF
        if fresh:
            import shutil
            import tempfile
            fd, path = tempfile.mkstemp(suffix=".so", prefix="pedenum_")
            os.close(fd)
            shutil.copyfile(REF, path)
        self.lib = C.CDLL(path)
        if fresh:
            os.unlink(path)

    @staticmethod
    def _i(v):
        return C.byref(C.c_int32(v))

    def sqminv(self, v, b):
        v = np.asfortranarray(v, dtype=np.float64).copy()
        b = np.asfortranarray(b, dtype=np.float64).copy()
        n = len(b)
        nr = C.c_int32()
        diag = np.zeros(n)
        nxt = np.zeros(n, dtype=np.int32)
        self.lib.sqminv_(_p(v, D), _p(b, D), self._i(n), C.byref(nr), _p(diag, D), _p(nxt, I32))
        return v, b, nr.value

    def sqminl(self, v, b):
        v = np.asfortranarray(v, dtype=np.float64).copy()
        b = np.asfortranarray(b, dtype=np.float64).copy()
        n = len(b)
        nr = C.c_int32()
        diag, vk = np.zeros(n), np.zeros(n)
        nxt = np.zeros(n, dtype=np.int32)
        self.lib.sqminl_(_p(v, D), _p(b, D), self._i(n), C.byref(nr), _p(diag, D), _p(nxt, I32), _p(vk, D),
                         self._i(0))
        return v, b, nr.value

    def chdec2(self, g, n):
        g = np.asfortranarray(g, dtype=np.float64).copy()
        nr = C.c_int32()
        e1, e2 = C.c_double(), C.c_double()
        self.lib.chdec2_(_p(g, D), self._i(n), C.byref(nr), C.byref(e1), C.byref(e2), self._i(0))
        return g, nr.value, e1.value, e2.value

    def chslv2(self, g, x):
        x = np.asfortranarray(x, dtype=np.float64).copy()
        g = np.asfortranarray(g, dtype=np.float64)
        self.lib.chslv2_(_p(g, D), _p(x, D), self._i(len(x)))
        return x

    def dbavat(self, v, a, w, n, m):
        w = np.asfortranarray(w, dtype=np.float64).copy()
        self.lib.dbavat_(_p(np.asfortranarray(v, dtype=np.float64), D), _p(np.asfortranarray(a, dtype=np.float64), D),
                         _p(w, D), self._i(n), self._i(m), self._i(1))
        return w

    def sqmibb(self, v, b, nbdr, nbnd, inv, side=1):
        v = np.asfortranarray(v, dtype=np.float64).copy()
        b = np.asfortranarray(b, dtype=np.float64).copy()
        n = len(b)
        nr = C.c_int32()
        e1, e2 = C.c_double(), C.c_double()
        vbnd = np.zeros(n * (nbnd + 1))
        vbdr = np.zeros(max(1, n * nbdr))
        aux = np.zeros(max(1, n * nbdr) + n)
        vbk = np.zeros(max(1, (nbdr * nbdr + nbdr) // 2))
        vzru = np.zeros(max(1, nbdr))
        scd = np.zeros(max(1, nbdr))
        scf = np.zeros(max(1, nbdr), np.int32)
        f = self.lib.sqmibb_ if side == 1 else self.lib.sqmibb2_
        f(_p(v, D), _p(b, D), self._i(n), self._i(nbdr), self._i(nbnd), self._i(inv), C.byref(nr), _p(vbnd, D),
          _p(vbdr, D), _p(aux, D), _p(vbk, D), _p(vzru, D), _p(scd, D), _p(scf, I32), C.byref(e1), C.byref(e2))
        return v, b, nr.value, e1.value, e2.value

    def chindl(self, n, nd):
        f = self.lib.chindl_
        f.restype = C.c_float
        return f(self._i(n), self._i(nd))

    def lltdec(self, n, c, india, iopt=0):
        c = np.asfortranarray(c, dtype=np.float64).copy()
        india = np.asfortranarray(india, dtype=np.int32)
        nr = C.c_int32()
        self.lib.lltdec_(self._i(n), _p(c, D), _p(india, I32), C.byref(nr), self._i(iopt))
        return c, nr.value

    def equdecs(self, n, m, b, nm, ls, c, india, l):
        c = np.asfortranarray(c, dtype=np.float64).copy()
        india = np.asfortranarray(india, dtype=np.int32).copy()
        l = np.asfortranarray(l, dtype=np.int32)
        nr, nr2 = C.c_int32(), C.c_int32()
        self.lib.equdecs_(self._i(n), self._i(m), self._i(b), C.byref(C.c_int64(nm)), self._i(ls), _p(c, D),
                          _p(india, I32), _p(l, I32), C.byref(nr), C.byref(nr2))
        return c, india, nr.value, nr2.value

    def equslvs(self, n, m, b, nm, c, india, l, x):
        x = np.asfortranarray(x, dtype=np.float64).copy()
        self.lib.equslvs_(self._i(n), self._i(m), self._i(b), C.byref(C.c_int64(nm)),
                          _p(np.asfortranarray(c, dtype=np.float64), D), _p(np.asfortranarray(india, dtype=np.int32), I32),
                          _p(np.asfortranarray(l, dtype=np.int32), I32), _p(x, D))
        return x

    def precons(self, p, n, b, nm, c, a, l):
        cu = np.zeros(n)
        a = np.asfortranarray(a, dtype=np.float64).copy()
        s = np.zeros(max(1, (p * p + p) // 2))
        nr = C.c_int32()
        c = np.asfortranarray(c, dtype=np.float64)
        self.lib.precons_(self._i(p), self._i(n), self._i(b), C.byref(C.c_int64(nm)), _p(c, D), _p(cu, D), _p(a, D),
                          _p(np.asfortranarray(l, dtype=np.int32), I32), _p(s, D), C.byref(nr))
        return cu, a, s, nr.value

    def presols(self, p, n, b, nm, cu, a, l, s, y):
        x = np.zeros(n + p)
        y = np.asfortranarray(y, dtype=np.float64)
        self.lib.presols_(self._i(p), self._i(n), self._i(b), C.byref(C.c_int64(nm)), _p(cu, D), _p(a, D),
                          _p(np.asfortranarray(l, dtype=np.int32), I32), _p(s, D), _p(x, D), _p(y, D))
        return x

    # QL decomposition of constraints (src/mpqldec.f90); module state lives in the library
    def ql_decompose(self, npar, lay):
        """qlini + qldecb + qlgete for one parameter block; returns (emin, emax)."""
        self.lib.qlini_(self._i(npar), self._i(lay["ncon"]), self._i(1), C.byref(C.c_int64(lay["matsize"])),
                        self._i(0))
        self._ql_keep = [np.asfortranarray(lay["a"], dtype=np.float64), np.asfortranarray(lay["bpar"]),
                         np.asfortranarray(lay["bcon"]), np.asfortranarray(lay["rcon"])]
        a, bpar, bcon, rcon = self._ql_keep
        self.lib.qldecb_(_p(a, D), _p(bpar, I32), _p(bcon, I32), _p(rcon, I32))
        e1, e2 = C.c_double(), C.c_double()
        self.lib.qlgete_(C.byref(e1), C.byref(e2))
        self._ql_npar = npar
        return e1.value, e2.value

    def qlmlq(self, x, t):
        x = np.array(x, dtype=np.float64, order="F", copy=True)
        m = 1 if x.ndim == 1 else x.shape[1]
        self.lib.qlmlq_(_p(x, D), self._i(m), self._i(1 if t else 0))
        return x

    def qlbsub(self, d, ncon):
        d = np.asfortranarray(d, dtype=np.float64)
        y = np.zeros(ncon)
        self.lib.qlbsub_(_p(d, D), _p(y, D))
        return y

    def qlpssq(self, A, band, mb, t=True):
        """band part (pede's variable-band layout, band width mb incl. the diagonal) of
        Q^T A Q (t True) for a dense symmetric A, through pede's qlpssq with A*v products."""
        A = np.asarray(A, dtype=np.float64)
        n = A.shape[0]
        FAP = C.CFUNCTYPE(None, I32, I64, D, I32, I32, D)

        def _ap(np_, lp, xp, isp, iep, yp):
            nn = np_[0]
            xa = np.ctypeslib.as_array(xp, shape=(nn,))
            ya = np.ctypeslib.as_array(yp, shape=(nn,))
            i0, i1 = isp[0] - 1, iep[0]
            ya[:] += A[:, i0:i1] @ xa[i0:i1]

        fa = FAP(_ap)
        b = np.array(band, dtype=np.float64, order="F", copy=True)
        self.lib.qlpssq_(fa, _p(b, D), self._i(mb), self._i(1 if t else 0))
        return b

    def minres(self, n, aprod, msolve, b, rtol, itnlim=2000, precon=True):
        FOP = C.CFUNCTYPE(None, I32, D, D)

        def _ap(np_, xp, yp):
            nn = np_[0]
            xa = np.ctypeslib.as_array(xp, shape=(nn,))
            ya = np.ctypeslib.as_array(yp, shape=(nn,))
            ya[:] = aprod(xa)

        def _ms(np_, xp, yp):
            nn = np_[0]
            xa = np.ctypeslib.as_array(xp, shape=(nn,))
            ya = np.ctypeslib.as_array(yp, shape=(nn,))
            ya[:] = msolve(xa) if msolve else xa

        fa, fm = FOP(_ap), FOP(_ms)
        b = np.asfortranarray(b, dtype=np.float64)
        x = np.zeros(n)
        istop, itn = C.c_int32(), C.c_int32()
        out = [C.c_double() for _ in range(5)]
        f = self.lib._QMminresmodulePminres
        f(self._i(n), fa, fm, _p(b, D), C.byref(C.c_double(0.0)), C.byref(C.c_int32(0)),
          C.byref(C.c_int32(1 if precon else 0)), _p(x, D), self._i(itnlim), self._i(0), C.byref(C.c_double(rtol)),
          C.byref(istop), C.byref(itn), *[C.byref(o) for o in out])
        return x, istop.value, itn.value, np.array([o.value for o in out])

    def minresqlp(self, n, aprod, msolve, b, rtol, itnlim=2000, precon=True, maxxnorm=None, trancond=1.0e7,
                  acondlim=None, shift=None):
        """The reference MINRESQLP (src/minresqlpModule.f90:411); absent optional arguments are
        passed as null pointers, nout = 0 (a preconnected unit, so nothing is printed)."""
        FOP = C.CFUNCTYPE(None, I32, D, D)

        def _ap(np_, xp, yp):
            nn = np_[0]
            xa = np.ctypeslib.as_array(xp, shape=(nn,))
            ya = np.ctypeslib.as_array(yp, shape=(nn,))
            ya[:] = aprod(xa)

        def _ms(np_, xp, yp):
            nn = np_[0]
            xa = np.ctypeslib.as_array(xp, shape=(nn,))
            ya = np.ctypeslib.as_array(yp, shape=(nn,))
            ya[:] = msolve(xa)

        if maxxnorm is None:
            maxxnorm = n / np.sqrt(np.finfo(float).eps)
        fa = FOP(_ap)
        fm = FOP(_ms) if precon else C.cast(None, FOP)
        b = np.asfortranarray(b, dtype=np.float64)
        x = np.zeros(n)
        istop, itn = C.c_int32(), C.c_int32()
        out = [C.c_double() for _ in range(5)]
        opt = lambda v: None if v is None else C.byref(C.c_double(v))  # noqa: E731
        f = self.lib._QMminresqlpmodulePminresqlp
        f.restype = None
        f(self._i(n), fa, _p(b, D), opt(shift), fm, None, self._i(0), self._i(itnlim), C.byref(C.c_double(rtol)),
          C.byref(C.c_double(maxxnorm)), C.byref(C.c_double(trancond)), opt(acondlim), _p(x, D), C.byref(istop),
          C.byref(itn), *[C.byref(o) for o in out])
        return x, istop.value, itn.value, np.array([o.value for o in out])
