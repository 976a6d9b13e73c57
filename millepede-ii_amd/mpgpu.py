"""Python host mirror of the mpgpu C-ABI (include/mpgpu.h) and of the synthetic workload
generators (libmpgen).

The names follow the pede routines each call replaces: loop1/loop2 (structure),
loopn (data pass), mupdat, avprod and the metsol dispatch (solve).  Errors raise
MpgpuError carrying pede's exit code (doc/exitCodes.md).  The library is loaded from this
directory; when it is missing the import fails loudly -- there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpgpu.so")
GEN_PATH = os.path.join(HERE, "libmpgen.so")

D = C.POINTER(C.c_double)
F = C.POINTER(C.c_float)
I32 = C.POINTER(C.c_int32)
I64 = C.POINTER(C.c_int64)

INVERSION, DECOMPOSITION, MINRES, MINRES_QLP, LAPACK_PACKED, LAPACK_UNPACKED, CHOLESKY_FP32_IR = 1, 3, 4, 5, 7, 8, 11
FULL, SPARSE = 1, 2

# every entry point declared in include/mpgpu.h
EXPORTS = [
    "mpgpu_default_options", "mpgpu_create", "mpgpu_destroy", "mpgpu_last_error", "mpgpu_set_options",
    "mpgpu_add_records", "mpgpu_read_mille_file", "mpgpu_set_steering_parameters", "mpgpu_add_equations",
    "mpgpu_build_structure", "mpgpu_get_labels", "mpgpu_matrix_info", "mpgpu_get_block_structure", "mpgpu_get_tuning",
    "mpgpu_get_record_summary", "mpgpu_set_parameters",
    "mpgpu_get_parameters", "mpgpu_loopn", "mpgpu_get_vector", "mpgpu_get_matrix_packed", "mpgpu_mupdat",
    "mpgpu_solve", "mpgpu_avprod", "mpgpu_avprod_noalloc", "mpgpu_get_errors", "mpgpu_get_monitor", "mpgpu_run_steering",
    "mpgpu_server_create", "mpgpu_server_submit", "mpgpu_server_wait", "mpgpu_server_destroy",
]


class Options(C.Structure):
    _fields_ = [("metsol", C.c_int32), ("matsto", C.c_int32), ("mbandw", C.c_int32), ("lprecm", C.c_int32),
                ("lhuber", C.c_int32), ("lfitbb", C.c_int32), ("mreqenf", C.c_int32), ("mreqena", C.c_int32),
                ("mcount", C.c_int32), ("itnlim", C.c_int32), ("withmultipliers", C.c_int32),
                ("mrmode", C.c_int32), ("chicut", C.c_float), ("chirem", C.c_float), ("chhuge", C.c_float),
                ("dwcut", C.c_float), ("cndlmx", C.c_float), ("regula", C.c_float),
                ("default_presigma", C.c_double), ("mrestl", C.c_double), ("mrtcnd", C.c_double),
                ("nregul", C.c_int32), ("refine_steps", C.c_int32), ("monitor", C.c_int32), ("monitor_bins", C.c_int32),
                ("lapack_errors", C.c_int32), ("pair_entries", C.c_int32), ("skip_empty_cons", C.c_int32),
                ("resolve_redundancy_cons", C.c_int32), ("scale_errors", C.c_double * 2),
                ("refine_tol", C.c_double), ("fp32_min_n", C.c_int32), ("iteren", C.c_int32)]


class PassStats(C.Structure):
    _fields_ = [("chi2", C.c_double), ("ndf", C.c_int64), ("wndf", C.c_double), ("nrej", C.c_int64 * 6),
                ("naccept", C.c_int64), ("nrec", C.c_int64), ("ms_kernel", C.c_double), ("ms_total", C.c_double),
                ("ms_records", C.c_double), ("n_records", C.c_int64), ("ms_gather", C.c_double),
                ("n_gather", C.c_int64)]


class SolveStats(C.Structure):
    _fields_ = [("method", C.c_int32), ("nrank", C.c_int32), ("istop", C.c_int32), ("itn", C.c_int32),
                ("anorm", C.c_double), ("acond", C.c_double), ("rnorm", C.c_double), ("arnorm", C.c_double),
                ("ynorm", C.c_double), ("rel_residual", C.c_double), ("ms", C.c_double), ("evmax", C.c_double),
                ("evmin", C.c_double)]


class RecordSummary(C.Structure):
    _fields_ = [("nrec", C.c_int64), ("neq", C.c_int64), ("neq_glob", C.c_int64), ("nglob_deriv", C.c_int64),
                ("fp64_payload", C.c_int32), ("ntot", C.c_int32), ("nvar", C.c_int32), ("ngroup_tot", C.c_int32),
                ("ngroup_var", C.c_int32), ("ncon", C.c_int32), ("max_glob_rec", C.c_int32),
                ("max_loc_rec", C.c_int32), ("max_eq_rec", C.c_int32), ("nempty_rec", C.c_int64)]


class RunStats(C.Structure):
    _fields_ = [("ntot", C.c_int32), ("nvar", C.c_int32), ("ncon", C.c_int32), ("iterations", C.c_int32),
                ("nrec", C.c_int64), ("final_chi2", C.c_double), ("final_ndf", C.c_int64),
                ("ms_structure", C.c_double), ("ms_passes", C.c_double), ("ms_solves", C.c_double),
                ("exit_code", C.c_int32)]


class JobStats(C.Structure):
    _fields_ = [("status", C.c_int32), ("nvar", C.c_int32), ("latency_ms", C.c_double), ("queue_ms", C.c_double),
                ("run_ms", C.c_double), ("chi2", C.c_double), ("ndf", C.c_int64), ("istop", C.c_int32),
                ("itn", C.c_int32), ("rel_residual", C.c_double), ("t_arrival_s", C.c_double), ("t_done_s", C.c_double),
                ("method", C.c_int32), ("reserved", C.c_int32)]


class MpgpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[pede exit code {code}] {msg}")
        self.code = code


_lib = None


def lib():
    """Load libmpgpu.so (raises OSError when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: run millepede-ii_amd/build.py")
        L = C.CDLL(LIB_PATH)
        L.mpgpu_default_options.argtypes = [C.POINTER(Options)]
        L.mpgpu_create.argtypes = [C.POINTER(Options), C.c_int32, C.POINTER(C.c_void_p)]
        L.mpgpu_destroy.argtypes = [C.c_void_p]
        L.mpgpu_last_error.argtypes = [C.c_void_p]
        L.mpgpu_last_error.restype = C.c_char_p
        L.mpgpu_set_options.argtypes = [C.c_void_p, C.POINTER(Options)]
        L.mpgpu_add_records.argtypes = [C.c_void_p, C.c_int64, I64, I32, F, D]
        L.mpgpu_read_mille_file.argtypes = [C.c_void_p, C.c_char_p, C.c_int32, I64]
        L.mpgpu_set_steering_parameters.argtypes = [C.c_void_p, C.c_int32, I32, D, D]
        L.mpgpu_add_equations.argtypes = [C.c_void_p, C.c_int32, C.c_int32, I64, I32, D, D, D]
        L.mpgpu_build_structure.argtypes = [C.c_void_p, I32, I32, I32]
        L.mpgpu_get_labels.argtypes = [C.c_void_p, I32, I32, I64, I32]
        L.mpgpu_matrix_info.argtypes = [C.c_void_p, I64, I64, I32]
        L.mpgpu_get_block_structure.argtypes = [C.c_void_p, I32, I64, I32, I64]
        L.mpgpu_get_tuning.argtypes = [C.c_void_p, C.c_char_p, C.c_int32]
        L.mpgpu_get_record_summary.argtypes = [C.c_void_p, C.POINTER(RecordSummary)]
        L.mpgpu_set_parameters.argtypes = [C.c_void_p, D]
        L.mpgpu_get_parameters.argtypes = [C.c_void_p, D]
        L.mpgpu_loopn.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_float, C.POINTER(PassStats)]
        L.mpgpu_get_vector.argtypes = [C.c_void_p, D, I64]
        L.mpgpu_get_matrix_packed.argtypes = [C.c_void_p, D, C.c_int64]
        L.mpgpu_mupdat.argtypes = [C.c_void_p, C.c_int32, I32, I32, D]
        L.mpgpu_solve.argtypes = [C.c_void_p, D, D, C.POINTER(SolveStats)]
        L.mpgpu_avprod.argtypes = [C.c_void_p, D, D]
        L.mpgpu_avprod_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.mpgpu_get_errors.argtypes = [C.c_void_p, D, D]
        L.mpgpu_get_monitor.argtypes = [C.c_void_p, I32, I32, C.POINTER(C.c_float), I64]
        L.mpgpu_run_steering.argtypes = [C.c_char_p, C.c_char_p, C.c_int32, C.POINTER(RunStats)]
        L.mpgpu_server_create.argtypes = [C.c_int32, C.c_int32, C.c_int64, C.c_int32, C.POINTER(C.c_void_p)]
        L.mpgpu_server_submit.argtypes = [C.c_void_p, C.POINTER(Options), C.c_int64, I64, I32, F, C.c_int32, I32, D, D,
                                          C.c_int32, I64, I32, D, D, I64]
        L.mpgpu_server_wait.argtypes = [C.c_void_p, C.c_int64, C.c_int32, I32, D, C.POINTER(JobStats)]
        L.mpgpu_server_destroy.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def default_options(**kw):
    o = Options()
    lib().mpgpu_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Pede:
    """One mpgpu context: records, structure, passes and solves on one GPU."""

    def __init__(self, device=0, **options):
        self.L = lib()
        self.opt = default_options(**options)
        h = C.c_void_p()
        rc = self.L.mpgpu_create(C.byref(self.opt), device, C.byref(h))
        if rc != 0:
            raise MpgpuError(rc, "mpgpu_create failed (no HIP device?)")
        self.h = h
        self.ntot = self.nvar = self.ncon = 0

    def close(self):
        if self.h:
            self.L.mpgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            msg = self.L.mpgpu_last_error(self.h)
            raise MpgpuError(rc, msg.decode() if msg else "")

    # ---- input ----
    def add_records(self, rec_ptr, inder, glder):
        rec_ptr = np.ascontiguousarray(rec_ptr, dtype=np.int64)
        inder = np.ascontiguousarray(inder, dtype=np.int32)
        if glder.dtype == np.float32:
            g = np.ascontiguousarray(glder)
            self._chk(self.L.mpgpu_add_records(self.h, len(rec_ptr) - 1, _p(rec_ptr, I64), _p(inder, I32), _p(g, F), None))
        else:
            g = np.ascontiguousarray(glder, dtype=np.float64)
            self._chk(self.L.mpgpu_add_records(self.h, len(rec_ptr) - 1, _p(rec_ptr, I64), _p(inder, I32), None, _p(g, D)))

    def read_mille(self, path, fortran=False):
        n = C.c_int64()
        self._chk(self.L.mpgpu_read_mille_file(self.h, path.encode(), 1 if fortran else 0, C.byref(n)))
        return n.value

    def set_steering_parameters(self, labels, values=None, presigma=None):
        labels = np.ascontiguousarray(labels, dtype=np.int32)
        v = None if values is None else np.ascontiguousarray(values, dtype=np.float64)
        s = None if presigma is None else np.ascontiguousarray(presigma, dtype=np.float64)
        self._chk(self.L.mpgpu_set_steering_parameters(self.h, len(labels), _p(labels, I32), _p(v, D), _p(s, D)))

    def add_equations(self, kind, blocks):
        """blocks: list of (rhs, sigma, labels, factors); kind 0 constraint, 1 measurement"""
        ptr = [0]
        labs, facs, rhs, sig = [], [], [], []
        for (r, s, l, f) in blocks:
            labs += list(l)
            facs += list(f)
            ptr.append(len(labs))
            rhs.append(r)
            sig.append(s)
        ptr = np.array(ptr, dtype=np.int64)
        labs = np.array(labs, dtype=np.int32)
        facs = np.array(facs, dtype=np.float64)
        rhs = np.array(rhs, dtype=np.float64)
        sig = np.array(sig, dtype=np.float64)
        self._chk(self.L.mpgpu_add_equations(self.h, kind, len(blocks), _p(ptr, I64), _p(labs, I32), _p(facs, D),
                                             _p(rhs, D), _p(sig, D)))

    def add_constraints(self, blocks):
        self.add_equations(0, blocks)

    # ---- structure ----
    def build_structure(self):
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        self._chk(self.L.mpgpu_build_structure(self.h, C.byref(a), C.byref(b), C.byref(c)))
        self.ntot, self.nvar, self.ncon = a.value, b.value, c.value
        self.nall = self.nvar + self.ncon
        return self.ntot, self.nvar, self.ncon

    def labels(self):
        n = self.ntot
        lab = np.zeros(n, np.int32)
        var = np.zeros(n, np.int32)
        cnt = np.zeros(n, np.int64)
        grp = np.zeros(n, np.int32)
        self._chk(self.L.mpgpu_get_labels(self.h, _p(lab, I32), _p(var, I32), _p(cnt, I64), _p(grp, I32)))
        return lab, var, cnt, grp

    def matrix_info(self):
        a, b, c = C.c_int64(), C.c_int64(), C.c_int32()
        self._chk(self.L.mpgpu_matrix_info(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return dict(stored=a.value, nblocks=b.value, ngroups=c.value)

    def tuning(self):
        """the autotuner's measurements and choice for this structure (text)"""
        buf = C.create_string_buffer(512)
        self._chk(self.L.mpgpu_get_tuning(self.h, buf, 512))
        return buf.value.decode()

    def record_summary(self):
        """the counts of pede's loop1 / loop2 (millepede.log) for the built structure"""
        r = RecordSummary()
        self._chk(self.L.mpgpu_get_record_summary(self.h, C.byref(r)))
        return {k: getattr(r, k) for k, _ in RecordSummary._fields_}

    def block_structure(self):
        """(gfirst, brow, bcol, boff, nstored) of the group-block storage (sparse storage)"""
        info = self.matrix_info()
        ng, nb = info["ngroups"], info["nblocks"]
        gfirst = np.zeros(ng + 1, np.int32)
        brow = np.zeros(ng + 1, np.int64)
        bcol = np.zeros(max(nb, 1), np.int32)
        boff = np.zeros(max(nb, 1), np.int64)
        self._chk(self.L.mpgpu_get_block_structure(self.h, _p(gfirst, I32), _p(brow, I64), _p(bcol, I32), _p(boff, I64)))
        return gfirst, brow, bcol[:nb], boff[:nb], info["stored"]

    def set_options(self, **kw):
        """change options of this context (those the structure build used are fixed, code 30)"""
        for k, v in kw.items():
            setattr(self.opt, k, v)
        self._chk(self.L.mpgpu_set_options(self.h, C.byref(self.opt)))

    def set_parameters(self, par):
        p = np.ascontiguousarray(par, dtype=np.float64)
        assert len(p) == self.ntot
        self._chk(self.L.mpgpu_set_parameters(self.h, _p(p, D)))

    def parameters(self):
        p = np.zeros(self.ntot)
        self._chk(self.L.mpgpu_get_parameters(self.h, _p(p, D)))
        return p

    # ---- passes ----
    def loopn(self, icalcm=1, nloopn=1, chicut=0.0):
        st = PassStats()
        self._chk(self.L.mpgpu_loopn(self.h, icalcm, nloopn, chicut, C.byref(st)))
        return dict(chi2=st.chi2, ndf=st.ndf, nrej=list(st.nrej), naccept=st.naccept, nrec=st.nrec,
                    ms_kernel=st.ms_kernel, ms_total=st.ms_total,
                    ms_records=st.ms_records, n_records=st.n_records, ms_gather=st.ms_gather, n_gather=st.n_gather)

    def vector(self):
        v = np.zeros(self.nall)
        c = np.zeros(self.nvar, np.int64)
        self._chk(self.L.mpgpu_get_vector(self.h, _p(v, D), _p(c, I64)))
        return v, c

    def matrix_packed(self):
        n = self.nall
        m = np.zeros(n * (n + 1) // 2)
        self._chk(self.L.mpgpu_get_matrix_packed(self.h, _p(m, D), len(m)))
        return m

    def mupdat(self, i, j, v):
        i = np.ascontiguousarray(i, dtype=np.int32)
        j = np.ascontiguousarray(j, dtype=np.int32)
        v = np.ascontiguousarray(v, dtype=np.float64)
        self._chk(self.L.mpgpu_mupdat(self.h, len(i), _p(i, I32), _p(j, I32), _p(v, D)))

    # ---- solve ----
    def solve(self, rhs=None):
        x = np.zeros(self.nall)
        st = SolveStats()
        r = None if rhs is None else np.ascontiguousarray(rhs, dtype=np.float64)
        self._chk(self.L.mpgpu_solve(self.h, _p(r, D), _p(x, D), C.byref(st)))
        return x, dict(method=st.method, nrank=st.nrank, istop=st.istop, itn=st.itn, anorm=st.anorm,
                       acond=st.acond, rnorm=st.rnorm, rel_residual=st.rel_residual, ms=st.ms, evmax=st.evmax,
                       evmin=st.evmin)

    def avprod(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.nall)
        self._chk(self.L.mpgpu_avprod(self.h, _p(x, D), _p(y, D)))
        return y

    def errors(self):
        e = np.zeros(self.nvar)
        g = np.zeros(self.nvar)
        self._chk(self.L.mpgpu_get_errors(self.h, _p(e, D), _p(g, D)))
        return e, g

    def monitor(self):
        """residual-monitoring histograms of the last pass: (labels, mean errors, counts[ngroups, bins])"""
        ng = C.c_int32()
        self._chk(self.L.mpgpu_get_monitor(self.h, C.byref(ng), None, None, None))
        lab = np.zeros(ng.value, np.int32)
        err = np.zeros(ng.value, np.float32)
        bins = self.opt.monitor_bins if self.opt.monitor_bins > 0 else 100
        h = np.zeros(ng.value * bins, np.int64)
        self._chk(self.L.mpgpu_get_monitor(self.h, C.byref(ng), _p(lab, I32), err.ctypes.data_as(C.POINTER(C.c_float)),
                                           _p(h, I64)))
        return lab, err, h.reshape(ng.value, bins)


class Server:
    """Serving engine: jobs admitted by device-memory footprint, run concurrently (one context
    and stream per worker) on one GPU."""

    def __init__(self, device=0, nworkers=8, budget_bytes=200 << 30, batch_max_params=5000):
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.mpgpu_server_create(device, nworkers, budget_bytes, batch_max_params, C.byref(h))
        if rc:
            raise MpgpuError(rc, "server create failed")
        self.h = h
        self._keep = {}

    def submit(self, records, params=None, constraints=None, **options):
        rp, ii, gg = records
        rp = np.ascontiguousarray(rp, dtype=np.int64)
        ii = np.ascontiguousarray(ii, dtype=np.int32)
        gg = np.ascontiguousarray(gg, dtype=np.float32)
        o = default_options(**options)
        pl = pv = ps = None
        npar = 0
        if params is not None:
            pl = np.ascontiguousarray(params[0], dtype=np.int32)
            pv = np.ascontiguousarray(params[1], dtype=np.float64)
            ps = np.ascontiguousarray(params[2], dtype=np.float64)
            npar = len(pl)
        cp = cl = cf = cr = None
        ncon = 0
        if constraints:
            ptr, labs, facs, rhs = [0], [], [], []
            for (r, s, l, f) in constraints:
                labs += list(l)
                facs += list(f)
                ptr.append(len(labs))
                rhs.append(r)
            cp = np.array(ptr, np.int64)
            cl = np.array(labs, np.int32)
            cf = np.array(facs, np.float64)
            cr = np.array(rhs, np.float64)
            ncon = len(rhs)
        jid = C.c_int64()
        rc = self.L.mpgpu_server_submit(self.h, C.byref(o), len(rp) - 1, _p(rp, I64), _p(ii, I32), _p(gg, F), npar,
                                        _p(pl, I32), _p(pv, D), _p(ps, D), ncon, _p(cp, I64), _p(cl, I32), _p(cf, D),
                                        _p(cr, D), C.byref(jid))
        if rc:
            raise MpgpuError(rc, "submit failed")
        # the server reads these arrays while the job runs: keep them until it is waited for
        self._keep[jid.value] = (rp, ii, gg, pl, pv, ps, cp, cl, cf, cr)
        return jid.value

    def wait(self, job_id, capacity=0):
        st = JobStats()
        lab = np.zeros(max(capacity, 1), np.int32)
        cor = np.zeros(max(capacity, 1))
        rc = self.L.mpgpu_server_wait(self.h, job_id, capacity, _p(lab, I32) if capacity else None,
                                      _p(cor, D) if capacity else None, C.byref(st))
        self._keep.pop(job_id, None)
        out = {f: getattr(st, f) for f, _ in JobStats._fields_}
        n = min(capacity, st.nvar)
        return rc, out, lab[:n], cor[:n]

    def close(self):
        if self.h:
            self.L.mpgpu_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_steering(steering_file, output_dir=".", device=0):
    st = RunStats()
    rc = lib().mpgpu_run_steering(steering_file.encode(), output_dir.encode(), device, C.byref(st))
    out = {f: getattr(st, f) for f, _ in RunStats._fields_}
    if rc != 0:
        raise MpgpuError(rc, f"run_steering failed: {out}")
    return out


# ---------------------------------------------------------------------------------------------
# synthetic workloads (libmpgen)
# ---------------------------------------------------------------------------------------------
class _GenRecords(C.Structure):
    _fields_ = [("nrec", C.c_int64), ("nwords", C.c_int64), ("rec_ptr", I64), ("inder", I32), ("glder", F)]


_gen = None


def gen():
    global _gen
    if _gen is None:
        if not os.path.exists(GEN_PATH):
            raise OSError(f"{GEN_PATH} is missing: run millepede-ii_amd/build.py")
        G = C.CDLL(GEN_PATH)
        G.mpgen_toy.argtypes = [C.c_uint64, C.c_int64, C.POINTER(_GenRecords), D]
        G.mpgen_pixel.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                  C.c_double, C.c_int64, C.c_int, C.c_int, C.c_double, C.c_double,
                                  C.POINTER(_GenRecords), D]
        G.mpgen_brokenlines.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_int64, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                        C.POINTER(_GenRecords)]
        G.mpgen_dense.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int64, C.c_double, C.c_double,
                                  C.POINTER(_GenRecords), D]
        G.mpgen_write.argtypes = [C.c_char_p, C.POINTER(_GenRecords), C.c_int]
        G.mpgen_free.argtypes = [C.POINTER(_GenRecords)]
        _gen = G
    return _gen


def _take(r):
    n, nw = r.nrec, r.nwords
    rec_ptr = np.ctypeslib.as_array(r.rec_ptr, shape=(n + 1,)).copy()
    inder = np.ctypeslib.as_array(r.inder, shape=(max(nw, 1),))[:nw].copy()
    glder = np.ctypeslib.as_array(r.glder, shape=(max(nw, 1),))[:nw].copy()
    gen().mpgen_free(C.byref(r))
    return rec_ptr, inder, glder


def toy_records(seed, ntracks):
    """Planar toy detector: 10 planes x 5 modules, 2 globals each (100 labels)."""
    r = _GenRecords()
    truth = np.zeros(100)
    rc = gen().mpgen_toy(seed, ntracks, C.byref(r), _p(truth, D))
    if rc:
        raise MpgpuError(rc, "generator failed")
    return _take(r) + (truth,)


def toy_labels():
    return np.array([100 * p + 10 * m + k for p in range(1, 11) for m in range(5) for k in (1, 2)], dtype=np.int32)


def pixel_records(seed, nlayers, nx, ny, ntracks, hmin, hmax, wx=2.0, wy=5.0, z0=10.0, dz=5.0, slope_sigma=0.1,
                  sigma=0.002, truth=True):
    """Layered pixel detector with 6 rigid-body globals per module."""
    r = _GenRecords()
    tr = np.zeros(nlayers * nx * ny * 6) if truth else None
    rc = gen().mpgen_pixel(seed, nlayers, nx, ny, wx, wy, z0, dz, ntracks, hmin, hmax, slope_sigma, sigma,
                           C.byref(r), _p(tr, D))
    if rc:
        raise MpgpuError(rc, "generator failed")
    return _take(r) + (tr,)


def brokenlines_records(seed, nlayers, nx, ny, ntracks, pmin, pmax, wx=2.0, wy=5.0, z0=10.0, dz=1.0,
                        slope_sigma=0.05, sigma=0.002, theta0=2e-4):
    """Broken-lines tracks: 4 border + 2 offsets per plane as local parameters (bordered-band
    local fits), 6 rigid-body globals per hit module."""
    r = _GenRecords()
    rc = gen().mpgen_brokenlines(seed, nlayers, nx, ny, wx, wy, z0, dz, ntracks, pmin, pmax, slope_sigma, sigma,
                                 theta0, C.byref(r))
    if rc:
        raise MpgpuError(rc, "generator failed")
    return _take(r)


def dense_records(seed, nlayers, nx, ny, ntracks, slope_frac=0.75, sigma=0.002, truth=True):
    """Dense detector: every track crosses all nlayers layers of a periodic nx x ny module grid
    (6 rigid-body globals per module, 4 local parameters, 2 measurements per hit); the slope
    spread sweeps slope_frac of the grid across the stack, so the normal matrix is densely filled."""
    r = _GenRecords()
    tr = np.zeros(nlayers * nx * ny * 6) if truth else None
    rc = gen().mpgen_dense(seed, nlayers, nx, ny, ntracks, slope_frac, sigma, C.byref(r), _p(tr, D))
    if rc:
        raise MpgpuError(rc, "generator failed")
    return _take(r) + (tr,)


def pixel_labels(nlayers, nx, ny):
    return np.array([(l + 1) * 100000 + (m + 1) * 10 + k + 1 for l in range(nlayers) for m in range(nx * ny)
                     for k in range(6)], dtype=np.int32)


def pixel_constraints(nlayers, nx, ny):
    """sum of shifts per layer = 0 (du, dv, dw): 3 constraints per layer"""
    blocks = []
    for l in range(nlayers):
        for k in range(3):
            labs = [(l + 1) * 100000 + (m + 1) * 10 + k + 1 for m in range(nx * ny)]
            blocks.append((0.0, 0.0, labs, [1.0] * len(labs)))
    return blocks


def pixel_presigma(labels):
    """Loose pre-sigmas for the weakly determined module parameters: 0.1 cm for dw,
    0.01 rad for the rotations."""
    return [0.1 if l % 10 == 3 else 1e-2 for l in labels]


def pixel_weak_labels(nlayers, nx, ny):
    return [(l + 1) * 100000 + (m + 1) * 10 + k for l in range(nlayers) for m in range(nx * ny) for k in (3, 4, 5, 6)]


def write_mille(path, rec_ptr, inder, glder, fortran=False):
    r = _GenRecords()
    rp = np.ascontiguousarray(rec_ptr, dtype=np.int64)
    ii = np.ascontiguousarray(inder, dtype=np.int32)
    gg = np.ascontiguousarray(glder, dtype=np.float32)
    r.nrec, r.nwords = len(rp) - 1, int(rp[-1])
    r.rec_ptr, r.inder, r.glder = _p(rp, I64), _p(ii, I32), _p(gg, F)
    rc = gen().mpgen_write(path.encode(), C.byref(r), 1 if fortran else 0)
    if rc:
        raise MpgpuError(rc, f"cannot write {path}")
