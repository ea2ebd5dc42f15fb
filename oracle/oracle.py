"""ctypes binding of the oracle library (test infrastructure only).

Loaded by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg, as the checker or
the timed CPU restatement; the product never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_GOLDEN = os.path.join(HERE, "_ref", "qw_golden")
REF_QWALK = os.path.join(HERE, "_ref", "qwalk")

_dp = ctypes.POINTER(ctypes.c_double)
_lib = None


def build():
    subprocess.run(["make", "-s", "-f", os.path.join(HERE, "Makefile")], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_info.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        L.orc_eval.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp,
                               _dp, _dp, _dp, _dp, _dp]
        L.orc_mo_eval.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_nparm.argtypes = [ctypes.c_void_p, _dp]
        L.orc_parm_deriv.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.orc_set_parms.argtypes = [ctypes.c_void_p, _dp]
        L.orc_basis_eval.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
        L.orc_spline.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
        L.orc_mo_coeff.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
        L.orc_transpose_inverse.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp]
        L.orc_inverse_update_row.restype = ctypes.c_double
        L.orc_inverse_update_row.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int]
        L.orc_philox.argtypes = [ctypes.POINTER(ctypes.c_uint)] * 3
        L.orc_tmoves.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int, _dp, ctypes.c_int,
                                 ctypes.POINTER(ctypes.c_int), _dp, _dp, _dp]
        L.orc_tmove_step.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int, _dp, ctypes.c_double,
                                     ctypes.c_int, ctypes.c_ulonglong, ctypes.c_uint, ctypes.c_uint]
        L.orc_run.restype = ctypes.c_double
        L.orc_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _dp, _dp, _dp, _dp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp) if a is not None else None


class RunOpts(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int), ("nwalkers", ctypes.c_int), ("nblock", ctypes.c_int),
                ("nstep", ctypes.c_int), ("ndecorr", ctypes.c_int), ("timestep", ctypes.c_double),
                ("seed", ctypes.c_ulonglong), ("pp_stochastic", ctypes.c_int),
                ("feedback_interval", ctypes.c_int), ("feedback", ctypes.c_double),
                ("max_weight", ctypes.c_double), ("branch_start_cutoff", ctypes.c_double),
                ("nthreads", ctypes.c_int), ("step0", ctypes.c_uint), ("tmoves", ctypes.c_int)]


class Oracle:
    def __init__(self, deck):
        err = ctypes.create_string_buffer(512)
        self.h = lib().orc_create(deck.encode(), err, 512)
        if not self.h:
            raise RuntimeError("oracle: " + err.value.decode())
        info = (ctypes.c_int * 5)()
        lib().orc_info(self.h, info)
        self.nup, self.ndn, self.natoms, self.nbasis, self.has_pp = list(info)
        self.ne = self.nup + self.ndn

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def eval(self, pos, rot=None, pp_mode=0, pp_rand=None):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        K = pos.shape[0]
        out = dict(logpsi=np.zeros(K), sign=np.zeros(K), grad=np.zeros((K, self.ne, 3)),
                   lap=np.zeros((K, self.ne)), eparts=np.zeros((K, 3)))
        r = None if rot is None else np.ascontiguousarray(rot, dtype=np.float64)
        pr = None if pp_rand is None else np.ascontiguousarray(pp_rand, dtype=np.float64)
        lib().orc_eval(self.h, K, _p(pos), _p(r), pp_mode, _p(pr), _p(out["logpsi"]), _p(out["sign"]),
                       _p(out["grad"]), _p(out["lap"]), _p(out["eparts"]))
        return out

    def parms(self):
        """Values of the linear Jastrow parameters (getParmDeriv order)."""
        n = lib().orc_nparm(self.h, None)
        v = np.zeros(max(n, 1))
        lib().orc_nparm(self.h, _p(v))
        return v[:n]

    def set_parms(self, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        lib().orc_set_parms(self.h, _p(v))

    def parm_deriv(self, pos):
        """gradient[K][np], gradderiv[K][np][Ne][4], val_gradient[K][Ne][3] of getParmDeriv."""
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        K = pos.shape[0]
        n = lib().orc_nparm(self.h, None)
        g = np.zeros((K, max(n, 1)))
        gd = np.zeros((K, max(n, 1), self.ne, 4))
        vg = np.zeros((K, self.ne, 3))
        lib().orc_parm_deriv(self.h, K, _p(pos), _p(g), _p(gd), _p(vg))
        return dict(gradient=g[:, :n], gradderiv=gd[:, :n], val_gradient=vg)

    def tmoves(self, pos, rot=None, pp_mode=0, pp_rand=None):
        """T-move candidates of one configuration pos[Ne][3]: electron[M], vxx[M],
        displacement[M][3] in the reference's order, and the nonlocal energy without them."""
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        r = None if rot is None else np.ascontiguousarray(rot, dtype=np.float64)
        pr = None if pp_rand is None else np.ascontiguousarray(pp_rand, dtype=np.float64)
        cap = max(1, self.natoms * self.ne * 32)
        e = (ctypes.c_int * cap)()
        vxx, disp, vnl = np.zeros(cap), np.zeros((cap, 3)), np.zeros(1)
        m = lib().orc_tmoves(self.h, _p(pos), _p(r), pp_mode, _p(pr), cap, e, _p(vxx), _p(disp), _p(vnl))
        assert m >= 0
        return dict(e=np.array(e[:m], dtype=np.int64), vxx=vxx[:m], disp=disp[:m], vnl=vnl[0])

    def tmove_step(self, pos, tau, sc, seed, walker, step, rot=None, pp_mode=0, pp_rand=None):
        """One T-move step of configuration pos[Ne][3]; returns (new positions, electrons moved)."""
        pos = np.ascontiguousarray(pos, dtype=np.float64).copy()
        r = None if rot is None else np.ascontiguousarray(rot, dtype=np.float64)
        pr = None if pp_rand is None else np.ascontiguousarray(pp_rand, dtype=np.float64)
        n = lib().orc_tmove_step(self.h, _p(pos), _p(r), pp_mode, _p(pr), tau, 1 if sc else 0, seed, walker, step)
        return pos, n

    def mo(self, spin, pts):
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        nmo = self.nup if spin == 0 else self.ndn
        out = np.zeros((pts.shape[0], 5, nmo))
        lib().orc_mo_eval(self.h, spin, pts.shape[0], _p(pts), _p(out))
        return out

    def basis(self, pts):
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        out = np.zeros((pts.shape[0], self.nbasis, 5))
        lib().orc_basis_eval(self.h, pts.shape[0], _p(pts), _p(out))
        return out

    def spline(self, shell):
        meta = np.zeros(3)
        n = lib().orc_spline(self.h, shell, None, _p(meta))
        if n < 0:
            return None
        c = np.zeros((n, 4))
        lib().orc_spline(self.h, shell, _p(c), _p(meta))
        return c, meta

    def mo_coeff(self, spin):
        nmo = self.nup if spin == 0 else self.ndn
        C = np.zeros((self.nbasis, nmo))
        rc = np.zeros(self.nbasis)
        lib().orc_mo_coeff(self.h, spin, _p(C), _p(rc))
        return C, rc

    def run(self, pos, method, nblock, nstep, timestep, seed, weights=None, ndecorr=2,
            pp_stochastic=1, nthreads=1, step0=0, state=None, feedback_interval=5,
            feedback=1.0, max_weight=7.0, branch_start_cutoff=10.0, tmoves=0):
        pos = np.ascontiguousarray(pos, dtype=np.float64).copy()
        nw = pos.shape[0]
        w = np.ones(nw) if weights is None else np.ascontiguousarray(weights, dtype=np.float64).copy()
        eb = np.zeros((nblock, 5))
        st = np.zeros(4) if state is None else np.ascontiguousarray(state, dtype=np.float64)
        o = RunOpts(1 if method.upper() == "DMC" else 0, nw, nblock, nstep, ndecorr, timestep,
                    seed, pp_stochastic, feedback_interval, feedback, max_weight,
                    branch_start_cutoff, nthreads, step0, tmoves)
        acc = lib().orc_run(self.h, ctypes.byref(o), _p(pos), _p(w), _p(eb), _p(st))
        return dict(pos=pos, weights=w, eblock=eb, acceptance=acc, state=st)


def transpose_inverse(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    n = a.shape[0]
    a1 = np.zeros_like(a)
    ld = np.zeros(1)
    sg = np.zeros(1)
    ok = lib().orc_transpose_inverse(n, _p(a), _p(a1), _p(ld), _p(sg))
    return ok, a1, ld[0], sg[0]


def inverse_update_row(a1, newrow, row):
    a1 = np.ascontiguousarray(a1, dtype=np.float64).copy()
    v = np.ascontiguousarray(newrow, dtype=np.float64)
    f = lib().orc_inverse_update_row(a1.shape[0], _p(a1), _p(v), row)
    return a1, f


def philox(ctr, key):
    c = (ctypes.c_uint * 4)(*ctr)
    k = (ctypes.c_uint * 2)(*key)
    o = (ctypes.c_uint * 4)()
    lib().orc_philox(c, k, o)
    return list(o)


def parse_golden(text):
    """Parse qw_golden output into a list of dicts."""
    out = []
    cur = None
    for line in text.splitlines():
        f = line.split()
        if not f:
            continue
        if f[0] == "config":
            cur = dict(elec=[], parm=[], pdgrad=[], pdgd=[], pdvg=[], tmove=[])
            out.append(cur)
        elif f[0] == "elec":
            cur["elec"].append([float(x) for x in f[2:6]])
        elif f[0] in ("parm", "pdgrad"):
            cur[f[0]].append(float(f[2]))
        elif f[0] == "pdgd":
            cur["pdgd"].append([float(x) for x in f[3:7]])
        elif f[0] == "pdvg":
            cur["pdvg"].append([float(x) for x in f[2:5]])
        elif f[0] == "tmove":
            cur["tmove"].append([float(x) for x in f[1:6]])
        elif f[0] in ("pdok", "nparm"):
            cur[f[0]] = int(f[1])
        elif f[0] == "sign":
            cur["sign"] = int(f[1])
        else:
            cur[f[0]] = float(f[1])
    for c in out:
        c["elec"] = np.array(c["elec"])
        npar = c.get("nparm", 0)
        ne = len(c["elec"])
        c["parm"] = np.array(c["parm"])
        c["pdgrad"] = np.array(c["pdgrad"])
        c["pdgd"] = np.array(c["pdgd"]).reshape(npar, ne, 4)
        c["pdvg"] = np.array(c["pdvg"]).reshape(-1, 3)
        c["tmove"] = np.array(c["tmove"]).reshape(-1, 5)  # e, vxx, displacement
    return out
