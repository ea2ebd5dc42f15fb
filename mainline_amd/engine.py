"""ctypes binding of libqwalk_amd.so (the C ABI in include/qwalk_amd.h).

The product path: every call goes to the HIP engine.  If the shared library is missing or
fails to load, the import of this module's Engine raises; there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqwalk_amd.so")

PREC = {"fp64": 0, "mixed_fp32": 1, "mixed_bf16": 2, "mixed_bf16x6": 3}
STATUS = {0: "ok", -1: "bad argument", -2: "bad input", -3: "unsupported", -4: "device error", -5: "out of memory"}

BLOCK_FIELDS = 12
_dp = ctypes.POINTER(ctypes.c_double)
_lib = None


class EngineOpts(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("device", ctypes.c_int), ("seed", ctypes.c_ulonglong),
                ("refresh_interval", ctypes.c_int)]


class MethodDesc(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int), ("nconfig", ctypes.c_int), ("nblock", ctypes.c_int),
                ("nstep", ctypes.c_int), ("ndecorr", ctypes.c_int), ("timestep", ctypes.c_double),
                ("feedback_interval", ctypes.c_int), ("feedback", ctypes.c_double),
                ("max_weight", ctypes.c_double), ("branch_start_cutoff", ctypes.c_double),
                ("pp_stochastic", ctypes.c_int), ("tmoves", ctypes.c_int), ("start_feedback", ctypes.c_double)]


class ServerOpts(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("hbm_budget", ctypes.c_ulonglong), ("max_active", ctypes.c_int)]


class JobDesc(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("seed", ctypes.c_ulonglong), ("nwalkers", ctypes.c_int),
                ("equil_steps", ctypes.c_int), ("method", MethodDesc)]


JOB_STATS = 10


class QwError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s (%s)" % (msg, STATUS.get(status, status)))
        self.status = status


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libqwalk_amd.so is not built; run __graft_entry__.build() or "
                              "make -f mainline_amd/Makefile")
        L = ctypes.CDLL(LIB_PATH)
        L.qw_engine_create.argtypes = [ctypes.c_char_p, ctypes.POINTER(EngineOpts), ctypes.POINTER(ctypes.c_void_p)]
        L.qw_engine_destroy.argtypes = [ctypes.c_void_p]
        L.qw_engine_clone.argtypes = [ctypes.c_void_p, ctypes.POINTER(EngineOpts), ctypes.POINTER(ctypes.c_void_p)]
        L.qw_last_error.restype = ctypes.c_char_p
        L.qw_last_error.argtypes = [ctypes.c_void_p]
        L.qw_engine_info.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        L.qw_eval_fixed.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp,
                                    _dp, _dp, _dp, _dp, _dp]
        L.qw_eval_fixed_static.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 2 + [ctypes.c_int] + \
            [ctypes.c_void_p] * 7
        L.qw_eval_fixed_static_max.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        L.qw_method_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(MethodDesc)]
        L.qw_method_count.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]
        L.qw_method_read_at.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(MethodDesc), ctypes.c_char_p,
                                        ctypes.c_int]
        L.qw_walkers_set.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
        L.qw_walkers_get.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
        L.qw_walkers_state.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.qw_run.argtypes = [ctypes.c_void_p, ctypes.POINTER(MethodDesc), _dp, _dp]
        L.qw_walkers_random.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_ulonglong, ctypes.c_double, _dp]
        L.qw_kernel_times.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, _dp,
                                      ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.qw_engine_walker_bytes.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_ulonglong)]
        L.qw_engine_tuning.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        L.qw_parm_count.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), _dp]
        L.qw_parm_deriv.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        L.qw_batch_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p)]
        L.qw_batch_destroy.argtypes = [ctypes.c_void_p]
        L.qw_batch_error.restype = ctypes.c_char_p
        L.qw_batch_error.argtypes = [ctypes.c_void_p]
        L.qw_batch_run.argtypes = [ctypes.c_void_p, ctypes.POINTER(MethodDesc), _dp, _dp]
        L.qw_batch_open.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        L.qw_batch_join.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(MethodDesc),
                                    ctypes.POINTER(ctypes.c_int)]
        L.qw_batch_step.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.qw_batch_leave.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
        L.qw_batch_kernel_times.argtypes = [ctypes.c_void_p, _dp, ctypes.POINTER(ctypes.c_int),
                                            ctypes.POINTER(ctypes.c_int)]
        L.qw_server_create.argtypes = [ctypes.POINTER(ServerOpts), ctypes.POINTER(ctypes.c_void_p)]
        L.qw_server_submit.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(JobDesc),
                                       ctypes.POINTER(ctypes.c_longlong)]
        L.qw_server_wait.argtypes = [ctypes.c_void_p, ctypes.c_longlong, _dp, _dp]
        L.qw_server_error.restype = ctypes.c_char_p
        L.qw_server_error.argtypes = [ctypes.c_void_p, ctypes.c_longlong]
        L.qw_server_destroy.argtypes = [ctypes.c_void_p]
        _ip = ctypes.POINTER(ctypes.c_int)
        L.qw_configs_read.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _ip, _dp, _dp, _ip]
        L.qw_configs_write.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _dp, _dp, _ip]
        L.qw_method_files.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        L.qw_configs_error.restype = ctypes.c_char_p
        L.qw_configs_error.argtypes = []
        _lib = L
    return _lib


def exported_symbols():
    """Names of the qw_* functions the loaded library exports (no device calls)."""
    L = lib()
    names = []
    for n in ["qw_engine_create", "qw_engine_clone", "qw_engine_destroy", "qw_last_error", "qw_engine_info", "qw_eval_fixed",
              "qw_eval_fixed_static", "qw_eval_fixed_static_max",
              "qw_method_read", "qw_method_count", "qw_method_read_at", "qw_walkers_set", "qw_walkers_get", "qw_walkers_state", "qw_walkers_random", "qw_run",
              "qw_kernel_times", "qw_engine_walker_bytes", "qw_engine_tuning", "qw_parm_count",
              "qw_parm_deriv", "qw_batch_create", "qw_batch_destroy", "qw_batch_error", "qw_batch_run",
              "qw_batch_open", "qw_batch_join", "qw_batch_step", "qw_batch_leave",
              "qw_batch_kernel_times", "qw_server_create", "qw_server_submit",
              "qw_server_wait", "qw_server_error", "qw_server_destroy", "qw_configs_read", "qw_configs_write",
              "qw_method_files", "qw_configs_error"]:
        if hasattr(L, n):
            names.append(n)
    return names


def _p(a):
    return None if a is None else a.ctypes.data_as(_dp)


def read_method(deck):
    m = MethodDesc()
    st = lib().qw_method_read(deck.encode(), ctypes.byref(m))
    if st:
        raise QwError(st, lib().qw_last_error(None).decode())
    return m


def read_methods(deck):
    """Every METHOD section of a deck in input order: [(MethodDesc, label)]."""
    n = ctypes.c_int()
    st = lib().qw_method_count(deck.encode(), ctypes.byref(n))
    if st:
        raise QwError(st, lib().qw_last_error(None).decode())
    out = []
    for i in range(n.value):
        m = MethodDesc()
        lab = ctypes.create_string_buffer(64)
        st = lib().qw_method_read_at(deck.encode(), i, ctypes.byref(m), lab, 64)
        if st:
            raise QwError(st, lib().qw_last_error(None).decode())
        out.append((m, lab.value.decode()))
    return out


def method_files(deck, index=0):
    """(READCONFIG, STORECONFIG) of the index-th METHOD section (<deck>.config by default)."""
    r, w = ctypes.create_string_buffer(4096), ctypes.create_string_buffer(4096)
    st = lib().qw_method_files(deck.encode(), index, r, w, 4096)
    if st:
        raise QwError(st, lib().qw_configs_error().decode())
    return r.value.decode(), w.value.decode()


def read_configs(path, ne):
    """Walker file -> (pos[nw][ne][3], weights[nw], signs[nw]) (weights 1 without DMC records)."""
    L, n = lib(), ctypes.c_int()
    st = L.qw_configs_read(path.encode(), ne, 0, ctypes.byref(n), None, None, None)
    if st:
        raise QwError(st, L.qw_configs_error().decode())
    pos = np.zeros((n.value, ne, 3))
    w = np.zeros(n.value)
    sg = np.zeros(n.value, dtype=np.int32)
    st = L.qw_configs_read(path.encode(), ne, n.value, ctypes.byref(n), _p(pos), _p(w),
                           sg.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    if st:
        raise QwError(st, L.qw_configs_error().decode())
    return pos, w, sg


def write_configs(path, pos, weights=None, signs=None):
    """Walker file in the reference's layout: VMC records, or DMC records when weights are given."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    nw, ne = pos.shape[0], pos.shape[1]
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    sg = None if signs is None else np.ascontiguousarray(signs, dtype=np.int32)
    st = lib().qw_configs_write(path.encode(), nw, ne, _p(pos), _p(w),
                                None if sg is None else sg.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    if st:
        raise QwError(st, lib().qw_configs_error().decode())


def method(kind, nconfig, nblock, nstep, timestep, ndecorr=2, feedback_interval=5, feedback=1.0,
           max_weight=7.0, branch_start_cutoff=10.0, pp_stochastic=1, tmoves=0, start_feedback=1.0):
    """tmoves (DMC): 0 = off, 1 = TMOVES, 2 = TMOVESSC (Casula T-moves)."""
    return MethodDesc(1 if kind.upper() == "DMC" else 0, nconfig, nblock, nstep, ndecorr, timestep,
                      feedback_interval, feedback, max_weight, branch_start_cutoff, pp_stochastic, tmoves,
                      start_feedback)


class Engine:
    """One walker engine on one device for the SYSTEM / PSEUDO / TRIALFUNC of a deck."""

    def __init__(self, deck, precision="fp64", device=0, seed=20261202, refresh_interval=0, _from=None):
        """_from: an Engine whose parsed deck is reused (qw_engine_clone) instead of reading `deck`."""
        h = ctypes.c_void_p()
        opts = EngineOpts(PREC[precision], device, seed, refresh_interval)
        if _from is not None:
            st = lib().qw_engine_clone(_from.h, ctypes.byref(opts), ctypes.byref(h))
        else:
            st = lib().qw_engine_create(deck.encode(), ctypes.byref(opts), ctypes.byref(h))
        if st:
            raise QwError(st, lib().qw_last_error(None).decode())
        self.h = h
        info = (ctypes.c_int * 6)()
        lib().qw_engine_info(self.h, info)
        self.nup, self.ndn, self.natoms, self.nbasis, self.has_pp = list(info)[:5]
        self.ne = self.nup + self.ndn
        self.precision = precision

    def clone(self, precision=None, device=0, seed=20261202, refresh_interval=0):
        """A new engine for the same deck (not read again) with other options."""
        return Engine(None, precision or self.precision, device, seed, refresh_interval, _from=self)

    def close(self):
        if getattr(self, "h", None):
            lib().qw_engine_destroy(self.h)
            self.h = None

    __del__ = close

    def _check(self, st):
        if st:
            raise QwError(st, lib().qw_last_error(self.h).decode())

    def eval(self, pos, rot=None, pp_mode=0, pp_rand=None):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        K = pos.shape[0]
        out = dict(logpsi=np.zeros(K), sign=np.zeros(K), grad=np.zeros((K, self.ne, 3)),
                   lap=np.zeros((K, self.ne)), eparts=np.zeros((K, 3)))
        r = None if rot is None else np.ascontiguousarray(rot, dtype=np.float64)
        pr = None if pp_rand is None else np.ascontiguousarray(pp_rand, dtype=np.float64)
        self._check(lib().qw_eval_fixed(self.h, K, _p(pos), _p(r), pp_mode, _p(pr), _p(out["logpsi"]),
                                        _p(out["sign"]), _p(out["grad"]), _p(out["lap"]), _p(out["eparts"])))
        return out

    def eval_static(self, K, pos, rot, pp_mode, pp_rand, outs, stream=None):
        """qw_eval_fixed_static on device buffers: pos, rot, pp_rand and the five outputs
        (logpsi, sign, grad, lap, eparts) are device addresses (int) or None."""
        self._check(lib().qw_eval_fixed_static(self.h, K, pos, rot, pp_mode, pp_rand, *outs, stream))

    def eval_static_max(self):
        k = ctypes.c_int()
        self._check(lib().qw_eval_fixed_static_max(self.h, ctypes.byref(k)))
        return k.value

    def parms(self):
        """Values of the linear Jastrow parameters (getParmDeriv order)."""
        n = ctypes.c_int()
        self._check(lib().qw_parm_count(self.h, ctypes.byref(n), None))
        v = np.zeros(max(n.value, 1))
        self._check(lib().qw_parm_count(self.h, ctypes.byref(n), _p(v)))
        return v[:n.value]

    def parm_deriv(self, pos):
        """gradient[K][np], hessian[K][np][np], gradderiv[K][np][Ne][4], val_gradient[K][Ne][3]
        of Wavefunction::getParmDeriv at fixed configurations."""
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        K = pos.shape[0]
        n = len(self.parms())
        m = max(n, 1)
        out = dict(gradient=np.zeros((K, m)), hessian=np.zeros((K, m, m)), gradderiv=np.zeros((K, m, self.ne, 4)),
                   val_gradient=np.zeros((K, self.ne, 3)))
        self._check(lib().qw_parm_deriv(self.h, K, _p(pos), _p(out["gradient"]), _p(out["hessian"]),
                                        _p(out["gradderiv"]), _p(out["val_gradient"])))
        out["gradient"] = out["gradient"][:, :n]
        out["hessian"] = out["hessian"][:, :n, :n]
        out["gradderiv"] = out["gradderiv"][:, :n]
        return out

    def set_walkers(self, pos, weights=None):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        self._check(lib().qw_walkers_set(self.h, pos.shape[0], _p(pos), _p(w)))

    def get_walkers(self, nw):
        pos = np.zeros((nw, self.ne, 3))
        w = np.zeros(nw)
        self._check(lib().qw_walkers_get(self.h, nw, _p(pos), _p(w)))
        return pos, w

    def walker_state(self, nw):
        """logdet[nw][2], sign[nw][2] and the last local-energy parts elast[nw][4] (kinetic,
        Coulomb, nonlocal, total) of the resident walkers, and dgl[nw][Ne][4], the determinant
        gradient and laplacian ratios the last sweep of each spin ended with."""
        ld, sg, el = np.zeros((nw, 2)), np.zeros((nw, 2)), np.zeros((nw, 4))
        dgl = np.zeros((nw, self.ne, 4))
        self._check(lib().qw_walkers_state(self.h, nw, _p(ld), _p(sg), _p(el), _p(dgl)))
        return dict(logdet=ld, sign=sg, elast=el, dgl=dgl)

    def random_walkers(self, nw, seed=1, inirange=3.0):
        pos = np.zeros((nw, self.ne, 3))
        self._check(lib().qw_walkers_random(self.h, nw, seed, inirange, _p(pos)))
        return pos

    def run(self, m):
        """eblock[b] = E, kinetic, Coulomb, nonlocal, weight sum, points, var(E), var(KE),
        var(V), var(NL), mean weight, var(weight) (QW_BLOCK_FIELDS)."""
        eb = np.zeros((max(m.nblock, 1), BLOCK_FIELDS))
        stats = np.zeros(4)
        self._check(lib().qw_run(self.h, ctypes.byref(m), _p(eb), _p(stats)))
        return dict(eblock=eb[:m.nblock], acceptance=stats[0], moves=stats[1], seconds=stats[2],
                    moves_per_s=stats[3])

    def tuning(self):
        """Launch choices: sweep (U, plus 100 when the coefficients are staged in LDS; 0 while
        the tuner measures), HIP graphs (0/1), block-per-walker path (0/1), inverse tile width."""
        v = (ctypes.c_int * 4)()
        self._check(lib().qw_engine_tuning(self.h, v))
        return dict(sweep=v[0], graphs=v[1], big=v[2], ns=v[3])

    def kernel_times(self):
        buf = ctypes.create_string_buffer(4096)
        avg = np.zeros(64)
        cnt = (ctypes.c_int * 64)()
        nk = ctypes.c_int()
        self._check(lib().qw_kernel_times(self.h, buf, 4096, _p(avg), cnt, 64, ctypes.byref(nk)))
        names = buf.value.decode().split(",") if nk.value else []
        return {n: (avg[i], cnt[i]) for i, n in enumerate(names)}


class Batch:
    """Several engines (one device, one precision) running DMC in shared launches
    (qw_batch_*); each engine's results equal its own qw_run bit for bit."""

    def __init__(self, engines=None, precision=None, device=0, refresh_interval=0):
        """Batch(engines): a fixed set for run(); Batch(precision=...): an empty batch that
        engines join() and leave between step() calls."""
        h = ctypes.c_void_p()
        if engines is not None:
            self.engines = list(engines)
            arr = (ctypes.c_void_p * len(self.engines))(*[e.h.value for e in self.engines])
            st = lib().qw_batch_create(len(self.engines), arr, ctypes.byref(h))
        else:
            self.engines = []
            st = lib().qw_batch_open(device, PREC[precision], refresh_interval, ctypes.byref(h))
        if st:
            raise QwError(st, lib().qw_batch_error(None).decode())
        self.h = h
        self.nblock = {}

    def join(self, engine, runs):
        """The engine (walkers set) joins with its runs (MethodDesc list); returns its id."""
        arr = (MethodDesc * len(runs))(*runs)
        mid = ctypes.c_int()
        st = lib().qw_batch_join(self.h, engine.h, len(runs), arr, ctypes.byref(mid))
        if st:
            raise QwError(st, lib().qw_batch_error(self.h).decode())
        self.nblock[mid.value] = runs[-1].nblock
        self.engines.append(engine)
        return mid.value

    def step(self):
        """One step of every member; returns the ids of members whose runs ended."""
        fin = (ctypes.c_int * 64)()
        nf = ctypes.c_int()
        st = lib().qw_batch_step(self.h, fin, ctypes.byref(nf))
        if st:
            raise QwError(st, lib().qw_batch_error(self.h).decode())
        return [fin[i] for i in range(nf.value)]

    def leave(self, mid):
        eb = np.zeros((self.nblock[mid], BLOCK_FIELDS))
        stats = np.zeros(4)
        st = lib().qw_batch_leave(self.h, mid, _p(eb), _p(stats))
        if st:
            raise QwError(st, lib().qw_batch_error(self.h).decode())
        return dict(eblock=eb, acceptance=stats[0], moves=stats[1], seconds=stats[2], moves_per_s=stats[3])

    def run(self, m):
        """eblock[engine][block][QW_BLOCK_FIELDS], and the run's moves, seconds,
        electron-moves/s, tuner state and per-engine acceptance."""
        n = len(self.engines)
        eb = np.zeros((n, max(m.nblock, 1), BLOCK_FIELDS))
        stats = np.zeros(4 + n)
        st = lib().qw_batch_run(self.h, ctypes.byref(m), _p(eb), _p(stats))
        if st:
            raise QwError(st, lib().qw_batch_error(self.h).decode())
        return dict(eblock=eb[:, :m.nblock], moves=stats[0], seconds=stats[1], moves_per_s=stats[2],
                    tuned=bool(stats[3]), acceptance=stats[4:])

    def kernel_times(self):
        avg = np.zeros(4)
        cnt = (ctypes.c_int * 4)()
        tun = (ctypes.c_int * 10)()
        st = lib().qw_batch_kernel_times(self.h, _p(avg), cnt, tun)
        if st:
            raise QwError(st, "batch kernel times")
        names = ["sweep", "recompute", "local", "branch"]
        buckets = {tun[2 * i]: tun[2 * i + 1] for i in range(5) if tun[2 * i]}
        return {names[k]: (avg[k], cnt[k]) for k in range(4)}, buckets

    def close(self):
        if getattr(self, "h", None):
            lib().qw_batch_destroy(self.h)
            self.h = None

    __del__ = close


class Server:
    """The serving loop (qw_server_*): jobs run concurrently on one device, admitted by HBM
    budget in arrival order."""

    def __init__(self, device=0, hbm_budget=0, max_active=0):
        h = ctypes.c_void_p()
        o = ServerOpts(device, hbm_budget, max_active)
        st = lib().qw_server_create(ctypes.byref(o), ctypes.byref(h))
        if st:
            raise QwError(st, "server create failed")
        self.h = h
        self.nblock = {}

    def submit(self, deck, m, nwalkers, precision="fp64", seed=1, equil_steps=0):
        j = JobDesc(PREC[precision], seed, nwalkers, equil_steps, m)
        jid = ctypes.c_longlong()
        st = lib().qw_server_submit(self.h, deck.encode(), ctypes.byref(j), ctypes.byref(jid))
        if st:
            raise QwError(st, "submit failed")
        self.nblock[jid.value] = m.nblock
        return jid.value

    def wait(self, jid):
        """Block records of the job and its stats: submit, admitted, finished, latency (s),
        electron moves, run seconds, electron-moves/s, device bytes, set-up seconds, 1 = ran in
        a batch."""
        eb = np.zeros((self.nblock[jid], BLOCK_FIELDS))
        stats = np.zeros(JOB_STATS)
        st = lib().qw_server_wait(self.h, jid, _p(eb), _p(stats))
        if st:
            raise QwError(st, lib().qw_server_error(self.h, jid).decode())
        return eb, stats

    def close(self):
        if getattr(self, "h", None):
            lib().qw_server_destroy(self.h)
            self.h = None

    __del__ = close
