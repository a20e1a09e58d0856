"""Python binding of libdiskfft (the C ABI in include/diskfft.h).

Mirrors the reference's C++ interface for the diskfft time-step loop:
``Disk.timestep/step/potential/forces`` follow ``DiskHydro`` (diskfft/src/
hydro.hpp:20-180) and ``KalnajsSolver`` (diskfft/src/kalnajs.hpp:19-117);
``Ensemble.run`` follows the loop of diskfft/src/diskmain.cpp:86-113.

The library is the HIP build in lib/libdiskfft.so.  There is no CPU fallback:
a missing library raises ImportError, a missing GPU raises DiskfftError.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "lib", "libdiskfft.so")

_D = ctypes.c_double
_I = ctypes.c_int
_LL = ctypes.c_longlong
_P = ctypes.c_void_p
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)

DFK_OK, DFK_EINVAL, DFK_EHIP, DFK_ENOMEM, DFK_EUNSUPPORTED, DFK_ESTATE = range(6)

# every entry point declared in include/diskfft.h
SYMBOLS = [
    "dfk_last_error", "dfk_device_count", "dfk_create", "dfk_destroy", "dfk_num_disks",
    "dfk_geometry", "dfk_set_state", "dfk_get_state", "dfk_timestep", "dfk_potential",
    "dfk_forces", "dfk_step", "dfk_stage", "dfk_khat", "dfk_modes", "dfk_disk_mass",
    "dfk_run_init", "dfk_advance", "dfk_samples", "dfk_get_status", "dfk_set_shapes",
    "dfk_get_shapes", "dfk_set_fft_options", "dfk_get_fft_options", "dfk_last_advance_ms", "dfk_set_graph_steps", "dfk_profile_step",
    "dfk_set_threads", "dfk_get_threads", "dfk_set_update_kernel", "dfk_get_update_kernel",
    "dfk_set_chains", "dfk_get_chains", "dfk_set_row_transform", "dfk_get_row_transform",
    "dfk_set_fused_rows", "dfk_get_fused_rows",
    "dfk_snapshot_save", "dfk_snapshot_load", "dfk_set_eos", "dfk_potential_noalloc",
]


class DiskParams(ctypes.Structure):
    _fields_ = [("Nu", _I), ("Nphi", _I), ("Rin", _D), ("Rout", _D), ("K", _D),
                ("gamma", _D), ("gmstar", _D), ("eps", _D), ("G", _D), ("selfgrav", _I)]


class RunParams(ctypes.Structure):
    _fields_ = [("tlim", _D), ("dtout", _D), ("cfl", _D), ("mD", _D), ("max_steps", _LL),
                ("stop_on_nonfinite", _I)]


class Sample(ctypes.Structure):
    _fields_ = [("t", _D), ("c0", _D), ("c2", _D), ("ph2", _D)]


class Status(ctypes.Structure):
    _fields_ = [("t", _D), ("steps", _LL), ("done", _I), ("flags", _I), ("c2_final", _D),
                ("ph2_final", _D)]


class DiskfftError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"diskfft error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise ImportError(f"{LIBPATH} is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIBPATH)
        H = ctypes.POINTER(_P)
        sig = {
            "dfk_last_error": (ctypes.c_char_p, []),
            "dfk_device_count": (_I, []),
            "dfk_create": (_I, [ctypes.POINTER(DiskParams), _I, _I, H]),
            "dfk_destroy": (None, [_P]),
            "dfk_num_disks": (_I, [_P]),
            "dfk_geometry": (_I, [_P, _I, _DP, _DP, _DP]),
            "dfk_set_state": (_I, [_P, _I, _DP, _DP, _DP]),
            "dfk_get_state": (_I, [_P, _I, _DP, _DP, _DP]),
            "dfk_timestep": (_I, [_P, _I, _D, _DP]),
            "dfk_potential": (_I, [_P, _I, _DP, _DP]),
            "dfk_forces": (_I, [_P, _I, _DP, _DP, _DP]),
            "dfk_step": (_I, [_P, _I, _D, _DP, _DP]),
            "dfk_stage": (_I, [_P, _I, _I, _D, _DP, _DP]),
            "dfk_khat": (_I, [_P, _I, _DP, _DP]),
            "dfk_modes": (_I, [_P, _I, _I, _D, _DP, _DP]),
            "dfk_disk_mass": (_I, [_P, _I, _DP]),
            "dfk_run_init": (_I, [_P, ctypes.POINTER(RunParams), _I]),
            "dfk_advance": (_I, [_P, _I, _IP]),
            "dfk_samples": (_I, [_P, _I, ctypes.POINTER(Sample), _I, _IP]),
            "dfk_get_status": (_I, [_P, _I, ctypes.POINTER(Status)]),
            "dfk_set_shapes": (_I, [_P, _I, _I, _I, _I, _I]),
            "dfk_get_shapes": (_I, [_P, _I, _IP, _IP, _IP, _IP]),
            "dfk_set_fft_options": (_I, [_P, _I, _I, _I, _I]),
            "dfk_get_fft_options": (_I, [_P, _I, _IP, _IP, _IP]),
            "dfk_last_advance_ms": (_I, [_P, _DP]),
            "dfk_set_graph_steps": (_I, [_P, _I]),
            "dfk_profile_step": (_I, [_P, _I, _DP]),
            "dfk_set_threads": (_I, [_P, _I, _I, _I]),
            "dfk_get_threads": (_I, [_P, _IP, _IP, _IP]),
            "dfk_set_update_kernel": (_I, [_P, _I]),
            "dfk_get_update_kernel": (_I, [_P, _IP]),
            "dfk_set_chains": (_I, [_P, _I]),
            "dfk_set_row_transform": (_I, [_P, _I]),
            "dfk_get_row_transform": (_I, [_P, _IP]),
            "dfk_set_fused_rows": (_I, [_P, _I]),
            "dfk_get_fused_rows": (_I, [_P, _IP]),
            "dfk_get_chains": (_I, [_P, _IP]),
            "dfk_snapshot_save": (_I, [_P, _I, ctypes.c_char_p]),
            "dfk_snapshot_load": (_I, [_P, _I, ctypes.c_char_p]),
            "dfk_set_eos": (_I, [_P, _I, _D, _D]),
            "dfk_potential_noalloc": (_I, [_P, _I, _P, _P, _P]),
        }
        for k, (res, args) in sig.items():
            f = getattr(L, k)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != DFK_OK:
        raise DiskfftError(rc, lib().dfk_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(_DP)


def _arr(x, shape):
    a = np.ascontiguousarray(x, dtype=np.float64)
    if a.shape != shape:
        a = a.reshape(shape)
    return a


def device_count():
    return lib().dfk_device_count()


def disk_params(Nu, Nphi, Rin=0.25, Rout=1.0, K=0.25, gamma=2.0, gmstar=0.6, eps=0.02, G=1.0,
                selfgrav=1):
    """Defaults are the reference driver's LKA98 constants (diskmain.cpp:28-29)."""
    return dict(Nu=Nu, Nphi=Nphi, Rin=Rin, Rout=Rout, K=K, gamma=gamma, gmstar=gmstar, eps=eps,
                G=G, selfgrav=selfgrav)


class Ensemble:
    """One or more disks resident on one GPU (a single disk is an ensemble of one)."""

    def __init__(self, params, device=0):
        if isinstance(params, dict):
            params = [params]
        self.params = [dict(p) for p in params]
        arr = (DiskParams * len(params))(*[DiskParams(**p) for p in params])
        h = _P()
        _check(lib().dfk_create(arr, len(params), device, ctypes.byref(h)))
        self.h = h
        self.n = len(params)

    def close(self):
        if getattr(self, "h", None):
            lib().dfk_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def shape(self, d):
        return self.params[d]["Nu"], self.params[d]["Nphi"]

    # ---- DiskHydro / KalnajsSolver mirror -----------------------------
    def geometry(self, d=0):
        Nu, _ = self.shape(d)
        R, Rf, V = np.empty(Nu), np.empty(Nu + 1), np.empty(Nu)
        _check(lib().dfk_geometry(self.h, d, _dp(R), _dp(Rf), _dp(V)))
        return R, Rf, V

    def set_state(self, Sig, vR, vp, d=0):
        s = self.shape(d)
        a, b, c = (_arr(x, s) for x in (Sig, vR, vp))
        _check(lib().dfk_set_state(self.h, d, _dp(a), _dp(b), _dp(c)))

    def get_state(self, d=0):
        s = self.shape(d)
        a, b, c = (np.empty(s) for _ in range(3))
        _check(lib().dfk_get_state(self.h, d, _dp(a), _dp(b), _dp(c)))
        return a, b, c

    def timestep(self, cfl, d=0):
        dt = _D()
        _check(lib().dfk_timestep(self.h, d, cfl, ctypes.byref(dt)))
        return dt.value

    def potential(self, Sig, d=0):
        s = self.shape(d)
        a = _arr(Sig, s)
        out = np.empty(s)
        _check(lib().dfk_potential(self.h, d, _dp(a), _dp(out)))
        return out

    def potential_noalloc(self, sigma_dev, phi_dev, stream=None, d=0):
        """dfk_potential_noalloc: Phi from Sigma, both device addresses (ints)
        on the handle's device; enqueued on `stream` (a hipStream_t address,
        None: the device's null stream), nothing allocated."""
        _check(lib().dfk_potential_noalloc(self.h, d, _P(sigma_dev), _P(phi_dev),
                                           None if stream is None else _P(stream)))

    def forces(self, Phi, d=0):
        s = self.shape(d)
        a = _arr(Phi, s)
        gR, gp = np.empty(s), np.empty(s)
        _check(lib().dfk_forces(self.h, d, _dp(a), _dp(gR), _dp(gp)))
        return gR, gp

    def step(self, dt, gR=None, gp=None, d=0):
        s = self.shape(d)
        a = None if gR is None else _arr(gR, s)
        b = None if gp is None else _arr(gp, s)
        _check(lib().dfk_step(self.h, d, dt, None if a is None else _dp(a),
                              None if b is None else _dp(b)))

    def stage(self, stage, dt, gR=None, gp=None, d=0):
        s = self.shape(d)
        a = None if gR is None else _arr(gR, s)
        b = None if gp is None else _arr(gp, s)
        _check(lib().dfk_stage(self.h, d, stage, dt, None if a is None else _dp(a),
                               None if b is None else _dp(b)))

    def khat(self, d=0):
        Nu, Nphi = self.shape(d)
        M, Nc = 2 * Nu, Nphi // 2 + 1
        re, im = np.empty((M, Nc)), np.empty((M, Nc))
        _check(lib().dfk_khat(self.h, d, _dp(re), _dp(im)))
        return re + 1j * im

    def modes(self, m, mD, d=0):
        mag, ph = _D(), _D()
        _check(lib().dfk_modes(self.h, d, m, mD, ctypes.byref(mag), ctypes.byref(ph)))
        return mag.value, ph.value

    def disk_mass(self, d=0):
        m = _D()
        _check(lib().dfk_disk_mass(self.h, d, ctypes.byref(m)))
        return m.value

    # ---- the driver loop (diskmain.cpp:86-113) ------------------------
    def run_init(self, tlim, dtout=0.25, cfl=0.3, mD=None, max_steps=0, stop_on_nonfinite=0,
                 sample_cap=256):
        def per(x):
            return list(x) if isinstance(x, (list, tuple, np.ndarray)) else [x] * self.n
        tl, do, cf, ms, st = per(tlim), per(dtout), per(cfl), per(max_steps), per(stop_on_nonfinite)
        if mD is None:
            mD = [self.disk_mass(d) for d in range(self.n)]
        md = per(mD)
        rp = (RunParams * self.n)(*[RunParams(tl[d], do[d], cf[d], md[d], int(ms[d]), int(st[d]))
                                     for d in range(self.n)])
        _check(lib().dfk_run_init(self.h, rp, sample_cap))

    def advance(self, nsteps):
        done = _I()
        _check(lib().dfk_advance(self.h, nsteps, ctypes.byref(done)))
        return bool(done.value)

    def samples(self, d=0, cap=1 << 16):
        buf = (Sample * cap)()
        n = _I()
        _check(lib().dfk_samples(self.h, d, buf, cap, ctypes.byref(n)))
        return np.array([(buf[i].t, buf[i].c0, buf[i].c2, buf[i].ph2)
                         for i in range(n.value)]).reshape(-1, 4)

    def status(self, d=0):
        s = Status()
        _check(lib().dfk_get_status(self.h, d, ctypes.byref(s)))
        return dict(t=s.t, steps=s.steps, done=bool(s.done), flags=s.flags, c2_final=s.c2_final,
                    ph2_final=s.ph2_final)

    def set_shapes(self, ti=0, tj=0, pr=0, cb=0, d=None):
        for q in (range(self.n) if d is None else [d]):
            _check(lib().dfk_set_shapes(self.h, q, ti, tj, pr, cb))

    def get_shapes(self, d=0):
        v = [_I() for _ in range(4)]
        _check(lib().dfk_get_shapes(self.h, d, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def set_fft_options(self, rows_tw_lds=1, cols_tw_lds=1, cols_wave=0, d=None):
        """Twiddle tables of the row / column transforms in LDS (1) or read
        through the caches (0); column transforms workgroup-cooperative (0),
        wave-local in place (1) or workgroup in place (2); see
        dfk_set_fft_options."""
        for q in (range(self.n) if d is None else [d]):
            _check(lib().dfk_set_fft_options(self.h, q, int(rows_tw_lds), int(cols_tw_lds),
                                             int(cols_wave)))

    def get_fft_options(self, d=0):
        v = [_I() for _ in range(3)]
        _check(lib().dfk_get_fft_options(self.h, d, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def last_advance_ms(self):
        ms = _D()
        _check(lib().dfk_last_advance_ms(self.h, ctypes.byref(ms)))
        return ms.value

    def set_threads(self, rows=0, cols=0, update=0):
        _check(lib().dfk_set_threads(self.h, rows, cols, update))

    def get_threads(self):
        v = [_I() for _ in range(3)]
        _check(lib().dfk_get_threads(self.h, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def set_update_kernel(self, kind):
        """0: LDS tile kernel, 1: register-streaming kernel."""
        _check(lib().dfk_set_update_kernel(self.h, kind))

    def get_update_kernel(self):
        k = _I()
        _check(lib().dfk_get_update_kernel(self.h, ctypes.byref(k)))
        return k.value

    def set_row_transform(self, inplace):
        """Row transforms workgroup-cooperative (0) or in place (1)."""
        _check(lib().dfk_set_row_transform(self.h, int(inplace)))

    def get_row_transform(self):
        k = _I()
        _check(lib().dfk_get_row_transform(self.h, ctypes.byref(k)))
        return k.value

    def set_fused_rows(self, threads):
        """Time loop as column convolution + fused row kernels of `threads`
        (256, 512 or 1024) threads per workgroup, or four launches (0)."""
        _check(lib().dfk_set_fused_rows(self.h, int(threads)))

    def get_fused_rows(self):
        k = _I()
        _check(lib().dfk_get_fused_rows(self.h, ctypes.byref(k)))
        return k.value

    def set_chains(self, k):
        """Step the disks as k parallel graph chains (dfk_set_chains)."""
        _check(lib().dfk_set_chains(self.h, k))

    def get_chains(self):
        k = _I()
        _check(lib().dfk_get_chains(self.h, ctypes.byref(k)))
        return k.value

    def set_eos(self, K, gamma=2.0, d=0):
        _check(lib().dfk_set_eos(self.h, d, K, gamma))
        self.params[d]["K"] = K
        self.params[d]["gamma"] = gamma

    def snapshot_save(self, path, d=0):
        _check(lib().dfk_snapshot_save(self.h, d, os.fsencode(path)))

    def snapshot_load(self, path, d=0):
        _check(lib().dfk_snapshot_load(self.h, d, os.fsencode(path)))

    def set_graph_steps(self, n):
        _check(lib().dfk_set_graph_steps(self.h, n))

    def profile_step(self, nrep=50):
        """Average ms per launch of (control, r2c, conv, c2r, update)."""
        ms = np.zeros(5)
        _check(lib().dfk_profile_step(self.h, nrep, _dp(ms)))
        return ms

    def run(self, tlim, dtout=0.25, cfl=0.3, mD=None, max_steps=0, chunk=2000):
        """Whole loop to completion; returns per-disk (samples, status)."""
        self.run_init(tlim, dtout, cfl, mD, max_steps)
        out = [[] for _ in range(self.n)]
        while True:
            done = self.advance(chunk)
            for d in range(self.n):
                s = self.samples(d)
                if len(s):
                    out[d].append(s)
            if done:
                break
        res = []
        for d in range(self.n):
            s = np.concatenate(out[d]) if out[d] else np.zeros((0, 4))
            res.append((s, self.status(d)))
        return res


# ---------------------------------------------------------------------------
# disk models (include/diskfft_model.h; host code, no GPU needed)
MODEL_LKA98, MODEL_POWERLAW = 0, 1
MODEL_SYMBOLS = ["dfk_model_defaults", "dfk_model_sigma0", "dfk_model_state", "dfk_model_mass",
                 "dfk_splitmix64", "dfk_normal", "dfk_extract_growth", "dfk_extract_growth_ex"]


class Growth(ctypes.Structure):
    _fields_ = [("gamma", _D), ("omega_p", _D), ("sat", _D), ("t_lo", _D), ("t_hi", _D),
                ("n_fit", _I), ("grew", _I), ("sat_peak", _D)]


class Model(ctypes.Structure):
    _fields_ = [("kind", _I), ("Rin", _D), ("Rout", _D), ("gmstar", _D), ("G", _D),
                ("gamma", _D), ("K", _D), ("amp", _D), ("mass_ratio", _D), ("Qmin", _D),
                ("slope", _D), ("taper_in", _D), ("taper_out", _D), ("taper_w", _D),
                ("q_in", _D), ("q_out", _D), ("Qmin_got", _D), ("noise_rms", _D), ("mmin", _I), ("mmax", _I), ("seed", ctypes.c_ulonglong)]


def _model_lib():
    L = lib()
    if not getattr(L, "_model_sig", False):
        L.dfk_model_defaults.argtypes = [_I, ctypes.POINTER(Model)]
        L.dfk_model_defaults.restype = None
        L.dfk_model_sigma0.argtypes = [ctypes.POINTER(Model), _I, _I, _DP]
        L.dfk_model_sigma0.restype = _I
        L.dfk_model_state.argtypes = [ctypes.POINTER(Model), _I, _I, _DP, _DP, _DP, _DP, _DP]
        L.dfk_model_state.restype = _I
        L.dfk_model_mass.argtypes = [ctypes.POINTER(Model), _I, _I, _DP]
        L.dfk_model_mass.restype = _D
        L.dfk_splitmix64.argtypes = [ctypes.POINTER(ctypes.c_ulonglong)]
        L.dfk_splitmix64.restype = ctypes.c_ulonglong
        L.dfk_normal.argtypes = [ctypes.POINTER(ctypes.c_ulonglong)]
        L.dfk_normal.restype = _D
        L.dfk_extract_growth.argtypes = [_DP, _DP, _DP, _DP, _I, _D, _D, _D,
                                         ctypes.POINTER(Growth)]
        L.dfk_extract_growth.restype = _I
        L.dfk_extract_growth_ex.argtypes = [_DP, _DP, _DP, _DP, _I, _D, _D, _D, _D, _D, _D,
                                            ctypes.POINTER(Growth)]
        L.dfk_extract_growth_ex.restype = _I
        L._model_sig = True
    return L


def model_defaults(kind, **over):
    m = Model()
    _model_lib().dfk_model_defaults(kind, ctypes.byref(m))
    for k, v in over.items():
        setattr(m, k, v)
    return m


def model_sigma0(m, Nu, Nphi):
    out = np.empty(Nu)
    if _model_lib().dfk_model_sigma0(ctypes.byref(m), Nu, Nphi, _dp(out)):
        raise ValueError("dfk_model_sigma0 failed")
    return out


def model_state(m, Nu, Nphi, gRdisk=None):
    """Equilibrium state; returns (Sig, vR, vp, Q).  m.K is updated (power-law Q solve)."""
    S, U, W = (np.empty((Nu, Nphi)) for _ in range(3))
    Q = np.empty(Nu)
    g = None if gRdisk is None else np.ascontiguousarray(gRdisk, dtype=np.float64)
    rc = _model_lib().dfk_model_state(ctypes.byref(m), Nu, Nphi, None if g is None else _dp(g),
                                      _dp(S), _dp(U), _dp(W), _dp(Q))
    if rc:
        raise ValueError(f"dfk_model_state failed ({rc})")
    return S, U, W, Q


def model_mass(m, Sig):
    Sig = np.ascontiguousarray(Sig, dtype=np.float64)
    return _model_lib().dfk_model_mass(ctypes.byref(m), Sig.shape[0], Sig.shape[1], _dp(Sig))


def model_ic(m, Nu, Nphi, eps, selfgrav=1, potential_forces=None, device=0):
    """Initial condition of model `m` on an Nu x Nphi grid (diskmain.cpp:36-70 flow).

    potential_forces(Sigma) -> (gR, gphi) computes the self-gravity of the
    axisymmetric disk; default: this library on the GPU.  Returns
    (disk params dict with the solved K, Sig, vR, vp, mD, Q)."""
    sig0 = model_sigma0(m, Nu, Nphi)
    gRdisk = None
    if selfgrav:
        Sax = np.repeat(sig0[:, None], Nphi, axis=1)
        if potential_forces is None:
            tmp = Ensemble(disk_params(Nu, Nphi, Rin=m.Rin, Rout=m.Rout, K=1.0, gamma=m.gamma,
                                       gmstar=m.gmstar, eps=eps, G=m.G, selfgrav=1), device)
            gR, _ = tmp.forces(tmp.potential(Sax))
            tmp.close()
        else:
            gR, _ = potential_forces(Sax)
        gRdisk = np.cumsum(gR, axis=1)[:, -1] / Nphi   # serial sum as diskmain.cpp:44-46
    S, U, W, Q = model_state(m, Nu, Nphi, gRdisk)
    p = disk_params(Nu, Nphi, Rin=m.Rin, Rout=m.Rout, K=m.K, gamma=m.gamma, gmstar=m.gmstar,
                    eps=eps, G=m.G, selfgrav=selfgrav)
    return p, S, U, W, model_mass(m, S), Q


def extract_growth(samples, lo_factor=3.0, hi_frac=0.3, grow_factor=10.0, hi_factor=None,
                   sat_frac=0.25, peak_window=None):
    """gamma, Omega_p and saturated |C2|/|C0| from an (n, 4) array of
    (t, |C0|, |C2|, arg C2) samples (see dfk_extract_growth; hi_factor,
    sat_frac and peak_window select dfk_extract_growth_ex, which also
    returns sat_peak)."""
    a = np.ascontiguousarray(np.asarray(samples, dtype=np.float64))
    cols = [np.ascontiguousarray(a[:, k]) for k in range(4)]
    g = Growth()
    ex = not (hi_factor is None and sat_frac == 0.25 and peak_window is None)
    if not ex:
        rc = _model_lib().dfk_extract_growth(*[_dp(c) for c in cols], len(a), lo_factor, hi_frac,
                                             grow_factor, ctypes.byref(g))
    else:
        rc = _model_lib().dfk_extract_growth_ex(
            *[_dp(c) for c in cols], len(a), lo_factor,
            np.inf if hi_factor is None else hi_factor, hi_frac, grow_factor, sat_frac,
            np.inf if peak_window is None else peak_window, ctypes.byref(g))
    if rc:
        raise ValueError("the growth extractor needs at least 4 samples and 0 < sat_frac <= 1")
    out = dict(gamma=g.gamma, omega_p=g.omega_p, sat=g.sat, t_lo=g.t_lo, t_hi=g.t_hi,
               n_fit=g.n_fit, grew=bool(g.grew))
    if ex:
        out["sat_peak"] = g.sat_peak
    return out
