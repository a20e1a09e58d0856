"""ctypes binding of the diskfft oracle -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, and
nowhere else.  It is the checker, never the product: the HIP library does not
link or call it.  The functions restate diskfft/src/{hydro.hpp,kalnajs.hpp,
diskmain.cpp}; see diskfft_oracle.h for the reference line of each one and
for what pins it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_D = ctypes.c_double
_I = ctypes.c_int
_P = ctypes.c_void_p
_DP = ctypes.POINTER(ctypes.c_double)


def _dp(a):
    return a.ctypes.data_as(_DP)


class Sample(ctypes.Structure):
    _fields_ = [("t", _D), ("c0", _D), ("c2", _D), ("ph2", _D)]


def load(fast=False):
    name = "liboracle_fast.so" if fast else "liboracle.so"
    lib = ctypes.CDLL(os.path.join(HERE, name))
    sig = {
        "orc_hydro_new": (_P, [_I, _I, _D, _D, _D, _D, _D]),
        "orc_hydro_free": (None, [_P]),
        "orc_hydro_timestep": (_D, [_P, _D]),
        "orc_hydro_source": (None, [_P, _D, _DP, _DP]),
        "orc_hydro_transportR": (None, [_P, _D]),
        "orc_hydro_transportPhi": (None, [_P, _D]),
        "orc_hydro_floor": (None, [_P]),
        "orc_hydro_step": (None, [_P, _D, _DP, _DP]),
        "orc_hydro_ptr": (_DP, [_P, _I]),
        "orc_fft": (_I, [_I, _DP, _DP, _I]),
        "orc_kal_new": (_P, [_I, _I, _D, _D, _D, _D]),
        "orc_kal_free": (None, [_P]),
        "orc_kal_potential": (None, [_P, _DP, _DP]),
        "orc_kal_forces": (None, [_P, _DP, _DP, _DP]),
        "orc_kal_potential_direct": (None, [_P, _DP, _DP]),
        "orc_kal_r2c2d": (None, [_P, _DP, _DP, _DP]),
        "orc_kal_c2r2d": (None, [_P, _DP, _DP, _DP]),
        "orc_kal_khat": (_DP, [_P, _I]),
        "orc_cm": (None, [_P, _I, _D, _DP, _DP]),
        "orc_disk_mass": (_D, [_P]),
        "orc_ic_lka98": (None, [_P, _P, _I, _D]),
        "orc_run": (_I, [_P, _P, _I, _D, _D, _D, ctypes.c_long, _D,
                         ctypes.POINTER(Sample), _I, _DP, _DP,
                         ctypes.POINTER(ctypes.c_long)]),
        "orc_run_from": (_I, [_P, _P, _I, _D, _D, _D, ctypes.c_long, _D,
                              ctypes.POINTER(Sample), _I, _DP, _DP,
                              ctypes.POINTER(ctypes.c_long), _DP]),
        "orc_set_threads": (None, [_I]),
        "orc_get_threads": (_I, []),
    }
    for k, (res, args) in sig.items():
        f = getattr(lib, k)
        f.restype = res
        f.argtypes = args
    return lib


_LIB = {}


def lib(fast=False):
    if fast not in _LIB:
        _LIB[fast] = load(fast)
    return _LIB[fast]


class Hydro:
    """DiskHydro restated (hydro.hpp:20-180)."""

    def __init__(self, Nu, Nphi, Rin=0.25, Rout=1.0, K=0.25, gamma=2.0, gmstar=0.6, fast=False):
        self.L = lib(fast)
        self.Nu, self.Nphi = Nu, Nphi
        self.h = self.L.orc_hydro_new(Nu, Nphi, Rin, Rout, K, gamma, gmstar)
        self.N = Nu * Nphi

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_hydro_free(self.h)
            self.h = None

    def _arr(self, which, n):
        p = self.L.orc_hydro_ptr(self.h, which)
        return np.ctypeslib.as_array(p, shape=(n,))

    # pointers move after source() swaps buffers: always fetch fresh views
    @property
    def Sig(self):
        return self._arr(0, self.N).reshape(self.Nu, self.Nphi)

    @property
    def vR(self):
        return self._arr(1, self.N).reshape(self.Nu, self.Nphi)

    @property
    def vp(self):
        return self._arr(2, self.N).reshape(self.Nu, self.Nphi)

    @property
    def R(self):
        return self._arr(3, self.Nu).copy()

    @property
    def Rf(self):
        return self._arr(4, self.Nu + 1).copy()

    @property
    def Vol(self):
        return self._arr(5, self.Nu).copy()

    def set_state(self, Sig, vR, vp):
        self.Sig[...] = Sig
        self.vR[...] = vR
        self.vp[...] = vp

    def state(self):
        return self.Sig.copy(), self.vR.copy(), self.vp.copy()

    def timestep(self, cfl):
        return self.L.orc_hydro_timestep(self.h, cfl)

    @staticmethod
    def _g(g):
        if g is None:
            return None, None
        a = np.ascontiguousarray(g, dtype=np.float64)
        return a, _dp(a)

    def source(self, dt, gR=None, gp=None):
        a, pa = self._g(gR)
        b, pb = self._g(gp)
        self.L.orc_hydro_source(self.h, dt, pa, pb)

    def transportR(self, dt):
        self.L.orc_hydro_transportR(self.h, dt)

    def transportPhi(self, dt):
        self.L.orc_hydro_transportPhi(self.h, dt)

    def floor(self):
        self.L.orc_hydro_floor(self.h)

    def step(self, dt, gR=None, gp=None):
        a, pa = self._g(gR)
        b, pb = self._g(gp)
        self.L.orc_hydro_step(self.h, dt, pa, pb)

    def cm(self, m, mD):
        mag, ph = _D(), _D()
        self.L.orc_cm(self.h, m, mD, ctypes.byref(mag), ctypes.byref(ph))
        return mag.value, ph.value

    def disk_mass(self):
        return self.L.orc_disk_mass(self.h)


class Kalnajs:
    """KalnajsSolver restated (kalnajs.hpp:19-117)."""

    def __init__(self, Nu, Nphi, Rin=0.25, Rout=1.0, eps=0.02, G=1.0, fast=False):
        self.L = lib(fast)
        self.Nu, self.Nphi = Nu, Nphi
        self.M, self.Nc = 2 * Nu, Nphi // 2 + 1
        self.k = self.L.orc_kal_new(Nu, Nphi, Rin, Rout, eps, G)
        if not self.k:
            raise ValueError("unsupported grid")

    def __del__(self):
        if getattr(self, "k", None):
            self.L.orc_kal_free(self.k)
            self.k = None

    def potential(self, Sig):
        s = np.ascontiguousarray(Sig, dtype=np.float64)
        out = np.empty((self.Nu, self.Nphi))
        self.L.orc_kal_potential(self.k, _dp(s), _dp(out))
        return out

    def potential_direct(self, Sig):
        s = np.ascontiguousarray(Sig, dtype=np.float64)
        out = np.empty((self.Nu, self.Nphi))
        self.L.orc_kal_potential_direct(self.k, _dp(s), _dp(out))
        return out

    def forces(self, Phi):
        p = np.ascontiguousarray(Phi, dtype=np.float64)
        gR = np.empty((self.Nu, self.Nphi))
        gp = np.empty((self.Nu, self.Nphi))
        self.L.orc_kal_forces(self.k, _dp(p), _dp(gR), _dp(gp))
        return gR, gp

    def r2c2d(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        re = np.empty((self.M, self.Nc))
        im = np.empty((self.M, self.Nc))
        self.L.orc_kal_r2c2d(self.k, _dp(x), _dp(re), _dp(im))
        return re + 1j * im

    def c2r2d(self, X):
        re = np.ascontiguousarray(X.real)
        im = np.ascontiguousarray(X.imag)
        out = np.empty((self.M, self.Nphi))
        self.L.orc_kal_c2r2d(self.k, _dp(re), _dp(im), _dp(out))
        return out

    def khat(self):
        n = self.M * self.Nc
        re = np.ctypeslib.as_array(self.L.orc_kal_khat(self.k, 0), shape=(n,)).copy()
        im = np.ctypeslib.as_array(self.L.orc_kal_khat(self.k, 1), shape=(n,)).copy()
        return (re + 1j * im).reshape(self.M, self.Nc)


def fft(x, sign=-1):
    x = np.asarray(x, dtype=np.complex128)
    re = np.ascontiguousarray(x.real)
    im = np.ascontiguousarray(x.imag)
    rc = lib().orc_fft(len(x), _dp(re), _dp(im), sign)
    if rc:
        raise ValueError("fft plan failed")
    return re + 1j * im


def ic_lka98(h, k, selfgrav=1, amp=1e-3):
    h.L.orc_ic_lka98(h.h, k.k if k is not None else None, selfgrav, amp)


def run(h, k, selfgrav, tlim, dtout=0.25, cfl=0.3, max_steps=0, mD=None, cap=100000):
    if mD is None:
        mD = h.disk_mass()
    buf = (Sample * cap)()
    tf, c2f = _D(), _D()
    st = ctypes.c_long()
    n = h.L.orc_run(h.h, k.k if k is not None else None, selfgrav, tlim, dtout, cfl,
                    max_steps, mD, buf, cap, ctypes.byref(tf), ctypes.byref(c2f),
                    ctypes.byref(st))
    samples = np.array([(buf[i].t, buf[i].c0, buf[i].c2, buf[i].ph2)
                        for i in range(min(n, cap))]).reshape(-1, 4)
    return samples, tf.value, c2f.value, st.value


def run_chunked(h, k, selfgrav, tlim, dtout=0.25, cfl=0.3, max_steps=0, mD=None, chunk=20000,
                progress=None):
    """orc_run in pieces of `chunk` steps (same result); progress(t, steps)
    is called after each piece."""
    if mD is None:
        mD = h.disk_mass()
    t, tn, st = _D(0.0), _D(0.0), ctypes.c_long(0)
    c2f = _D()
    out = []
    cap = 1 << 16
    buf = (Sample * cap)()
    while True:
        cap_steps = st.value + chunk if max_steps <= 0 else min(max_steps, st.value + chunk)
        n = h.L.orc_run_from(h.h, k.k if k is not None else None, selfgrav, tlim, dtout, cfl,
                             cap_steps, mD, buf, cap, ctypes.byref(t), ctypes.byref(tn),
                             ctypes.byref(st), ctypes.byref(c2f))
        out.extend((buf[i].t, buf[i].c0, buf[i].c2, buf[i].ph2) for i in range(min(n, cap)))
        if progress:
            progress(t.value, st.value)
        if not (t.value < tlim) or (max_steps > 0 and st.value >= max_steps):
            break
    return np.array(out).reshape(-1, 4), t.value, c2f.value, st.value


def set_threads(n):
    lib().orc_set_threads(n)
    lib(True).orc_set_threads(n) if os.path.exists(os.path.join(HERE, "liboracle_fast.so")) else None


class RefHydro:
    """The reference's own hydro.hpp, compiled from its source into oracle/_ref
    (only where the reference checkout exists)."""

    PATH = os.path.join(HERE, "_ref", "libref_hydro.so")

    @classmethod
    def available(cls):
        return os.path.exists(cls.PATH)

    def __init__(self, Nu, Nphi, Rin=0.25, Rout=1.0, K=0.25, gamma=2.0, gmstar=0.6):
        L = ctypes.CDLL(self.PATH)
        L.ref_hydro_new.restype = _P
        L.ref_hydro_new.argtypes = [_I, _I, _D, _D, _D, _D, _D]
        L.ref_hydro_free.argtypes = [_P]
        L.ref_hydro_set.argtypes = [_P, _DP, _DP, _DP]
        L.ref_hydro_get.argtypes = [_P, _DP, _DP, _DP]
        L.ref_hydro_geom.argtypes = [_P, _DP, _DP, _DP]
        L.ref_hydro_timestep.restype = _D
        L.ref_hydro_timestep.argtypes = [_P, _D]
        L.ref_hydro_stage.argtypes = [_P, _I, _D, _DP, _DP]
        L.ref_fft2d.argtypes = [_DP, _DP, _I, _I, _I]
        self.L = L
        self.Nu, self.Nphi = Nu, Nphi
        self.h = L.ref_hydro_new(Nu, Nphi, Rin, Rout, K, gamma, gmstar)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ref_hydro_free(self.h)
            self.h = None

    def set_state(self, Sig, vR, vp):
        a, b, c = (np.ascontiguousarray(x, dtype=np.float64) for x in (Sig, vR, vp))
        self.L.ref_hydro_set(self.h, _dp(a), _dp(b), _dp(c))

    def state(self):
        a, b, c = (np.empty((self.Nu, self.Nphi)) for _ in range(3))
        self.L.ref_hydro_get(self.h, _dp(a), _dp(b), _dp(c))
        return a, b, c

    def geom(self):
        R, Rf, V = np.empty(self.Nu), np.empty(self.Nu + 1), np.empty(self.Nu)
        self.L.ref_hydro_geom(self.h, _dp(R), _dp(Rf), _dp(V))
        return R, Rf, V

    def timestep(self, cfl):
        return self.L.ref_hydro_timestep(self.h, cfl)

    def stage(self, stage, dt, gR=None, gp=None):
        a = None if gR is None else np.ascontiguousarray(gR, dtype=np.float64)
        b = None if gp is None else np.ascontiguousarray(gp, dtype=np.float64)
        self.L.ref_hydro_stage(self.h, stage, dt, None if a is None else _dp(a),
                               None if b is None else _dp(b))

    def fft2d(self, x, sign=-1):
        x = np.asarray(x, dtype=np.complex128)
        re = np.ascontiguousarray(x.real)
        im = np.ascontiguousarray(x.imag)
        self.L.ref_fft2d(_dp(re), _dp(im), x.shape[0], x.shape[1], sign)
        return re + 1j * im
