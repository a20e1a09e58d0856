"""Surrogate of the hydro: disk parameters -> (gamma, Omega_p, saturated |C2|/|C0|).

An ensemble of BF16 MLPs trained on the GPU by libdiskfft's own fused kernels
(include/diskfft_mlp.h).  Inputs are (M_d/M*, Q_min, softening s, Sigma
slope), mapped linearly onto [-1, 1] from the ensemble's parameter box.  The
three targets are heavy tailed, so each is compressed as asinh(y / s), s being
the training set's median |y|, then standardised with the mean and deviation
of the compressed training values.
Where the members disagree by more than `spread_tol` (relative deviation of
any output across members) the caller should fall back to the hydro, which
remains the measure of the surrogate's error.
"""
import ctypes
import os

import numpy as np

import diskfft_amd as dfk

LO = np.array([0.3, 1.0, 0.05, -1.75])
HI = np.array([1.5, 2.0, 0.20, -1.25])
NPAR = 34051

_P = ctypes.c_void_p
_FP = ctypes.POINTER(ctypes.c_float)
_I = ctypes.c_int


class MlpError(RuntimeError):
    pass


def _lib():
    L = dfk.lib()
    if not getattr(L, "_mlp_sig", False):
        L.dfk_mlp_last_error.restype = ctypes.c_char_p
        L.dfk_mlp_param_count.restype = _I
        L.dfk_mlp_create.argtypes = [_I, _I, ctypes.c_ulonglong, ctypes.POINTER(_P)]
        L.dfk_mlp_create.restype = _I
        L.dfk_mlp_destroy.argtypes = [_P]
        L.dfk_mlp_destroy.restype = None
        L.dfk_mlp_get_params.argtypes = [_P, _I, _FP]
        L.dfk_mlp_set_params.argtypes = [_P, _I, _FP]
        L.dfk_mlp_forward.argtypes = [_P, _FP, _I, _FP]
        L.dfk_mlp_grad.argtypes = [_P, _FP, _FP, _I, _FP, _FP]
        L.dfk_mlp_train.argtypes = [_P, _FP, _FP, _I, _I, _I, ctypes.c_float, ctypes.c_ulonglong,
                                    _FP]
        L.dfk_mlp_mfma_probe.argtypes = [_FP, _FP, _FP]
        for f in ("dfk_mlp_get_params", "dfk_mlp_set_params", "dfk_mlp_forward", "dfk_mlp_grad",
                  "dfk_mlp_train", "dfk_mlp_mfma_probe"):
            getattr(L, f).restype = _I
        L._mlp_sig = True
    return L


def _chk(rc):
    if rc:
        raise MlpError(_lib().dfk_mlp_last_error().decode())


def _fp(a):
    return a.ctypes.data_as(_FP)


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a if shape is None else a.reshape(shape)


SYMBOLS = ["dfk_mlp_last_error", "dfk_mlp_param_count", "dfk_mlp_create", "dfk_mlp_destroy",
           "dfk_mlp_get_params", "dfk_mlp_set_params", "dfk_mlp_forward", "dfk_mlp_grad",
           "dfk_mlp_train", "dfk_mlp_mfma_probe"]


def mfma_probe(A, B):
    A, B = _f32(A, (16, 16)), _f32(B, (16, 16))
    D = np.zeros((16, 16), dtype=np.float32)
    _chk(_lib().dfk_mlp_mfma_probe(_fp(A), _fp(B), _fp(D)))
    return D


class MLP:
    """An ensemble of `members` networks on one GPU."""

    def __init__(self, members=5, seed=1, device=0):
        h = _P()
        _chk(_lib().dfk_mlp_create(members, device, seed, ctypes.byref(h)))
        self.h, self.members = h, members

    def close(self):
        if getattr(self, "h", None):
            _lib().dfk_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def get_params(self, m):
        w = np.zeros(NPAR, dtype=np.float32)
        _chk(_lib().dfk_mlp_get_params(self.h, m, _fp(w)))
        return w

    def set_params(self, m, w):
        w = _f32(w)
        _chk(_lib().dfk_mlp_set_params(self.h, m, _fp(w)))

    def forward(self, x):
        x = _f32(x)
        n = x.shape[0]
        y = np.zeros((self.members, n, 3), dtype=np.float32)
        _chk(_lib().dfk_mlp_forward(self.h, _fp(x), n, _fp(y)))
        return y

    def grad(self, x, t):
        x, t = _f32(x), _f32(t)
        g = np.zeros((self.members, NPAR), dtype=np.float32)
        loss = np.zeros(self.members, dtype=np.float32)
        _chk(_lib().dfk_mlp_grad(self.h, _fp(x), _fp(t), x.shape[0], _fp(g), _fp(loss)))
        return loss, g

    def train(self, x, t, epochs, batch=256, lr=1e-3, seed=7):
        x, t = _f32(x), _f32(t)
        hist = np.zeros(epochs, dtype=np.float32)
        _chk(_lib().dfk_mlp_train(self.h, _fp(x), _fp(t), x.shape[0], epochs, batch, lr, seed,
                                  _fp(hist)))
        return hist


class Surrogate:
    """Normalisation + MLP ensemble + spread-based fallback decision."""

    def __init__(self, members=5, seed=1, device=0):
        self.net = MLP(members, seed, device)
        self.mu = np.zeros(3)
        self.sd = np.ones(3)
        self.sc = np.ones(3)

    def _fwd(self, Y):
        return np.arcsinh(np.asarray(Y, dtype=np.float64) / self.sc)

    def _inv(self, Z):
        return self.sc * np.sinh(Z)

    @staticmethod
    def norm_x(X):
        return 2.0 * (np.asarray(X, dtype=np.float64) - LO) / (HI - LO) - 1.0

    def fit(self, X, Y, epochs=400, batch=256, lr=2e-3, seed=7):
        Y = np.asarray(Y, dtype=np.float64)
        self.sc = np.median(np.abs(Y), axis=0) + 1e-12
        Z = self._fwd(Y)
        self.mu = Z.mean(0)
        self.sd = Z.std(0) + 1e-12
        return self.net.train(self.norm_x(X), (Z - self.mu) / self.sd, epochs, batch, lr, seed)

    def predict_members(self, X):
        z = self.net.forward(self.norm_x(X)).astype(np.float64)
        return self._inv(z * self.sd + self.mu)

    def predict(self, X, spread_tol=0.05):
        """Mean prediction, relative member spread, and the mask of points
        where the ensemble disagrees by more than spread_tol (hydro fallback)."""
        ym = self.predict_members(X)
        mean = ym.mean(0)
        spread = ym.std(0) / np.maximum(np.abs(mean), 1e-12)
        return mean, spread, np.any(spread > spread_tol, axis=1)
