"""GPU: dfk_potential_noalloc, the entry point that allocates nothing (its
half spectrum in the library's per-device scratch array), against
dfk_potential on the same inputs: the same bits, on two calls in a row, the
second on a smaller disk, and again after dfk_potential and a run have used
the handle.  The device buffers are the caller's, allocated here through the
HIP runtime the library itself links (/opt/rocm), as a caller would."""
import ctypes
import os

import numpy as np
import pytest

from helpers import relerr

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def dfk():
    import diskfft_amd
    if diskfft_amd.device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")
    return diskfft_amd


@pytest.fixture(scope="module")
def hip(dfk):
    dfk.lib()
    L = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so.7")   # the runtime libdiskfft.so links
    L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    L.hipFree.argtypes = [ctypes.c_void_p]
    L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    L.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    L.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    L.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    L.hipSetDevice.argtypes = [ctypes.c_int]
    for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipStreamCreate", "hipStreamSynchronize",
              "hipStreamDestroy", "hipSetDevice"):
        getattr(L, f).restype = ctypes.c_int
    assert L.hipSetDevice(0) == 0
    return L


class DevBuf:
    H2D, D2H = 1, 2

    def __init__(self, hip, shape):
        self.hip, self.shape = hip, shape
        self.n = int(np.prod(shape)) * 8
        p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(p), self.n) == 0
        self.p = p.value

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.nbytes == self.n
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, self.n, self.H2D) == 0

    def get(self):
        a = np.empty(self.shape)
        assert self.hip.hipMemcpy(a.ctypes.data, self.p, self.n, self.D2H) == 0
        return a

    def free(self):
        self.hip.hipFree(self.p)


@pytest.mark.parametrize("modes", [None, (2, 3), (3, 4)])
def test_potential_noalloc_matches_potential(dfk, hip, orc, modes):
    """modes: the handle's default transforms, the two-pass register row (2)
    and column (3) kernels, or the three-pass ones (3, 4)."""
    second = {None: (64, 96), (2, 3): (64, 128), (3, 4): (128, 128)}[modes]
    shapes = [(256, 256, 0.01), second + (0.02,)]   # then a smaller disk
    print("case potential_noalloc", shapes, "modes", modes, flush=True)
    e = dfk.Ensemble([dfk.disk_params(a, b, eps=c, Rin=0.1) for a, b, c in shapes])
    if modes is not None:
        e.set_row_transform(modes[0])
        for d in range(len(shapes)):
            e.set_fft_options(1, 0, modes[1], d=d)
        assert e.get_row_transform() == modes[0]
        assert all(e.get_fft_options(d)[2] == modes[1] for d in range(len(shapes)))
    s = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(s)) == 0
    rng = np.random.default_rng(3)
    sig = [rng.random(x[:2]) for x in shapes]
    bufs = [(DevBuf(hip, x[:2]), DevBuf(hip, x[:2])) for x in shapes]
    try:
        for rnd in range(2):
            for d, (a, b, _) in enumerate(shapes):
                bs, bp = bufs[d]
                bs.put(sig[d])
                bp.put(np.full((a, b), np.nan))
                # the caller's stream in the first round, the null stream (which
                # the blocking hipMemcpy below waits for) in the second
                e.potential_noalloc(bs.p, bp.p, stream=s.value if rnd == 0 else None, d=d)
                if rnd == 0:
                    assert hip.hipStreamSynchronize(s.value) == 0
                got = bp.get()            # hipMemcpy waits for the device
                ref = e.potential(sig[d], d=d)
                assert np.array_equal(got, ref), (rnd, d, relerr(got, ref))
                k = orc.Kalnajs(a, b, 0.1, 1.0, shapes[d][2])
                assert relerr(got, k.potential(sig[d])) < 1e-13
            # between the rounds the handle runs its own loop (its spectra,
            # control blocks and state stay its own)
            if rnd == 0:
                for d, (a, b, _) in enumerate(shapes):
                    e.set_state(sig[d], np.zeros((a, b)), np.ones((a, b)), d=d)
                e.run_init(1e9, mD=1.0, max_steps=3)
                e.advance(3)
                assert all(e.status(d)["steps"] == 3 for d in range(len(shapes)))
    finally:
        for bs, bp in bufs:
            bs.free()
            bp.free()
        hip.hipStreamDestroy(s.value)
        e.close()


def test_potential_noalloc_errors(dfk, hip):
    print("case potential_noalloc errors", flush=True)
    e = dfk.Ensemble([dfk.disk_params(32, 32), dfk.disk_params(32, 32, selfgrav=0)])
    b = DevBuf(hip, (32, 32))
    try:
        with pytest.raises(dfk.DiskfftError) as ei:
            e.potential_noalloc(b.p, b.p, d=1)        # no self-gravity: no K-hat
        assert ei.value.code == dfk.DFK_ESTATE
        with pytest.raises(dfk.DiskfftError) as ei:
            e.potential_noalloc(b.p, b.p, d=2)
        assert ei.value.code == dfk.DFK_EINVAL
        assert dfk.lib().dfk_potential_noalloc(e.h, 0, None, ctypes.c_void_p(b.p), None) == \
            dfk.DFK_EINVAL
    finally:
        b.free()
        e.close()
