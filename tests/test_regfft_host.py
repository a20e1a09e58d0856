"""The two-pass register transforms of heavy-disks_amd/csrc/dfk_regfft.h run on
the CPU, lane by lane with the kernels' LDS layouts (bin/regfft_check), at
every length the kernels are built for, against numpy's FFT: the row r2c
(two real rows packed as one complex row, split into their half spectra), the
row c2r (FFTW's convention: the imaginary parts of the DC and Nyquist bins
ignored) and the column convolution (zero-padded forward transform, times a
real even kernel, inverse, first half kept), the latter also by the
three-pass form (one column per wavefront)."""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(os.path.dirname(HERE), "heavy-disks_amd", "bin", "regfft_check")
LENGTHS = [128, 192, 256, 384, 512, 768, 1024]
TOL = 2e-15


def run(mode, n, vals):
    assert os.path.exists(EXE), "build first: make -C heavy-disks_amd"
    r = subprocess.run([EXE, mode, str(n)], input="\n".join("%.17g" % x for x in vals),
                       capture_output=True, text=True, check=True)
    return np.array(r.stdout.split(), float)


def cplx(z):
    return np.c_[z.real, z.imag].ravel()


@pytest.mark.parametrize("n", LENGTHS)
def test_rows_r2c(n):
    rng = np.random.default_rng(n)
    a, b = rng.standard_normal(n), rng.standard_normal(n)
    v = run("r2c", n, np.r_[a, b]).reshape(-1, 2)
    X = v[:, 0] + 1j * v[:, 1]
    Xa, Xb = X[:n // 2 + 1], X[n // 2 + 1:]
    scale = np.max(np.abs(np.fft.rfft(a)))
    assert np.max(np.abs(Xa - np.fft.rfft(a))) / scale < TOL
    assert np.max(np.abs(Xb - np.fft.rfft(b))) / scale < TOL


@pytest.mark.parametrize("n", LENGTHS)
def test_rows_c2r(n):
    rng = np.random.default_rng(n + 1)
    A, B = np.fft.rfft(rng.standard_normal(n)), np.fft.rfft(rng.standard_normal(n))
    ra, rb = np.fft.irfft(A, n) * n, np.fft.irfft(B, n) * n
    A[0] += 0.3j          # dropped, as FFTW's c2r drops them
    B[-1] -= 0.2j
    y = run("c2r", n, np.r_[cplx(A), cplx(B)])
    scale = np.max(np.abs(ra))
    assert np.max(np.abs(y[:n] - ra)) / scale < TOL
    assert np.max(np.abs(y[n:] - rb)) / scale < TOL


@pytest.mark.parametrize("M", LENGTHS)
def test_column_convolution(M):
    rng = np.random.default_rng(M + 2)
    Nu = M // 2
    x = rng.standard_normal(Nu) + 1j * rng.standard_normal(Nu)
    K = rng.standard_normal(Nu + 1)
    Kfull = np.r_[K, K[1:Nu][::-1]]          # real and even: K[M - p] = K[p]
    ref = np.fft.ifft(Kfull * np.fft.fft(np.r_[x, np.zeros(Nu)]))[:Nu] * M
    w = run("conv", M, np.r_[cplx(x), K]).reshape(-1, 2)
    z = w[:, 0] + 1j * w[:, 1]
    assert np.max(np.abs(z - ref)) / np.max(np.abs(ref)) < TOL


@pytest.mark.parametrize("n", [128, 192, 256, 384, 512])
def test_rows_three_pass(n):
    """The three-pass row transforms (one row pair per wavefront, P-, 8- and
    8-point DFTs) against numpy: r2c of two real rows, and c2r with the DC
    and Nyquist imaginary parts dropped; impulses at the rows' ends too."""
    rng = np.random.default_rng(n + 9)
    for a, b in ((rng.standard_normal(n), rng.standard_normal(n)),
                 (np.eye(n)[0], np.eye(n)[n - 1])):
        v = run("r2c3", n, np.r_[a, b]).reshape(-1, 2)
        X = v[:, 0] + 1j * v[:, 1]
        scale = max(np.max(np.abs(np.fft.rfft(a))), np.max(np.abs(np.fft.rfft(b))))
        assert np.max(np.abs(X[:n // 2 + 1] - np.fft.rfft(a))) / scale < TOL
        assert np.max(np.abs(X[n // 2 + 1:] - np.fft.rfft(b))) / scale < TOL
    A, B = np.fft.rfft(rng.standard_normal(n)), np.fft.rfft(rng.standard_normal(n))
    ra, rb = np.fft.irfft(A, n) * n, np.fft.irfft(B, n) * n
    A[0] += 0.3j
    B[-1] -= 0.2j
    y = run("c2r3", n, np.r_[cplx(A), cplx(B)])
    assert np.all(np.isfinite(y))
    scale = np.max(np.abs(ra))
    assert np.max(np.abs(y[:n] - ra)) / scale < TOL
    assert np.max(np.abs(y[n:] - rb)) / scale < TOL


@pytest.mark.parametrize("M", [256, 384, 512, 768, 1024])
def test_column_convolution_three_pass(M):
    """The three-pass column convolution (one column per wavefront, P-, 8-
    and 8-point DFTs, lanes above 8P repeating lane 8(P-1)'s work for P < 8)
    against numpy, and against the two-pass form of the same length."""
    rng = np.random.default_rng(M + 5)
    Nu = M // 2
    x = rng.standard_normal(Nu) + 1j * rng.standard_normal(Nu)
    K = rng.standard_normal(Nu + 1)
    Kfull = np.r_[K, K[1:Nu][::-1]]
    ref = np.fft.ifft(Kfull * np.fft.fft(np.r_[x, np.zeros(Nu)]))[:Nu] * M
    w = run("conv3", M, np.r_[cplx(x), K]).reshape(-1, 2)
    z = w[:, 0] + 1j * w[:, 1]
    assert np.all(np.isfinite(z))
    assert np.max(np.abs(z - ref)) / np.max(np.abs(ref)) < TOL
    w2 = run("conv", M, np.r_[cplx(x), K]).reshape(-1, 2)
    assert np.max(np.abs(z - (w2[:, 0] + 1j * w2[:, 1]))) / np.max(np.abs(ref)) < TOL


def test_column_convolution_three_pass_impulses():
    """Impulses at the first and last rows and a K that is 1 at one bin:
    the result is that bin's plane wave restricted to the rows (a transposed
    index, a dropped twiddle or a wrong K bin moves or scales it)."""
    M, Nu = 512, 256
    for r0, p0 in ((0, 0), (Nu - 1, 3), (77, Nu), (130, 101)):
        x = np.zeros(Nu, complex)
        x[r0] = 1.0
        K = np.zeros(Nu + 1)
        K[p0] = 1.0
        Kfull = np.r_[K, K[1:Nu][::-1]]
        ref = np.fft.ifft(Kfull * np.fft.fft(np.r_[x, np.zeros(Nu)]))[:Nu] * M
        w = run("conv3", M, np.r_[cplx(x), K]).reshape(-1, 2)
        z = w[:, 0] + 1j * w[:, 1]
        assert np.max(np.abs(z - ref)) < 1e-14, (r0, p0)


def test_unsupported_length_refused():
    r = subprocess.run([EXE, "r2c", "96"], input="", capture_output=True, text=True)
    assert r.returncode == 2 and "no r2c transform of length 96" in r.stderr
    r = subprocess.run([EXE, "conv3", "128"], input="", capture_output=True, text=True)
    assert r.returncode == 2 and "no conv3 transform of length 128" in r.stderr
