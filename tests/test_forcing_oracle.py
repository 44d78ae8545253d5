"""Pins the turbulent-forcing restatement (oracle/forcing.py) on the CPU:
Philox4x64-10 bit for bit against numpy's Philox generator, the Box-Muller
normals' moments, the spectrum's scaling along a k axis and its projection
(k . w_hat = 0), and the energy-injection identity of the amplitude
(_turbulent_forcing.py:77-104: the kinetic energy grows by Edot dt)."""

import math

import numpy as np
import pytest

from oracle import forcing as Fo


@pytest.mark.parametrize("key,start", [((5, 7), 0), ((2 ** 64 - 1, 3), 41), ((12345, 0), 2 ** 40)])
def test_philox_matches_numpy(key, start):
    # numpy's Philox advances its counter before producing a block
    bg = np.random.Philox(key=np.array(key, dtype=np.uint64), counter=np.array([start, 0, 0, 0], dtype=np.uint64))
    raw = bg.random_raw(12)
    for b in range(3):
        assert Fo.philox4x64([start + 1 + b, 0, 0, 0], key) == [int(v) for v in raw[4 * b:4 * b + 4]]


def test_normals_moments():
    z = np.array([Fo.normal_pair(e, 3, 42) for e in range(20000)], dtype=np.float64).ravel()
    assert abs(z.mean()) < 0.02 and abs(z.var() - 1) < 0.03
    assert abs(np.mean(z ** 4) - 3) < 0.15  # Gaussian kurtosis
    assert Fo.normal_pair(7, 3, 42) != Fo.normal_pair(7, 4, 42)


def test_spectrum_axis_scaling_and_projection():
    n, L = 8, 2.0
    cw = Fo.spectrum((n, n, n), L, seed=9, step=1)
    kx = Fo.wavenumbers(n, L)
    K = np.meshgrid(kx, kx, kx, indexing="ij")
    # solenoidal: k . w_hat = 0 everywhere, DC zero
    dot = K[0] * cw[0] + K[1] * cw[1] + K[2] * cw[2]
    assert np.abs(dot).max() < 1e-4 * np.abs(cw).max() * np.abs(kx).max()
    assert np.all(cw[:, 0, 0, 0] == 0)
    # on the kx axis the projection removes the x component and leaves
    # sqrt(P(k)) times the raw noise in y and z
    kp = 4 * math.pi / L
    for i in range(1, n):
        k = abs(float(kx[i]))
        sp = math.sqrt(k ** 6 * math.exp(-8 * k / kp))
        c = i * n * n  # flat index of (i, 0, 0)
        assert abs(cw[0, i, 0, 0]) < 1e-6 * sp
        for d in (1, 2):
            a, b = Fo.normal_pair(d * n ** 3 + c, 1, 9)
            np.testing.assert_allclose([cw[d, i, 0, 0].real, cw[d, i, 0, 0].imag], [sp * a, sp * b], rtol=2e-6)


def test_field_is_real_divergence_free():
    n, L = 8, 1.0
    w = Fo.field(Fo.spectrum((n, n, n), L, seed=1, step=0))
    assert w.shape == (3, n, n, n) and w.dtype == np.float32
    assert abs(w.mean()) < 1e-6 * np.abs(w).max()
    wh = np.fft.fftn(w, axes=(1, 2, 3))
    kx = Fo.wavenumbers(n, L)
    K = np.meshgrid(kx, kx, kx, indexing="ij")
    # the real part keeps the projection's k . w_hat = 0 (up to float32
    # round-off) away from the Nyquist planes, where fftfreq's -n/2 pairs a
    # mode with a partner that was projected along a different k
    dot = K[0] * wh[0] + K[1] * wh[1] + K[2] * wh[2]
    idx = np.indices((n, n, n))
    inner = np.all(idx != n // 2, axis=0)
    assert np.abs(dot[inner]).max() < 1e-4 * np.abs(wh).max() * np.abs(kx).max()


@pytest.mark.parametrize("edot", [0.5, 2.0])
def test_amplitude_injects_edot_dt(edot):
    rng = np.random.default_rng(2)
    n, dx, dt = 8, 1.0 / 8, 1e-3
    s = np.zeros((8, n, n, n), np.float32)
    s[0] = rng.uniform(0.5, 2, (n, n, n))
    s[1:4] = rng.normal(0, 0.3, (3, n, n, n))
    s[4] = 1.0
    w = Fo.field(Fo.spectrum((n, n, n), 1.0, seed=3, step=0))
    out, amp = Fo.apply(s, w, dt, edot, dx)

    def ke(x):
        return 0.5 * np.sum(x[0].astype(np.float64) * (x[1:4].astype(np.float64) ** 2).sum(0)) * dx ** 3

    assert amp > 0
    assert ke(out) - ke(s) == pytest.approx(edot * dt, rel=1e-3)
    assert np.array_equal(out[[0, 4, 5, 6, 7]], s[[0, 4, 5, 6, 7]])
