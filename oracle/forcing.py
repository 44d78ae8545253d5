"""CPU restatement of the turbulent forcing (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.

* ``philox4x64`` -- Philox4x64-10 (Salmon et al., SC 2011), the counter-based
  generator the native forcing draws its noise from; pinned bit for bit
  against ``numpy.random.Philox`` in tests/test_forcing_oracle.py.
* ``spectrum`` -- ``_create_forcing_field`` up to the inverse FFT
  (``_physics_modules/_turbulent_forcing/_turbulent_forcing.py:14-75``) with
  that noise: sqrt(P(k)) times complex normals, DC removed, compressive part
  projected out; ``field`` = Re(ifftn).
* ``apply`` -- ``_apply_forcing`` (``_turbulent_forcing.py:77-104``): the
  amplitude from the energy-injection quadratic and v += amp w.
"""

from __future__ import annotations

import math

import numpy as np

F32 = np.float32
M0, M1 = 0xD2E7470EE14C6C93, 0xCA5A826395121157
W0, W1 = 0x9E3779B97F4A7C15, 0xBB67AE8584CAA73B
MASK = (1 << 64) - 1


def philox4x64(ctr, key):
    """One Philox4x64-10 block: 4 words from a 4-word counter and 2-word key."""
    x = [int(v) & MASK for v in ctr]
    k0, k1 = int(key[0]) & MASK, int(key[1]) & MASK
    for _ in range(10):
        p0, p1 = M0 * x[0], M1 * x[2]
        x = [(p1 >> 64) ^ x[1] ^ k0, p1 & MASK, (p0 >> 64) ^ x[3] ^ k1, p0 & MASK]
        k0, k1 = (k0 + W0) & MASK, (k1 + W1) & MASK
    return x


def normal_pair(e, step, seed):
    """Box-Muller pair of element e at a step (float64 arithmetic, float32 out)."""
    c = philox4x64([e, step, 0, 0], [seed, 0])
    u1 = float((c[0] >> 11) + 1) * 2.0 ** -53
    u2 = float(c[1] >> 11) * 2.0 ** -53
    r = math.sqrt(-2.0 * math.log(u1))
    th = 6.283185307179586 * u2
    return F32(r * math.cos(th)), F32(r * math.sin(th))


def wavenumbers(n, box_size):
    """2 pi fftfreq(n, d = L / n) in float32."""
    return (F32(2.0 * math.pi / box_size) * np.fft.fftfreq(n, d=1.0 / n).astype(F32)).astype(F32)


def spectrum(shape, box_size, seed, step):
    nx, ny, nz = shape
    ncell = nx * ny * nz
    kx = wavenumbers(nx, box_size).reshape(nx, 1, 1)
    ky = wavenumbers(ny, box_size).reshape(1, ny, 1)
    kz = wavenumbers(nz, box_size).reshape(1, 1, nz)
    k2 = (kx * kx + ky * ky) + kz * kz
    kk = np.sqrt(k2)
    kp = F32(4.0 * math.pi / box_size)
    pk = (((kk * kk) * (kk * kk)) * (kk * kk)) * np.exp((F32(-8.0) * kk) / kp)
    sp = np.sqrt(pk).astype(F32)
    noise = np.empty((3, ncell), np.complex64)
    for d in range(3):
        for c in range(ncell):
            a, b = normal_pair(d * ncell + c, step, seed)
            noise[d, c] = complex(a, b)
    cw = (sp[None] * noise.reshape(3, nx, ny, nz)).astype(np.complex64)
    cw[:, 0, 0, 0] = 0
    safe = np.where(k2 == 0, F32(1), k2)
    div = (kx * cw[0] + ky * cw[1] + kz * cw[2]) / safe
    div[0, 0, 0] = 0
    out = np.stack([cw[0] - kx * div, cw[1] - ky * div, cw[2] - kz * div]).astype(np.complex64)
    return out


def field(cw):
    return np.real(np.fft.ifftn(cw, axes=(1, 2, 3))).astype(F32)


def apply(state, w, dt, edot, dx, rho_row=0, vel_row=1):
    """(new state, amp) of _apply_forcing for the real field w."""
    s = np.array(state, dtype=F32, copy=True)
    rho = s[rho_row]
    u, v, z = s[vel_row], s[vel_row + 1], s[vel_row + 2]
    ta = F32(0.5) * F32(np.sum((rho * ((w[0] * w[0] + w[1] * w[1]) + w[2] * w[2])).astype(np.float64)))
    tb = F32(np.sum(((rho * u * w[0] + rho * v * w[1]) + rho * z * w[2]).astype(np.float64)))
    tc = -((F32(edot) * F32(dt)) / F32(dx ** 3))
    disc = tb * tb - (F32(4) * ta) * tc
    amp = (-tb + np.sqrt(disc)) / (F32(2) * ta) if (disc >= 0 and abs(ta) > 1e-10) else F32(0)
    amp = F32(amp)
    for d in range(3):
        s[vel_row + d] = s[vel_row + d] + amp * w[d]
    return s, amp
