"""Seeded self-affine rough surfaces generated on the device: determinism,
rms slope, zero mean, PSD exponent -2(1+H), short-wavelength cutoff."""
import numpy as np
import pytest

import gfmd_hip

pytestmark = pytest.mark.gpu


def radial_psd(h):
    nx, ny = h.shape
    C = np.abs(np.fft.fft2(h)) ** 2
    qx = 2 * np.pi * np.fft.fftfreq(nx)
    qy = 2 * np.pi * np.fft.fftfreq(ny)
    q = np.hypot(qx[:, None], qy[None, :])
    return q.ravel(), C.ravel()


@pytest.mark.parametrize("H", [0.6, 0.8])
def test_rough_surface_statistics(H):
    n = 512
    h = gfmd_hip.rough_surface(n, n, hurst=H, rms_slope=0.1, lambda_short=4.0,
                               seed=1000)
    h2 = gfmd_hip.rough_surface(n, n, hurst=H, rms_slope=0.1, lambda_short=4.0,
                                seed=1000)
    h3 = gfmd_hip.rough_surface(n, n, hurst=H, rms_slope=0.1, lambda_short=4.0,
                                seed=1001)
    assert np.array_equal(h, h2)
    assert not np.allclose(h, h3)
    assert abs(h.mean()) < 1e-12 * np.abs(h).max()
    gx = (np.roll(h, -1, 0) - np.roll(h, 1, 0)) / 2
    gy = (np.roll(h, -1, 1) - np.roll(h, 1, 1)) / 2
    assert abs(np.sqrt(np.mean(gx ** 2 + gy ** 2)) - 0.1) < 1e-12
    q, C = radial_psd(h)
    qs = 2 * np.pi / 4.0
    assert C[q > qs * 1.0001].max() < 1e-20 * C.max()
    sel = (q > 4 * 2 * np.pi / n) & (q < qs / 2)
    slope = np.polyfit(np.log(q[sel]), np.log(C[sel]), 1)[0]
    assert abs(slope + 2 * (1 + H)) < 0.1
