"""Pins of the image-preparation oracle (oracle/imageprep.py) by closed forms and identities."""
import math

import numpy as np

from oracle import imageprep as IP


def test_translate_moves_a_delta_and_zero_fills():
    a = np.zeros((16, 16))
    a[5, 7] = 3.0
    b = IP.translate_int(a, 2, -3)            # out(y, x) = in(y - oy, x - ox)
    assert b[2, 9] == 3.0 and b.sum() == 3.0
    c = IP.translate_int(np.ones((8, 8)), 3, 0)
    assert np.all(c[:, :3] == 0) and np.all(c[:, 3:] == 1)


def test_centred_delta_has_a_flat_transform():
    n = 32
    a = np.zeros((n, n))
    a[n // 2, n // 2] = 1.0                   # the centre of ROME's real-space convention
    F = IP.fft_half(IP.center_fft(a, True))
    np.testing.assert_allclose(F, np.full(F.shape, 1.0 / (n * n)), atol=1e-18)


def test_shifted_delta_is_a_phase_ramp():
    n = 32
    a = np.zeros((n, n))
    a[n // 2 + 1, n // 2 + 2] = 1.0           # one pixel down, two right of the centre
    F = IP.fft_half(IP.center_fft(a, True))
    ky = np.fft.fftfreq(n)[:, None]
    kx = np.arange(n // 2 + 1)[None, :] / n
    np.testing.assert_allclose(F, np.exp(-2j * np.pi * (ky * 1 + kx * 2)) / n ** 2, atol=1e-15)


def test_window_shrink_then_grow_keeps_a_band_limited_image():
    n, m = 48, 32
    rng = np.random.default_rng(1)
    F = np.zeros((n, n // 2 + 1), complex)
    for _ in range(20):                        # random coefficients well inside radius m/2 - 1
        ky, kx = rng.integers(-10, 11), rng.integers(0, 11)
        F[ky % n, kx] += rng.normal() + 1j * rng.normal()
    F[0, 0] = F[0, 0].real
    G = IP.window_ft(IP.window_ft(F, n, m), m, n)
    np.testing.assert_allclose(G, F, atol=1e-15)


def test_soft_mask_of_a_constant_image_is_constant_and_noise_fills_outside():
    n = 32
    c = np.full((n, n), 2.5)
    np.testing.assert_allclose(IP.soft_mask_outside(c, 10.0, 3.0), c, atol=1e-14)
    rng = np.random.default_rng(2)
    img = rng.normal(size=(n, n))
    noise = rng.normal(size=(n, n))
    out = IP.soft_mask_outside(img, 8.0, 3.0, noise)
    ii, jj = np.meshgrid(np.arange(n) - n // 2, np.arange(n) - n // 2, indexing="ij")
    r = np.sqrt(ii ** 2 + jj ** 2)
    assert np.array_equal(out[r < 8.0], img[r < 8.0])
    assert np.array_equal(out[r > 11.0], noise[r > 11.0])
    mid = (r > 8.0) & (r < 11.0)
    lo, hi = np.minimum(img, noise)[mid], np.maximum(img, noise)[mid]
    assert np.all(out[mid] >= lo - 1e-12) and np.all(out[mid] <= hi + 1e-12)


def test_electron_wavelength_and_ctf_closed_form():
    # published relativistic electron wavelengths: 0.025079 A at 200 kV, 0.019687 A at 300 kV
    for kv, lam in ((200.0, 0.025079), (300.0, 0.019687)):
        lk = kv * 1e3
        assert abs(12.2643247 / math.sqrt(lk * (1 + lk * 0.978466e-6)) - lam) < 2e-6
    # Cs = 0, Q0 = 0, no astigmatism or B-factor: CTF(u) = sin(pi * lambda * defocus * u^2)
    dim, ori, apix, df = 64, 64, 1.5, 15000.0
    C = IP.ctf_image(dim, ori, apix, df, df, 0.0, 300.0, 0.0, 0.0, 0.0)
    lam = 0.0196875
    for (i, j) in ((0, 5), (3, 4), (dim - 7, 2), (10, 0)):
        ip = i if i < dim // 2 + 1 else i - dim
        u2 = (j / (ori * apix)) ** 2 + (ip / (ori * apix)) ** 2
        assert abs(C[i, j] - math.sin(math.pi * lam * df * u2)) < 2e-4
    # Friedel symmetry of the half-plane's x = 0 column, and the phase-flip / abs modes
    D = IP.ctf_image(dim, ori, apix, 12000.0, 18000.0, 30.0, 300.0, 2.7, 0.1, 50.0)
    np.testing.assert_allclose(D[1:dim // 2, 0], D[dim - 1:dim // 2:-1, 0], atol=1e-14)
    A = IP.ctf_image(dim, ori, apix, 12000.0, 18000.0, 30.0, 300.0, 2.7, 0.1, 50.0, do_abs=True)
    S = IP.ctf_image(dim, ori, apix, 12000.0, 18000.0, 30.0, 300.0, 2.7, 0.1, 50.0, only_flip_phases=True)
    np.testing.assert_allclose(A, np.abs(D), atol=1e-15)
    np.testing.assert_allclose(S, np.where(D < 0, -1.0, 1.0))


def test_power_spectrum_of_a_cosine_wave():
    n, ky, kx = 32, 3, 4
    y, x = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    img = np.cos(2 * np.pi * (ky * y + kx * x) / n)
    p = dict(ori=n, current=16, pixel_size=1.0, particle_diameter=4 * n, width_mask_edge=3.0,
             do_norm_correction=False, avg_norm_correction=1.0, do_ctf_correction=False, ctf_phase_flipped=False,
             only_flip_phases=False, intact_ctf_first_peak=False)
    r = IP.set_fourier_transforms(img[None], np.ones(1), np.zeros(1), np.zeros(1), None, p)
    spec = r["power"][0]
    # one of the two +-k coefficients (1/2 each) lies in the half-plane: 1/4 at radius 5
    assert abs(spec[5] - 0.25) < 1e-12 and abs(spec.sum() - 0.25) < 1e-12
    assert abs(r["highres_Xi2"][0] - spec[9:].sum()) < 1e-15
    assert r["highres_Xi2"][0] < 1e-28      # radius 5 < current_size / 2 + 1: rounding noise only


def test_pipeline_composition_without_mask_effects():
    n = 32
    rng = np.random.default_rng(3)
    img = rng.normal(size=(n, n))
    p = dict(ori=n, current=n, pixel_size=1.0, particle_diameter=4 * n, width_mask_edge=3.0,
             do_norm_correction=True, avg_norm_correction=0.8, do_ctf_correction=False, ctf_phase_flipped=False,
             only_flip_phases=False, intact_ctf_first_peak=False)
    r = IP.set_fourier_transforms(img[None], np.array([0.5]), np.array([1.6]), np.array([-2.5]), None, p)
    t = IP.translate_int(img * 0.8 / 0.5, 2, -3)            # C round: 1.6 -> 2, -2.5 -> -3
    np.testing.assert_allclose(r["fnomask"][0], np.fft.rfft2(np.roll(t, (n // 2, n // 2), (0, 1))) / n ** 2,
                               atol=1e-14)
    assert r["old_offx"][0] == 2.0 and r["old_offy"][0] == -3.0
    np.testing.assert_array_equal(r["fctf"][0], np.ones((n, n // 2 + 1)))
