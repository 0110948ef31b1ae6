"""Pins the oracle's pose / translation enumeration (src/sampling.cpp) and the product's host port of it.

Checks: the reference's own HEALPix (oracle/_ref, compiled from Healpix_2.15a), the HEALPix NEST
closed forms (Gorski et al. 2005: 12*4^order pixels, children of pixel p are 4p..4p+3), lattice
point counts of the translation disc, the ZYZ Euler matrix as a product of textbook axis rotations,
and the perturbation identity A(perturbed) = A(pose) * R(p, p, p).
"""
import numpy as np
import pytest

import oracle as O
import rome_amd


def rz(a):
    c, s = np.cos(np.deg2rad(a)), np.sin(np.deg2rad(a))
    return np.array([[c, s, 0], [-s, c, 0], [0, 0, 1.0]])


def ry(b):
    c, s = np.cos(np.deg2rad(b)), np.sin(np.deg2rad(b))
    return np.array([[c, 0, -s], [0, 1.0, 0], [s, 0, c]])


@pytest.mark.parametrize("order", range(0, 6))
def test_healpix_matches_reference_build(order):
    ref = O.ref_healpix()
    if ref is None:
        pytest.skip("reference HEALPix (oracle/_ref) not built")
    z, phi = O.pix2ang(order)
    n = 12 * 4 ** order
    zr, pr = np.zeros(n), np.zeros(n)
    assert ref.ref_healpix_pix2ang_z_phi(order, O.ptr(zr), O.ptr(pr)) == n
    assert np.array_equal(z, zr) and np.array_equal(phi, pr)
    rng = np.random.default_rng(order)
    import ctypes as C
    for pix in rng.integers(0, n, 50):
        ix, iy, f = C.c_int(), C.c_int(), C.c_int()
        ref.ref_healpix_nest2xyf(order, int(pix), C.byref(ix), C.byref(iy), C.byref(f))
        assert O.nest2xyf(order, int(pix)) == (ix.value, iy.value, f.value)
        assert O.xyf2nest(order, ix.value, iy.value, f.value) == pix


@pytest.mark.parametrize("order", range(0, 5))
def test_healpix_closed_forms(order):
    z, phi = O.pix2ang(order)
    assert len(z) == 12 * 4 ** order
    assert np.all(np.abs(z) < 1) and np.all((phi >= 0) & (phi < 2 * np.pi))
    # equal-area pixels: z is uniform on [-1, 1] -> mean 0, second moment 1/3 over the sphere
    assert abs(z.mean()) < 1e-12
    # NEST hierarchy: children of pixel p at order+1 are 4p .. 4p+3
    for p in range(0, len(z), max(1, len(z) // 17)):
        ix, iy, f = O.nest2xyf(order, p)
        kids = sorted(O.xyf2nest(order + 1, i, j, f) for j in (2 * iy, 2 * iy + 1) for i in (2 * ix, 2 * ix + 1))
        assert kids == [4 * p, 4 * p + 1, 4 * p + 2, 4 * p + 3]


@pytest.mark.parametrize("rng_,step,count", [(3, 1, 29), (4, 1, 49), (5, 1, 81), (4, 0.5, 197), (10, 2, 81)])
def test_translation_disc_counts(rng_, step, count):
    tx, ty = O.translations(rng_, step)
    assert len(tx) == count
    # brute-force lattice count inside the disc
    g = np.arange(-int(np.ceil(rng_ / step)), int(np.ceil(rng_ / step)) + 1) * step
    X, Y = np.meshgrid(g, g, indexing="ij")
    assert np.sum(X * X + Y * Y <= rng_ * rng_) == count
    assert np.all(tx * tx + ty * ty <= rng_ * rng_)


def test_euler_is_zyz_product():
    rng = np.random.default_rng(1)
    for _ in range(20):
        a, b, g = rng.uniform(-180, 180, 3)
        A = O.euler_matrix(a, b, g)
        np.testing.assert_allclose(A, rz(g) @ ry(b) @ rz(a), atol=1e-14)
        np.testing.assert_allclose(A @ A.T, np.eye(3), atol=1e-14)
        # angles -> matrix -> angles -> matrix round trip
        A2 = O.euler_matrix(*O.euler_angles(A))
        np.testing.assert_allclose(A2, A, atol=1e-12)


def test_pose_tables_geometry():
    order, p = 1, 0.37
    ang = 360.0 / (6 * 2 ** order)
    c0 = O.pose_table3d(order, 0, 0.0)
    c1 = O.pose_table3d(order, 0, p)
    z, phi = O.pix2ang(order)
    # unperturbed directions are the pixel centres, psi = i * step
    for d in range(0, len(z), 7):
        v = np.array([np.sin(np.arccos(z[d])) * np.cos(phi[d]), np.sin(np.arccos(z[d])) * np.sin(phi[d]), z[d]])
        A = O.euler_matrix(*c0[d, 3])
        np.testing.assert_allclose(A.T @ np.array([0, 0, 1.0]), v, atol=1e-12)
        np.testing.assert_allclose(c0[d, :, 2], np.arange(c0.shape[1]) * ang)
    # perturbation: A(perturbed) = A(pose) R(p*ang, p*ang, p*ang)
    R = O.euler_matrix(p * ang, p * ang, p * ang)
    for d in range(0, len(z), 5):
        for s in (0, 3):
            np.testing.assert_allclose(O.euler_matrix(*c1[d, s]), O.euler_matrix(*c0[d, s]) @ R, atol=1e-12)
    # fine grid: 4 children directions x psi +- step/4
    f0 = O.pose_table3d(order, 1, 0.0)
    zf, phif = O.pix2ang(order + 1)
    for d in range(0, len(z), 9):
        dirs = set()
        for io in range(8):
            rot, tilt, psi = f0[d, 2, io]
            assert abs(psi - (2 * ang + (-0.25 if io % 2 == 0 else 0.25) * ang)) < 1e-12
            dirs.add((round(rot, 9), round(tilt, 9)))
        want = set()
        for q in range(4 * d, 4 * d + 4):
            r, t = np.rad2deg(phif[q]), np.rad2deg(np.arccos(zf[q]))
            r = r - 360 if r > 180 else r
            want.add((round(r, 9), round(t, 9)))
        assert dirs == want


def test_shift_tables():
    x0, y0 = O.shift_table(4, 1, 0, 0.0)
    x1, y1 = O.shift_table(4, 1, 1, 0.0)
    xp, yp = O.shift_table(4, 1, 1, 0.3)
    tx, ty = O.translations(4, 1)
    np.testing.assert_array_equal(x0[:, 0], tx)
    np.testing.assert_allclose(x1 - tx[:, None], np.tile([-0.25, 0.25, -0.25, 0.25], (len(tx), 1)))
    np.testing.assert_allclose(y1 - ty[:, None], np.tile([-0.25, -0.25, 0.25, 0.25], (len(tx), 1)))
    np.testing.assert_allclose(xp - x1, 0.3)
    np.testing.assert_allclose(yp - y1, 0.3)


def test_product_sampling_is_bit_identical_to_oracle():
    for order in (1, 2):
        for os_ in (0, 1):
            for p in (0.0, 0.25, -0.41):
                a = rome_amd.pose_table3d(order, os_, p)
                b = O.pose_table3d(order, os_, p)
                assert np.array_equal(a, b), (order, os_, p)
    for os_ in (0, 1):
        assert np.array_equal(rome_amd.pose_table2d(10.0, os_, 0.3), O.pose_table2d(10.0, os_, 0.3))
        assert np.array_equal(np.stack(rome_amd.shift_table(5, 1, os_, 0.2)),
                              np.stack(O.shift_table(5, 1, os_, 0.2)))
    for order in range(0, 5):
        z, phi = O.pix2ang(order)
        for p in range(0, len(z), 11):
            assert rome_amd.healpix_pix2ang(order, p) == (z[p], phi[p])
