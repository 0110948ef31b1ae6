"""GPU parity of the image preparation (rome_prepare_images) against the FP64 oracle
(oracle/imageprep.py): both sides FP64 (hipFFT vs pocketfft), so the transforms, spectra and CTF
images must agree to 1e-12 of their largest magnitude; rounded offsets exactly.  Cases: norm
correction, fractional offsets (C rounding), background-average fill (-zero_mask) and a given noise
image, current_size < ori_size (power spectrum, highres_Xi2), the three CTF modes.  Then the
prepared pool through the E-step: host and device outputs give the same results."""
import numpy as np
import pytest

import oracle.imageprep as IP
from rome_amd import synth

from test_gpu_parity import _case

pytestmark = pytest.mark.gpu


def _cfg(ori, cur, do_norm, do_ctf):
    return dict(ref_dim=2, ori_size=ori, current_size=cur, coarse_size=min(cur, 24), nr_classes=1, nr_groups=1,
                adaptive_oversampling=0, nr_dir=1, nr_psi=36, nr_trans=9, adaptive_fraction=0.999, sigma2_fudge=1.0,
                sigma2_offset=4.0, avg_norm_correction=0.8, do_ctf_correction=int(do_ctf), refs_are_ctf_corrected=1,
                do_norm_correction=int(do_norm), do_map=1, precision=0)


def _inputs(n, ori, seed):
    rng = np.random.default_rng(seed)
    imgs = rng.normal(size=(n, ori, ori))
    yy, xx = np.meshgrid(np.arange(ori) - ori // 2, np.arange(ori) - ori // 2, indexing="ij")
    imgs += 3.0 * np.exp(-(xx ** 2 + yy ** 2) / (2 * (ori / 8) ** 2))   # a particle-like blob
    norm = rng.uniform(0.6, 1.4, n)
    xoff = rng.uniform(-4.5, 4.5, n)
    yoff = rng.uniform(-4.5, 4.5, n)
    xoff[0], yoff[0] = 2.5, -2.5                                          # halves round away from zero
    ctf = np.column_stack([rng.uniform(8000, 25000, n), rng.uniform(8000, 25000, n), rng.uniform(0, 180, n),
                           np.full(n, 300.0), np.full(n, 2.7), np.full(n, 0.1), rng.uniform(0, 80, n)])
    return imgs, norm, xoff, yoff, ctf


CASES = [
    # ori, cur, do_norm, do_ctf, noise, phase_flipped, only_flip, intact
    (64, 64, True, True, False, False, False, False),
    (64, 40, True, True, True, True, False, True),
    (96, 64, False, True, False, False, True, False),
    (128, 128, True, False, True, False, False, False),
]


@pytest.mark.parametrize("case", CASES, ids=[f"ori{c[0]}_cur{c[1]}_{i}" for i, c in enumerate(CASES)])
def test_prepare_images_matches_oracle(engine, case):
    ori, cur, do_norm, do_ctf, with_noise, flip, only, intact = case
    n = 5
    print(f"[imageprep case ori={ori} cur={cur}]", flush=True)
    imgs, norm, xoff, yoff, ctf = _inputs(n, ori, ori + cur)
    noise = np.random.default_rng(7).normal(size=(n, ori, ori)) if with_noise else None
    engine.configure(**_cfg(ori, cur, do_norm, do_ctf))
    apix, diam, width = 1.5, 0.7 * ori * 1.5, 5.0
    got = engine.prepare_images(imgs, xoff, yoff, norm=norm, ctf=ctf if do_ctf else None, noise=noise,
                                pixel_size=apix, particle_diameter=diam, width_mask_edge=width,
                                ctf_phase_flipped=flip, only_flip_phases=only, intact_ctf_first_peak=intact)
    p = dict(ori=ori, current=cur, pixel_size=apix, particle_diameter=diam, width_mask_edge=width,
             do_norm_correction=do_norm, avg_norm_correction=0.8, do_ctf_correction=do_ctf, ctf_phase_flipped=flip,
             only_flip_phases=only, intact_ctf_first_peak=intact)
    ref = IP.set_fourier_transforms(imgs, norm, xoff, yoff, ctf, p, noise)
    cF = cur * (cur // 2 + 1)
    for key, part in (("fnomask", "fnomask"), ("fmask", "fmask")):
        o = ref[key].reshape(n, cF)
        g = got[part + "_re"] + 1j * got[part + "_im"]
        assert np.max(np.abs(g - o)) <= 1e-12 * np.max(np.abs(o)), key
    np.testing.assert_allclose(got["fctf"], ref["fctf"].reshape(n, cF), rtol=0, atol=1e-12)
    assert np.array_equal(got["old_offx"], ref["old_offx"]) and np.array_equal(got["old_offy"], ref["old_offy"])
    assert got["old_offx"][0] == 3.0 and got["old_offy"][0] == -3.0
    if cur < ori:
        np.testing.assert_allclose(got["power_img"], ref["power"], rtol=1e-11, atol=1e-15 * ref["power"].max())
        np.testing.assert_allclose(got["highres_Xi2"], ref["highres_Xi2"], rtol=1e-11)
        assert np.all(got["highres_Xi2"] > 0)
    else:
        assert np.all(got["power_img"] == 0) and np.all(got["highres_Xi2"] == 0)


class _Dev:
    """Device buffers from the HIP runtime librome_gpu.so itself links (plain hipMalloc / hipMemcpy)."""

    def __init__(self):
        import ctypes as C
        self.C = C
        self.hip = C.CDLL("libamdhip64.so.7")
        self.bufs = []

    def alloc(self, nbytes):
        p = self.C.c_void_p()
        assert self.hip.hipMalloc(self.C.byref(p), self.C.c_size_t(nbytes)) == 0
        self.bufs.append(p)
        return p.value

    def upload(self, a):
        a = np.ascontiguousarray(a)
        d = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(self.C.c_void_p(d), a.ctypes.data_as(self.C.c_void_p), self.C.c_size_t(a.nbytes),
                                  1) == 0
        return d

    def download(self, d, shape, dtype):
        a = np.empty(shape, dtype)
        assert self.hip.hipMemcpy(a.ctypes.data_as(self.C.c_void_p), self.C.c_void_p(d), self.C.c_size_t(a.nbytes),
                                  2) == 0
        return a

    def free(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def test_prepared_pool_through_the_estep(engine):
    """Real-space particles of a small 3D problem, prepared on the GPU (host and device outputs), then
    the E-step on each: identical results; and the same E-step from the oracle-prepared inputs."""
    prob, batch, _ = _case("3d")
    n, ori = batch["n"], prob["ori_size"]
    cur = prob["current_size"]
    cF = cur * (cur // 2 + 1)
    # real-space particles: inverse transforms of the problem's own (unshifted) images, re-centred
    F = (batch["fmask_re"] + 1j * batch["fmask_im"]).reshape(n, cur, cur // 2 + 1) * ori * ori
    imgs = np.stack([IP.center_fft(np.fft.irfft2(F[i], s=(ori, ori)), forward=False) for i in range(n)])
    rng = np.random.default_rng(5)
    xoff, yoff = rng.uniform(-1.4, 1.4, n), rng.uniform(-1.4, 1.4, n)
    ctf = np.column_stack([np.full(n, 15000.0), np.full(n, 14000.0), np.full(n, 20.0), np.full(n, 300.0),
                           np.full(n, 2.7), np.full(n, 0.1), np.zeros(n)])
    synth.load_engine(engine, prob)
    kw = dict(norm=batch["norm"], ctf=ctf if prob["do_ctf_correction"] else None, pixel_size=3.0,
              particle_diameter=0.8 * ori * 3.0, width_mask_edge=3.0)
    D = _Dev()
    dev = {k: D.alloc(n * cF * 8) for k in ("fmask", "fnomask")}
    dev["fctf"] = D.alloc(n * cF * 4)
    for k, m in (("highres_Xi2", n), ("power_img", n * (ori // 2 + 1)), ("old_offx", n), ("old_offy", n)):
        dev[k] = D.alloc(m * 8)
    got = engine.prepare_images(imgs, xoff, yoff, device_out=dev, **kw)
    fm32 = D.download(dev["fmask"], (n, cF), np.complex64)
    assert np.max(np.abs(fm32 - (got["fmask_re"] + 1j * got["fmask_im"]))) <= 1e-6 * np.max(np.abs(fm32))
    assert np.array_equal(D.download(dev["old_offx"], (n,), np.float64), got["old_offx"])
    b = dict(batch)
    for k in ("fmask_re", "fmask_im", "fnomask_re", "fnomask_im", "fctf", "highres_Xi2", "old_offx", "old_offy"):
        b[k] = got[k]
    b["power_img"] = got["power_img"]
    engine.reset_sums()
    r_host = engine.expectation(b)
    s_host = engine.get_sums()
    engine.reset_sums()
    norm_d = D.upload(np.asarray(batch["norm"], np.float64))
    group_d = D.upload(np.asarray(batch["group"], np.int32))
    r_dev = engine.expectation_device(n, dev["fmask"], dev["fnomask"], dev["fctf"], dev["highres_Xi2"],
                                      dev["power_img"], dev["old_offx"], dev["old_offy"], norm_d, group_d,
                                      results=True)
    s_dev = engine.get_sums()
    D.free()
    # the host path converts the FP64 outputs to FP32 exactly as the device outputs were rounded
    assert np.array_equal(r_host["argmax"], r_dev["argmax"])
    np.testing.assert_allclose(r_host["meta"], r_dev["meta"], rtol=1e-9, atol=1e-12)
    for k in s_host:
        np.testing.assert_allclose(s_host[k], s_dev[k], rtol=1e-9, atol=1e-12, err_msg=k)
    # the oracle's preparation of the same pool: same poses
    p = dict(ori=ori, current=cur, pixel_size=3.0, particle_diameter=0.8 * ori * 3.0, width_mask_edge=3.0,
             do_norm_correction=bool(prob["do_norm_correction"]), avg_norm_correction=prob["avg_norm_correction"],
             do_ctf_correction=bool(prob["do_ctf_correction"]), ctf_phase_flipped=False, only_flip_phases=False,
             intact_ctf_first_peak=False)
    ref = IP.set_fourier_transforms(imgs, batch["norm"], xoff, yoff, ctf, p)
    b2 = dict(batch)
    b2.update(fmask_re=ref["fmask"].real.reshape(n, cF), fmask_im=ref["fmask"].imag.reshape(n, cF),
              fnomask_re=ref["fnomask"].real.reshape(n, cF), fnomask_im=ref["fnomask"].imag.reshape(n, cF),
              fctf=ref["fctf"].reshape(n, cF), highres_Xi2=ref["highres_Xi2"], power_img=ref["power"],
              old_offx=ref["old_offx"], old_offy=ref["old_offy"])
    engine.reset_sums()
    r_or = engine.expectation(b2)
    assert np.array_equal(r_or["argmax"], r_host["argmax"])
    np.testing.assert_allclose(r_or["meta"], r_host["meta"], rtol=1e-6, atol=1e-9)
