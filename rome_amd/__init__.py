"""rome_amd -- MI355X-native MAP E-step + Fourier backprojection engine.

Python binding of librome_gpu.so (C-ABI declared in include/rome_gpu.h).  The HIP
library is the only implementation: importing works without it, but every call
raises if the library was not built (``python -m rome_amd.build``).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librome_gpu.so")

_lib = None
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_longlong)

META_FIELDS = ("PMAX", "CLASS", "ROT", "TILT", "PSI", "XOFF", "YOFF", "NORM", "DLL")

EXPORTED = (
    "rome_create", "rome_destroy", "rome_last_error", "rome_configure", "rome_set_sampling", "rome_set_model",
    "rome_set_reference", "rome_reset_sums", "rome_expectation", "rome_expectation_device", "rome_get_sums",
    "rome_get_backprojector", "rome_get_coarse_diff2", "rome_get_fine_count", "rome_get_fine_list", "rome_project",
    "rome_sampling_pose3d", "rome_sampling_pose2d", "rome_sampling_translations", "rome_sampling_shifts",
    "rome_healpix_pix2ang", "rome_project_device", "rome_set_profiling", "rome_get_profile",
    "rome_get_batch_stats", "rome_maximization_other", "rome_update_model", "rome_fourier_map",
    "rome_reconstruct", "rome_set_coarse_mode", "rome_get_coarse_stats", "rome_prepare_images",
    "rome_set_significance_cap", "rome_get_significance_stats", "rome_proposal_configure", "rome_proposal_mode",
    "rome_proposal_train", "rome_proposal_stats", "rome_proposal_get_params", "rome_proposal_set_params",
    "rome_proposal_forward", "rome_proposal_gradients", "rome_proposal_training_rows", "rome_backproject_images",
    "rome_project_device_noalloc", "rome_set_execution", "rome_get_execution_stats", "rome_expectation_fine_device",
    "rome_autotune", "rome_set_persistent_grids", "rome_proposal_set_thresholds", "rome_set_class_range",
)

EXEC_HOST, EXEC_DEVICE, EXEC_GRAPH = 0, 1, 2

COARSE_LIST, COARSE_FULL = 0, 1

STAGES = ("prep", "coarse_norms", "coarse_cross", "coarse_weights", "task_build", "fine_project", "fine_diff2",
          "fine_weights", "post", "store", "backproject", "coarse_recompute", "proposal")


class RomeError(RuntimeError):
    pass


class rome_config(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("ref_dim", "ori_size", "current_size", "coarse_size", "nr_classes",
                                        "nr_groups", "adaptive_oversampling", "nr_dir", "nr_psi", "nr_trans")] + \
               [(n, C.c_double) for n in ("adaptive_fraction", "sigma2_fudge", "sigma2_offset",
                                           "avg_norm_correction")] + \
               [(n, C.c_int) for n in ("do_ctf_correction", "refs_are_ctf_corrected", "do_norm_correction",
                                        "do_map", "precision")]

PRECISIONS = {"fp32": 0, "bf16": 1, "fp16": 2}


class rome_images(C.Structure):
    _fields_ = [("n", C.c_int)] + [(n, _dp) for n in (
        "fmask_re", "fmask_im", "fnomask_re", "fnomask_im", "fctf", "highres_Xi2", "power_img", "old_offx",
        "old_offy", "norm")] + [("group", _ip)]


class rome_device_images(C.Structure):
    _fields_ = [("n", C.c_int), ("fmask", C.c_void_p), ("fnomask", C.c_void_p), ("fctf", C.c_void_p),
                ("highres_Xi2", C.c_void_p), ("power_img", C.c_void_p), ("old_offx", C.c_void_p),
                ("old_offy", C.c_void_p), ("norm", C.c_void_p), ("group", C.c_void_p)]


class rome_results(C.Structure):
    _fields_ = [("meta", _dp), ("nr_sign", _ip), ("argmax", _ip)] + [(n, _dp) for n in (
        "min_diff2", "sum_weight", "sig_weight", "min_diff2_coarse", "sum_weight_coarse", "sig_weight_coarse",
        "wsum_norm_correction")]


class rome_sums(C.Structure):
    _fields_ = [(n, _dp) for n in ("wsum_sigma2_noise", "wsum_sumw_group", "wsum_pdf_class", "wsum_pdf_direction",
                                   "wsum_prior_offx", "wsum_prior_offy", "wsum_scalars")]


class rome_model(C.Structure):
    _fields_ = [(n, _dp) for n in ("sigma2_noise", "pdf_class", "pdf_direction", "prior_offx", "prior_offy",
                                   "scalars")]


def _model_arrays(K, G, D, F):
    return {"sigma2_noise": np.zeros((G, F)), "pdf_class": np.zeros(K), "pdf_direction": np.zeros((K, D)),
            "prior_offx": np.zeros(K), "prior_offy": np.zeros(K), "scalars": np.zeros(4)}


def maximization_other(ref_dim, ori_size, sums, nr_dir, do_norm_correction, avg_norm_in, prior_offx, prior_offy):
    """rome_maximization_other (host FP64): the model update of maximizationOtherParameters from E-step
    sums (same dict layout as Engine.get_sums).  scalars = sigma2_offset, avg_norm_correction, LL, ave_Pmax."""
    K = len(sums["wsum_pdf_class"])
    G, F = np.asarray(sums["wsum_sigma2_noise"]).shape
    keep = {k: _f64(v) for k, v in sums.items()}
    cs = rome_sums(**{k: _p(v) for k, v in keep.items()})
    out = _model_arrays(K, G, nr_dir, F)
    cm = rome_model(**{k: _p(v) for k, v in out.items()})
    px, py = _f64(prior_offx), _f64(prior_offy)
    rc = lib().rome_maximization_other(ref_dim, ori_size, K, G, nr_dir, int(do_norm_correction),
                                       C.c_double(avg_norm_in), _p(px), _p(py), C.byref(cs), C.byref(cm))
    if rc != 0:
        raise RomeError(f"rome_maximization_other failed ({rc})")
    return out


class rome_recon_params(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("max_iter_preweight", "do_map", "minres_map", "update_tau2_with_fsc",
                                        "is_whole_instead_of_half")] + \
               [(n, C.c_double) for n in ("tau2_fudge", "normalise")]


class rome_raw_images(C.Structure):
    _fields_ = [("n", C.c_int), ("images", _dp), ("norm", _dp), ("xoff", _dp), ("yoff", _dp), ("ctf", _dp),
                ("noise", _dp)]


class rome_prep_params(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("pixel_size", "particle_diameter", "width_mask_edge")] + \
               [(n, C.c_int) for n in ("ctf_phase_flipped", "only_flip_phases", "intact_ctf_first_peak")]


class rome_prepared(C.Structure):
    _fields_ = [(n, _dp) for n in ("fmask_re", "fmask_im", "fnomask_re", "fnomask_im", "fctf", "highres_Xi2",
                                   "power_img", "old_offx", "old_offy")]


class rome_prepared_device(C.Structure):
    _fields_ = [("fmask", C.c_void_p), ("fnomask", C.c_void_p), ("fctf", C.c_void_p), ("highres_Xi2", C.c_void_p),
                ("power_img", C.c_void_p), ("old_offx", C.c_void_p), ("old_offy", C.c_void_p)]


class rome_proposal_config(C.Structure):
    _fields_ = [("hidden", C.c_int), ("nr_coef", C.c_int), ("nr_psi_bins", C.c_int), ("top_k", C.c_int),
                ("min_mass", C.c_double), ("audit_fraction", C.c_double), ("min_audit_mass", C.c_double),
                ("target_cells", C.c_int), ("shift_weight", C.c_double), ("max_train", C.c_int),
                ("seed", C.c_ulonglong)]


PROPOSAL_PARAMS = ("W1", "b1", "W2", "b2", "W3", "b3", "W4", "b4")


class rome_proposal_params(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_float)) for n in PROPOSAL_PARAMS]


PROPOSAL_MODES = {"off": 0, "collect": 1, "use": 2}


def lib():
    """Load librome_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RomeError(f"{LIB_PATH} not found: build it with `python -m rome_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.rome_last_error.restype = C.c_char_p
        L.rome_get_fine_count.restype = C.c_longlong
        L.rome_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
        for name in EXPORTED:
            getattr(L, name)  # every declared symbol must resolve
        _lib = L
    return _lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a, t=C.c_double):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


# ---------------- host sampling (HealpixSampler, C1) ----------------
def pose_table3d(order, oversampling, perturbation):
    nd, npsi = 12 * 4 ** order, 6 * 2 ** order
    nover = 8 if oversampling else 1
    out = np.zeros((nd, npsi, nover, 3))
    if lib().rome_sampling_pose3d(order, oversampling, C.c_double(perturbation), _p(out)) < 0:
        raise RomeError("rome_sampling_pose3d failed")
    return out if oversampling else out[:, :, 0, :]


def pose_table2d(psi_step, oversampling, perturbation):
    npsi = int(np.ceil(360.0 / psi_step))
    nover = 2 if oversampling else 1
    out = np.zeros((npsi, nover))
    if lib().rome_sampling_pose2d(C.c_double(psi_step), oversampling, C.c_double(perturbation), _p(out)) < 0:
        raise RomeError("rome_sampling_pose2d failed")
    return out if oversampling else out[:, 0]


def translations(offset_range, offset_step):
    n = lib().rome_sampling_translations(C.c_double(offset_range), C.c_double(offset_step), None, None)
    tx, ty = np.zeros(n), np.zeros(n)
    lib().rome_sampling_translations(C.c_double(offset_range), C.c_double(offset_step), _p(tx), _p(ty))
    return tx, ty


def shift_table(offset_range, offset_step, oversampling, perturbation):
    T = len(translations(offset_range, offset_step)[0])
    nover = 4 if oversampling else 1
    x, y = np.zeros((T, nover)), np.zeros((T, nover))
    lib().rome_sampling_shifts(C.c_double(offset_range), C.c_double(offset_step), oversampling,
                               C.c_double(perturbation), _p(x), _p(y))
    return x, y


def healpix_pix2ang(order, pix):
    z, phi = C.c_double(), C.c_double()
    if lib().rome_healpix_pix2ang(order, pix, C.byref(z), C.byref(phi)) != 0:
        raise RomeError("bad HEALPix pixel")
    return z.value, phi.value


class Engine:
    """One device context: the E-step of rome_map3d / rome_map2d on one GPU."""

    def __init__(self, device=0):
        self._h = C.c_void_p()
        rc = lib().rome_create(device, C.byref(self._h))
        if rc != 0:
            raise RomeError(f"rome_create failed ({rc}): no usable GPU")
        self.cfg = None

    def close(self):
        if self._h:
            lib().rome_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            raise RomeError(f"{what} failed ({rc}): {lib().rome_last_error(self._h).decode()}")

    # ---- iteration setup
    def configure(self, **kw):
        c = rome_config(**kw)
        self._check(lib().rome_configure(self._h, C.byref(c)), "rome_configure")
        self.cfg = dict(kw)
        self.is3d = c.ref_dim == 3
        self.K, self.D, self.PSI, self.T = c.nr_classes, c.nr_dir, c.nr_psi, c.nr_trans
        self.nor = (8 if self.is3d else 2) if c.adaptive_oversampling else 1
        self.nots = 4 if c.adaptive_oversampling else 1
        r_max = min(c.current_size // 2, c.ori_size // 2)
        self.pad = 2 * (2 * r_max + 1) + 1
        padF = self.pad // 2 + 1
        self.grid_n = self.pad * self.pad * padF if self.is3d else self.pad * padF
        self.oriF = c.ori_size // 2 + 1
        self.G = c.nr_groups
        self.F2 = c.current_size * (c.current_size // 2 + 1)

    def set_sampling(self, pose_coarse, pose_fine, shift_coarse_x, shift_coarse_y, shift_fine_x, shift_fine_y,
                     trans_x, trans_y):
        keep = [None if a is None else _f64(a) for a in (pose_coarse, pose_fine, shift_coarse_x, shift_coarse_y,
                                                         shift_fine_x, shift_fine_y, trans_x, trans_y)]
        self._check(lib().rome_set_sampling(self._h, *[_p(a) for a in keep]), "rome_set_sampling")

    def set_model(self, sigma2_noise, pdf_class, pdf_direction, prior_offx, prior_offy):
        keep = [_f64(a) for a in (sigma2_noise, pdf_class, pdf_direction, prior_offx, prior_offy)]
        self._check(lib().rome_set_model(self._h, *[_p(a) for a in keep]), "rome_set_model")

    def set_reference(self, iclass, grid):
        g = _f64(grid)
        if g.size != 2 * self.grid_n:
            raise RomeError(f"reference grid has {g.size} values, expected {2 * self.grid_n}")
        self._check(lib().rome_set_reference(self._h, iclass, _p(g)), "rome_set_reference")

    def reset_sums(self):
        self._check(lib().rome_reset_sums(self._h), "rome_reset_sums")

    # ---- E-step
    def _results(self, n):
        r = {"meta": np.zeros((n, 9)), "nr_sign": np.zeros(n, np.int32), "argmax": np.zeros((n, 6), np.int32)}
        for k in ("min_diff2", "sum_weight", "sig_weight", "min_diff2_coarse", "sum_weight_coarse",
                  "sig_weight_coarse", "wsum_norm_correction"):
            r[k] = np.zeros(n)
        cr = rome_results(meta=_p(r["meta"]), nr_sign=_p(r["nr_sign"], C.c_int), argmax=_p(r["argmax"], C.c_int),
                          **{k: _p(r[k]) for k in ("min_diff2", "sum_weight", "sig_weight", "min_diff2_coarse",
                                                   "sum_weight_coarse", "sig_weight_coarse",
                                                   "wsum_norm_correction")})
        return r, cr

    def expectation(self, b):
        """b: dict of host arrays (fmask_re, fmask_im, fnomask_re, fnomask_im, fctf, highres_Xi2, power_img,
        old_offx, old_offy, norm, group)."""
        n = int(b["n"])
        keep = {k: _f64(b[k]) for k in ("fmask_re", "fmask_im", "fnomask_re", "fnomask_im", "fctf", "highres_Xi2",
                                        "power_img", "old_offx", "old_offy", "norm")}
        grp = np.ascontiguousarray(b["group"], dtype=np.int32)
        im = rome_images(n=n, group=_p(grp, C.c_int), **{k: _p(v) for k, v in keep.items()})
        r, cr = self._results(n)
        self._check(lib().rome_expectation(self._h, C.byref(im), C.byref(cr)), "rome_expectation")
        return r

    def expectation_device(self, n, fmask, fnomask, fctf, highres_Xi2, power_img, old_offx, old_offy, norm, group,
                           results=True):
        """Device pointers (ints) of FP32 complex / FP32 / FP64 / int32 arrays already in HBM."""
        im = rome_device_images(n=n, fmask=fmask, fnomask=fnomask, fctf=fctf, highres_Xi2=highres_Xi2,
                                power_img=power_img, old_offx=old_offx, old_offy=old_offy, norm=norm, group=group)
        if results:
            r, cr = self._results(n)
            self._check(lib().rome_expectation_device(self._h, C.byref(im), C.byref(cr)), "rome_expectation_device")
            return r
        self._check(lib().rome_expectation_device(self._h, C.byref(im), None), "rome_expectation_device")
        return None

    def expectation_fine_device(self, n, fmask, fnomask, fctf, highres_Xi2, power_img, old_offx, old_offy, norm,
                                group, sig_keys, sig_cnt, cap, results=True):
        """The fine pass of a pool whose coarse significance is given: sig_keys [n][cap] int32 keys
        m * nr_trans + t (m = (class * nr_dir + dir) * nr_psi + psi), sig_cnt [n] int32, device pointers."""
        im = rome_device_images(n=n, fmask=fmask, fnomask=fnomask, fctf=fctf, highres_Xi2=highres_Xi2,
                                power_img=power_img, old_offx=old_offx, old_offy=old_offy, norm=norm, group=group)
        r, cr = self._results(n) if results else (None, None)
        self._check(lib().rome_expectation_fine_device(self._h, C.byref(im), C.c_void_p(sig_keys), C.c_void_p(sig_cnt),
                                                       int(cap), C.byref(cr) if results else None),
                    "rome_expectation_fine_device")
        return r

    def get_sums(self):
        s = {"wsum_sigma2_noise": np.zeros((self.G, self.oriF)), "wsum_sumw_group": np.zeros(self.G),
             "wsum_pdf_class": np.zeros(self.K), "wsum_pdf_direction": np.zeros((self.K, self.D)),
             "wsum_prior_offx": np.zeros(self.K), "wsum_prior_offy": np.zeros(self.K),
             "wsum_scalars": np.zeros(4)}
        cs = rome_sums(**{k: _p(v) for k, v in s.items()})
        self._check(lib().rome_get_sums(self._h, C.byref(cs)), "rome_get_sums")
        return s

    def fourier_map(self, iclass, vol, do_gridding=True, want_grid=True):
        """rome_fourier_map: projector grid of a real map, installed as reference iclass.
        Returns (power_spectrum, corrected_map, grid or None); grid is complex128 [pad][pad][pad/2+1]
        (3D) or [pad][pad/2+1] (2D)."""
        v = _f64(vol)
        ps = np.zeros(self.oriF)
        corr = np.zeros_like(v)
        grid = np.zeros(2 * self.grid_n) if want_grid else None
        self._check(lib().rome_fourier_map(self._h, iclass, _p(v), int(do_gridding), _p(ps), _p(corr), _p(grid)),
                    "rome_fourier_map")
        if grid is not None:
            padF = self.pad // 2 + 1
            shape = (self.pad, self.pad, padF) if self.is3d else (self.pad, padF)
            grid = (grid[0::2] + 1j * grid[1::2]).reshape(shape)
        return ps, corr, grid

    def reconstruct(self, iclass, tau2, data=None, weight=None, fsc=None, max_iter_preweight=10, do_map=True,
                    minres_map=5, update_tau2_with_fsc=False, is_whole_instead_of_half=False, tau2_fudge=1.0,
                    normalise=1.0):
        """rome_reconstruct: the class map from the backprojector grids (the context's own, or the given
        complex data / real weight grids).  Returns (vol, tau2, sigma2, data_vs_prior)."""
        prm = rome_recon_params(max_iter_preweight, int(do_map), minres_map, int(update_tau2_with_fsc),
                                int(is_whole_instead_of_half), tau2_fudge, normalise)
        t2 = _f64(tau2).copy()
        sig = np.zeros(self.oriF)
        dvp = np.zeros(self.oriF)
        ori = self.cfg["ori_size"]
        vol = np.zeros((ori, ori, ori) if self.is3d else (ori, ori))
        d = w = None
        if data is not None:
            d = np.ascontiguousarray(np.stack([np.real(data), np.imag(data)], axis=-1), dtype=np.float64)
            w = _f64(weight)
        f = None if fsc is None else _f64(fsc)
        self._check(lib().rome_reconstruct(self._h, iclass, C.byref(prm), _p(d), _p(w), _p(t2), _p(f), _p(sig),
                                           _p(dvp), _p(vol)), "rome_reconstruct")
        return vol, t2, sig, dvp

    def update_model(self, install=True):
        """rome_update_model: M-step model parameters from this context's sums (optionally installed)."""
        out = _model_arrays(self.K, self.G, self.D, self.oriF)
        cm = rome_model(**{k: _p(v) for k, v in out.items()})
        self._check(lib().rome_update_model(self._h, int(install), C.byref(cm)), "rome_update_model")
        return out

    def get_backprojector(self, iclass):
        d = np.zeros(2 * self.grid_n)
        w = np.zeros(self.grid_n)
        self._check(lib().rome_get_backprojector(self._h, iclass, _p(d), _p(w)), "rome_get_backprojector")
        return d, w

    def get_coarse_diff2(self, n):
        out = np.zeros((n, self.K, self.D, self.PSI, self.T))
        self._check(lib().rome_get_coarse_diff2(self._h, _p(out)), "rome_get_coarse_diff2")
        return out

    def get_fine_list(self):
        cnt = lib().rome_get_fine_count(self._h)
        keys = np.zeros(max(cnt, 1), np.int64)
        d2 = np.zeros(max(cnt, 1))
        self._check(lib().rome_get_fine_list(self._h, _p(keys, C.c_longlong), _p(d2)), "rome_get_fine_list")
        return keys[:cnt], d2[:cnt]

    def project_device(self, iclass, size, A, out_ptr):
        A = _f64(A).reshape(-1, 9)
        self._check(lib().rome_project_device(self._h, iclass, size, A.shape[0], _p(A), C.c_void_p(out_ptr)),
                    "rome_project_device")

    def project_device_noalloc(self, iclass, size, A, out_ptr):
        """rome_project_device without device allocation (rotations staged in the library's fixed array)."""
        A = _f64(A).reshape(-1, 9)
        self._check(lib().rome_project_device_noalloc(self._h, iclass, size, A.shape[0], _p(A), C.c_void_p(out_ptr)),
                    "rome_project_device_noalloc")

    def set_profiling(self, on=True):
        self._check(lib().rome_set_profiling(self._h, int(on)), "rome_set_profiling")

    def profile(self):
        ms = np.zeros(len(STAGES))
        calls = np.zeros(len(STAGES), np.int64)
        lib().rome_get_profile(self._h, _p(ms), _p(calls, C.c_longlong), len(STAGES))
        return {k: (float(ms[i]), int(calls[i])) for i, k in enumerate(STAGES)}

    def batch_stats(self):
        out = np.zeros(8, np.int64)
        self._check(lib().rome_get_batch_stats(self._h, _p(out, C.c_longlong)), "rome_get_batch_stats")
        return dict(zip(("n", "tasks", "blocks", "unique_rot", "Pc", "Pf", "M", "T"), out.tolist()))

    def prepare_images(self, images, xoff, yoff, norm=None, ctf=None, noise=None, pixel_size=1.0,
                       particle_diameter=None, width_mask_edge=5.0, ctf_phase_flipped=False, only_flip_phases=False,
                       intact_ctf_first_peak=False, device_out=None):
        """rome_prepare_images: the E-step inputs of a pool from real-space images [n][ori][ori] (dict of
        FP64 host arrays in the layout `expectation` takes).  device_out: optional dict of device pointers
        (fmask, fnomask, fctf, highres_Xi2, power_img, old_offx, old_offy) to fill as well."""
        img = _f64(images)
        n, ori = img.shape[0], img.shape[1]
        cur = self.cfg["current_size"]
        cF = cur * (cur // 2 + 1)
        out = {k: np.zeros((n, cF)) for k in ("fmask_re", "fmask_im", "fnomask_re", "fnomask_im", "fctf")}
        out.update(highres_Xi2=np.zeros(n), power_img=np.zeros((n, ori // 2 + 1)), old_offx=np.zeros(n),
                   old_offy=np.zeros(n))
        arrs = {"xoff": _f64(xoff), "yoff": _f64(yoff), "norm": None if norm is None else _f64(norm),
                "ctf": None if ctf is None else _f64(ctf), "noise": None if noise is None else _f64(noise)}
        raw = rome_raw_images(n, _p(img), _p(arrs["norm"]), _p(arrs["xoff"]), _p(arrs["yoff"]), _p(arrs["ctf"]),
                              _p(arrs["noise"]))
        if particle_diameter is None:
            particle_diameter = ori * pixel_size
        prm = rome_prep_params(pixel_size, particle_diameter, width_mask_edge, int(ctf_phase_flipped),
                               int(only_flip_phases), int(intact_ctf_first_peak))
        ho = rome_prepared(*[_p(out[k]) for k in ("fmask_re", "fmask_im", "fnomask_re", "fnomask_im", "fctf",
                                                  "highres_Xi2", "power_img", "old_offx", "old_offy")])
        do = None
        if device_out is not None:
            do = rome_prepared_device(*[C.c_void_p(device_out.get(k)) if device_out.get(k) else None
                                        for k in ("fmask", "fnomask", "fctf", "highres_Xi2", "power_img", "old_offx",
                                                  "old_offy")])
        self._check(lib().rome_prepare_images(self._h, C.byref(raw), C.byref(prm), C.byref(ho),
                                              C.byref(do) if do is not None else None), "rome_prepare_images")
        return out

    def backproject_images(self, iclass, images, euler, shift):
        """rome_backproject_images: transforms (dict with fnomask_re / fnomask_im / fctf [n][F2], as
        prepare_images returns them) at poses euler [n][3] and offsets shift [n][2] into class iclass's grid."""
        re_, im_ = _f64(images["fnomask_re"]), _f64(images["fnomask_im"])
        ctf = _f64(images["fctf"]) if images.get("fctf") is not None else None
        e, sh = _f64(euler), _f64(shift)
        self._check(lib().rome_backproject_images(self._h, int(iclass), re_.shape[0], _p(re_), _p(im_), _p(ctf),
                                                  _p(e), _p(sh)), "rome_backproject_images")

    # ---- learned pose proposal
    def proposal_configure(self, hidden=1024, nr_coef=812, nr_psi_bins=4, top_k=64, min_mass=0.9,
                           audit_fraction=0.02, min_audit_mass=0.99, target_cells=32, shift_weight=1.0,
                           max_train=20000, seed=1):
        cfg = rome_proposal_config(hidden, nr_coef, nr_psi_bins, top_k, min_mass, audit_fraction, min_audit_mass,
                                   target_cells, shift_weight, max_train, seed)
        self._check(lib().rome_proposal_configure(self._h, C.byref(cfg)), "rome_proposal_configure")
        self.prop = dict(hidden=hidden, nr_coef=nr_coef, ncell=self.K * 4 * self.D * nr_psi_bins,
                         target_cells=target_cells)
        self.prop_cfg = dict(top_k=top_k, min_mass=min_mass, min_audit_mass=min_audit_mass)

    def proposal_set_gate(self, min_mass, top_k=None, min_audit_mass=None):
        """Per-image predicted-mass gate (and optionally top_k, the audit's captured-mass bar) of the
        configured proposal, keeping the trained network."""
        c = self.prop_cfg
        c["min_mass"] = min_mass
        if top_k is not None:
            c["top_k"] = top_k
        if min_audit_mass is not None:
            c["min_audit_mass"] = min_audit_mass
        self._check(lib().rome_proposal_set_thresholds(self._h, int(c["top_k"]), C.c_double(c["min_mass"]),
                                                       C.c_double(c["min_audit_mass"])),
                    "rome_proposal_set_thresholds")

    def proposal_mode(self, mode):
        self._check(lib().rome_proposal_mode(self._h, PROPOSAL_MODES.get(mode, mode)), "rome_proposal_mode")

    def proposal_train(self, epochs, batch=256, lr=1e-3):
        out = np.zeros(max(epochs, 1))
        self._check(lib().rome_proposal_train(self._h, int(epochs), int(batch), C.c_double(lr), _p(out)),
                    "rome_proposal_train")
        return out[:epochs]

    def proposal_stats(self):
        out = np.zeros(8)
        self._check(lib().rome_proposal_stats(self._h, _p(out)), "rome_proposal_stats")
        return dict(zip(("pools", "images", "cand_rot_per_image", "images_all_rotations", "audited",
                         "min_audit_mass", "fallback_pools", "train_images"), out.tolist()))

    def _prop_shapes(self):
        H, Fin, O = self.prop["hidden"], 2 * self.prop["nr_coef"], self.prop["ncell"] + 2
        return {"W1": (H, Fin), "b1": (H,), "W2": (H, H), "b2": (H,), "W3": (H, H), "b3": (H,), "W4": (O, H),
                "b4": (O,)}

    def _prop_struct(self, arrs):
        return rome_proposal_params(*[arrs[k].ctypes.data_as(C.POINTER(C.c_float)) for k in PROPOSAL_PARAMS])

    def proposal_get_params(self):
        arrs = {k: np.zeros(v, np.float32) for k, v in self._prop_shapes().items()}
        self._check(lib().rome_proposal_get_params(self._h, C.byref(self._prop_struct(arrs))), "rome_proposal_get_params")
        return arrs

    def proposal_set_params(self, params):
        arrs = {k: np.ascontiguousarray(params[k], dtype=np.float32).reshape(v) for k, v in self._prop_shapes().items()}
        self._check(lib().rome_proposal_set_params(self._h, C.byref(self._prop_struct(arrs))), "rome_proposal_set_params")

    def proposal_forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        z = np.zeros((x.shape[0], self.prop["ncell"] + 2), np.float32)
        self._check(lib().rome_proposal_forward(self._h, x.shape[0], x.ctypes.data_as(C.POINTER(C.c_float)),
                                                z.ctypes.data_as(C.POINTER(C.c_float))), "rome_proposal_forward")
        return z

    def proposal_gradients(self, x, target, shift):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        sh = np.ascontiguousarray(shift, dtype=np.float32)
        g = {k: np.zeros(v, np.float32) for k, v in self._prop_shapes().items()}
        loss = C.c_double()
        fp = C.POINTER(C.c_float)
        self._check(lib().rome_proposal_gradients(self._h, x.shape[0], x.ctypes.data_as(fp), t.ctypes.data_as(fp),
                                                  sh.ctypes.data_as(fp), C.byref(loss), C.byref(self._prop_struct(g))),
                    "rome_proposal_gradients")
        return loss.value, g

    def proposal_training_rows(self, row0, n):
        x = np.zeros((n, 2 * self.prop["nr_coef"]), np.float32)
        t = np.zeros((n, self.prop["ncell"]), np.float32)
        sh = np.zeros((n, 2), np.float32)
        fp = C.POINTER(C.c_float)
        self._check(lib().rome_proposal_training_rows(self._h, int(row0), int(n), x.ctypes.data_as(fp),
                                                      t.ctypes.data_as(fp), sh.ctypes.data_as(fp)),
                    "rome_proposal_training_rows")
        return x, t, sh

    def set_coarse_mode(self, mode, list_cap=0):
        """ROME_COARSE_LIST (keep only the entries within reach of the weights) or ROME_COARSE_FULL
        (the whole coarse pose term, needed by get_coarse_diff2); list_cap <= 0: the default cap."""
        m = {"list": COARSE_LIST, "full": COARSE_FULL}.get(mode, mode)
        self._check(lib().rome_set_coarse_mode(self._h, int(m), int(list_cap)), "rome_set_coarse_mode")

    def coarse_stats(self):
        out = np.zeros(3, np.int64)
        self._check(lib().rome_get_coarse_stats(self._h, _p(out, C.c_longlong)), "rome_get_coarse_stats")
        return dict(zip(("mode", "list_max", "list_total"), out.tolist()))

    def set_class_range(self, iclass_min=-1, iclass_max=-1):
        """Classes the next pools run (the seeding iteration's single class per pool); (-1, -1): all."""
        self._check(lib().rome_set_class_range(self._h, int(iclass_min), int(iclass_max)), "rome_set_class_range")

    def set_execution(self, mode):
        """How pools run: "host" (host-synchronised launches), "device" (device-driven fine pass, one
        synchronisation per pool) or "graph" (device-driven, captured once per input set and replayed)."""
        m = {"host": EXEC_HOST, "device": EXEC_DEVICE, "graph": EXEC_GRAPH}.get(mode, mode)
        self._check(lib().rome_set_execution(self._h, int(m)), "rome_set_execution")

    def autotune_device(self, n, fmask, fnomask, fctf, highres_Xi2, power_img, old_offx, old_offy, norm, group,
                        sig_keys=None, sig_cnt=None, cap=0):
        """rome_autotune on a pool of device arrays (with sig_keys / sig_cnt / cap: its fine pass from that
        coarse significance): the persistent grids of the device-driven kernels (fine diff2, fine
        projection, post slices, backprojection) are set to the fastest of 4 candidates each.  Returns
        (chosen grids, stage ms [4][4]).  The runs add to the sums: reset them after."""
        im = rome_device_images(n=n, fmask=fmask, fnomask=fnomask, fctf=fctf, highres_Xi2=highres_Xi2,
                                power_img=power_img, old_offx=old_offx, old_offy=old_offy, norm=norm, group=group)
        chosen = np.zeros(4, np.int32)
        ms = np.zeros((4, 4))
        self._check(lib().rome_autotune(self._h, C.byref(im), C.c_void_p(sig_keys), C.c_void_p(sig_cnt), int(cap),
                                        _p(chosen, C.c_int), _p(ms)), "rome_autotune")
        return chosen.tolist(), ms

    def set_persistent_grids(self, blocks):
        b = np.ascontiguousarray(blocks, np.int32)
        self._check(lib().rome_set_persistent_grids(self._h, _p(b, C.c_int)), "rome_set_persistent_grids")

    def execution_stats(self):
        out = np.zeros(6, np.int64)
        self._check(lib().rome_get_execution_stats(self._h, _p(out, C.c_longlong)), "rome_get_execution_stats")
        return dict(zip(("last", "graph_launches", "graph_captures", "reruns", "launch_us", "wait_us"), out.tolist()))

    def set_significance_cap(self, cap=-1):
        """Significant samples per image kept as keys (<= 4096); 0: the dense mask always; < 0: default."""
        self._check(lib().rome_set_significance_cap(self._h, int(cap)), "rome_set_significance_cap")

    def significance_stats(self):
        out = np.zeros(2, np.int64)
        self._check(lib().rome_get_significance_stats(self._h, _p(out, C.c_longlong)),
                    "rome_get_significance_stats")
        return dict(zip(("sig_form", "sig_max"), out.tolist()))

    def project(self, iclass, size, A):
        A = _f64(A).reshape(-1, 9)
        n = size * (size // 2 + 1)
        re = np.zeros((A.shape[0], n))
        im = np.zeros((A.shape[0], n))
        self._check(lib().rome_project(self._h, iclass, size, A.shape[0], _p(A), _p(re), _p(im)), "rome_project")
        return re, im
