"""ctypes front end of the FP64 oracle (TEST INFRASTRUCTURE).

The oracle is a CPU restatement of ROME's MAP E-step / reconstruction hot
path (oracle/rome_oracle.cpp).  Only tests/, the smoke check and bench.py's
CPU-baseline leg may import this module; the product (rome_amd/) never does.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_ref = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build():
    """Compile the oracle (and the reference HEALPix checker when the reference tree exists)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "librome_oracle.so")
        if not os.path.exists(path):
            build()
        _lib = C.CDLL(path)
        _lib.or_estep.restype = C.c_int
        _lib.or_healpix_xyf2nest.restype = C.c_int
    return _lib


def ref_healpix():
    """The reference's own HEALPix (oracle/_ref), or None where it was not built."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libref_healpix.so")
        if not os.path.exists(path):
            return None
        _ref = C.CDLL(path)
        _ref.ref_healpix_max_pixrad.restype = C.c_double
    return _ref


def ptr(a, t=C.c_double):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(t))


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------- sampling ----------------
def pix2ang(order):
    npix = 12 * 4 ** order
    z = np.zeros(npix)
    phi = np.zeros(npix)
    lib().or_healpix_pix2ang_nest(order, ptr(z), ptr(phi))
    return z, phi


def nest2xyf(order, pix):
    ix, iy, f = C.c_int(), C.c_int(), C.c_int()
    lib().or_healpix_nest2xyf(order, pix, C.byref(ix), C.byref(iy), C.byref(f))
    return ix.value, iy.value, f.value


def xyf2nest(order, ix, iy, face):
    return lib().or_healpix_xyf2nest(order, ix, iy, face)


def translations(offset_range, offset_step):
    n = lib().or_translations(C.c_double(offset_range), C.c_double(offset_step), None, None)
    tx = np.zeros(n)
    ty = np.zeros(n)
    lib().or_translations(C.c_double(offset_range), C.c_double(offset_step), ptr(tx), ptr(ty))
    return tx, ty


def euler_matrix(rot, tilt, psi):
    A = np.zeros(9)
    lib().or_euler_angles2matrix(C.c_double(rot), C.c_double(tilt), C.c_double(psi), ptr(A))
    return A.reshape(3, 3)


def euler_angles(A):
    A = f64(A).reshape(9)
    r, t, p = C.c_double(), C.c_double(), C.c_double()
    lib().or_euler_matrix2angles(ptr(A), C.byref(r), C.byref(t), C.byref(p))
    return r.value, t.value, p.value


def pose_table3d(order, oversampling, perturb):
    nr_dir = 12 * 4 ** order
    nr_psi = 6 * 2 ** order
    nover = 8 ** oversampling
    out = np.zeros((nr_dir, nr_psi, nover, 3))
    lib().or_pose_table3d(order, oversampling, C.c_double(perturb), ptr(out))
    return out if oversampling else out[:, :, 0, :]


def pose_table2d(psi_step, oversampling, perturb):
    nr_psi = int(np.ceil(360.0 / psi_step))
    nover = 2 ** oversampling
    out = np.zeros((nr_psi, nover))
    lib().or_pose_table2d(C.c_double(psi_step), oversampling, C.c_double(perturb), ptr(out))
    return out if oversampling else out[:, 0]


def shift_table(offset_range, offset_step, oversampling, perturb):
    T = len(translations(offset_range, offset_step)[0])
    nover = 4 ** oversampling
    x = np.zeros((T, nover))
    y = np.zeros((T, nover))
    lib().or_shift_table(C.c_double(offset_range), C.c_double(offset_step), oversampling,
                         C.c_double(perturb), ptr(x), ptr(y))
    return x, y


# ---------------- image helpers ----------------
def shift_image(re, im, size, sx, sy, ori_size):
    re = f64(re); im = f64(im)
    ore = np.zeros_like(re); oim = np.zeros_like(im)
    lib().or_shift_image(ptr(re), ptr(im), ptr(ore), ptr(oim), size, C.c_double(sx), C.c_double(sy), ori_size)
    return ore, oim


def window_fourier(re, im, in_size, out_size):
    re = f64(re); im = f64(im)
    n = out_size * (out_size // 2 + 1)
    ore = np.zeros(n); oim = np.zeros(n)
    lib().or_window_fourier(ptr(re), ptr(im), in_size, ptr(ore), ptr(oim), out_size)
    return ore, oim


def mresol(size):
    out = np.zeros(size * (size // 2 + 1), dtype=np.int32)
    lib().or_mresol(size, ptr(out, C.c_int))
    return out


# ---------------- projector / backprojector ----------------
def pad_size(current_size, ori_size=None):
    r_max = current_size // 2 if ori_size is None else min(current_size // 2, ori_size // 2)
    return 2 * (2 * r_max + 1) + 1


def project3d(grid, pad, r_max, A, size):
    grid = f64(grid)
    n = size * (size // 2 + 1)
    re = np.zeros(n); im = np.zeros(n)
    lib().or_project3d(ptr(grid), pad, r_max, ptr(f64(A).reshape(9)), size, ptr(re), ptr(im))
    return re, im


def rotate2d(grid, pad, r_max, A, size):
    grid = f64(grid)
    n = size * (size // 2 + 1)
    re = np.zeros(n); im = np.zeros(n)
    lib().or_rotate2d(ptr(grid), pad, r_max, ptr(f64(A).reshape(9)), size, ptr(re), ptr(im))
    return re, im


def backproject3d(data, weight, pad, r_max, A, size, re, im, mw=None):
    lib().or_backproject3d(ptr(data), ptr(weight), pad, r_max, ptr(f64(A).reshape(9)), size,
                           ptr(f64(re)), ptr(f64(im)), ptr(None if mw is None else f64(mw)))


def backrotate2d(data, weight, pad, r_max, A, size, re, im, mw=None):
    lib().or_backrotate2d(ptr(data), ptr(weight), pad, r_max, ptr(f64(A).reshape(9)), size,
                          ptr(f64(re)), ptr(f64(im)), ptr(None if mw is None else f64(mw)))


# ---------------- E-step ----------------
class OrConfig(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("ori_size", "current_size", "coarse_size", "nr_classes", "nr_groups",
                                        "is_3d", "adaptive_oversampling", "nr_dir", "nr_psi", "nr_trans")] + \
               [(n, C.c_double) for n in ("adaptive_fraction", "sigma2_fudge", "sigma2_offset",
                                           "avg_norm_correction")] + \
               [(n, C.c_int) for n in ("do_ctf_correction", "refs_are_ctf_corrected", "do_norm_correction",
                                        "do_map", "pad", "nr_threads")] + \
               [(n, _dp) for n in ("sigma2_noise", "pdf_class", "pdf_direction", "prior_offx", "prior_offy",
                                   "grids", "pose_coarse", "pose_fine", "shift_coarse_x", "shift_coarse_y",
                                   "shift_fine_x", "shift_fine_y", "trans_x", "trans_y")]


class OrImages(C.Structure):
    _fields_ = [("n", C.c_int)] + [(n, _dp) for n in (
        "fmask_re", "fmask_im", "fnomask_re", "fnomask_im", "fctf", "highres_Xi2", "power_img",
        "old_offx", "old_offy")] + [("group", _ip), ("norm", _dp), ("coarse_sig_in", C.POINTER(C.c_ubyte))]


class OrOutputs(C.Structure):
    _fields_ = [("coarse_diff2", _dp), ("coarse_sig", C.POINTER(C.c_ubyte)), ("fine_diff2", _dp),
                ("fine_weight", _dp)] + \
               [(n, _dp) for n in ("min_diff2_coarse", "sum_weight_coarse", "sig_weight_coarse",
                                   "min_diff2", "sum_weight", "sig_weight")] + \
               [("nr_sign", _ip), ("argmax", _ip)] + \
               [(n, _dp) for n in ("meta", "wsum_norm_correction", "wsum_sigma2_noise", "wsum_sumw_group",
                                   "wsum_pdf_class", "wsum_pdf_direction", "wsum_scalars",
                                   "wsum_prior_offx", "wsum_prior_offy", "bp_data", "bp_weight")]


META_FIELDS = ("PMAX", "CLASS", "ROT", "TILT", "PSI", "XOFF", "YOFF", "NORM", "DLL")


def estep(prob, batch, dump_coarse=False, dump_fine=False, dump_fine_weight=False, nr_threads=0):
    """Run the oracle E-step.

    prob: dict with the model / sampling (see rome_amd.problem.Problem.as_dict()).
    batch: dict with per-image arrays.
    Returns a dict of numpy outputs.
    """
    keep = []

    def P(a, t=C.c_double):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=np.int32 if t is C.c_int else np.float64)
        keep.append(a)
        return a.ctypes.data_as(C.POINTER(t))

    is3d = int(prob["is_3d"])
    K, D, PSI, T = prob["nr_classes"], prob["nr_dir"], prob["nr_psi"], prob["nr_trans"]
    os_ = prob["adaptive_oversampling"]
    nor = (8 if is3d else 2) if os_ else 1
    nots = 4 if os_ else 1
    G = prob["nr_groups"]
    oriF = prob["ori_size"] // 2 + 1
    pad = prob["pad"]
    padF = pad // 2 + 1
    gridN = pad * pad * padF if is3d else pad * padF
    cfg = OrConfig(
        ori_size=prob["ori_size"], current_size=prob["current_size"], coarse_size=prob["coarse_size"],
        nr_classes=K, nr_groups=G, is_3d=is3d, adaptive_oversampling=os_, nr_dir=D, nr_psi=PSI, nr_trans=T,
        adaptive_fraction=prob["adaptive_fraction"], sigma2_fudge=prob["sigma2_fudge"],
        sigma2_offset=prob["sigma2_offset"], avg_norm_correction=prob["avg_norm_correction"],
        do_ctf_correction=int(prob["do_ctf_correction"]), refs_are_ctf_corrected=int(prob["refs_are_ctf_corrected"]),
        do_norm_correction=int(prob["do_norm_correction"]), do_map=int(prob["do_map"]), pad=pad,
        nr_threads=nr_threads,
        sigma2_noise=P(prob["sigma2_noise"]), pdf_class=P(prob["pdf_class"]),
        pdf_direction=P(prob["pdf_direction"]), prior_offx=P(prob["prior_offx"]), prior_offy=P(prob["prior_offy"]),
        grids=P(prob["grids"]), pose_coarse=P(prob["pose_coarse"]), pose_fine=P(prob.get("pose_fine")),
        shift_coarse_x=P(prob["shift_coarse_x"]), shift_coarse_y=P(prob["shift_coarse_y"]),
        shift_fine_x=P(prob.get("shift_fine_x")), shift_fine_y=P(prob.get("shift_fine_y")),
        trans_x=P(prob["trans_x"]), trans_y=P(prob["trans_y"]))
    n = int(batch["n"])
    ims = OrImages(n=n, fmask_re=P(batch["fmask_re"]), fmask_im=P(batch["fmask_im"]),
                   fnomask_re=P(batch["fnomask_re"]), fnomask_im=P(batch["fnomask_im"]), fctf=P(batch["fctf"]),
                   highres_Xi2=P(batch["highres_Xi2"]), power_img=P(batch["power_img"]),
                   old_offx=P(batch["old_offx"]), old_offy=P(batch["old_offy"]),
                   group=P(batch["group"], C.c_int), norm=P(batch["norm"]))
    if batch.get("coarse_sig_in") is not None:  # coarse significance given: the fine pass alone
        cs = np.ascontiguousarray(batch["coarse_sig_in"], dtype=np.uint8)
        assert cs.size == n * K * D * PSI * T
        keep.append(cs)
        ims.coarse_sig_in = cs.ctypes.data_as(C.POINTER(C.c_ubyte))
    r = {}

    def alloc(name, shape, dtype=np.float64, fill=0.0):
        a = np.full(shape, fill, dtype=dtype)
        r[name] = a
        return a.ctypes.data_as(C.POINTER({np.float64: C.c_double, np.int32: C.c_int,
                                           np.uint8: C.c_ubyte}[dtype]))

    out = OrOutputs()
    if dump_coarse:
        out.coarse_diff2 = alloc("coarse_diff2", (n, K, D, PSI, T))
        out.coarse_sig = alloc("coarse_sig", (n, K, D, PSI, T), np.uint8)
    if dump_fine:
        out.fine_diff2 = alloc("fine_diff2", (n, K, D, PSI, T, nor, nots))
    if dump_fine_weight:
        out.fine_weight = alloc("fine_weight", (n, K, D, PSI, T, nor, nots))
    for nm in ("min_diff2_coarse", "sum_weight_coarse", "sig_weight_coarse", "min_diff2", "sum_weight",
               "sig_weight", "wsum_norm_correction"):
        setattr(out, nm, alloc(nm, (n,)))
    out.nr_sign = alloc("nr_sign", (n,), np.int32)
    out.argmax = alloc("argmax", (n, 6), np.int32)
    out.meta = alloc("meta", (n, 9))
    out.wsum_sigma2_noise = alloc("wsum_sigma2_noise", (G, oriF))
    out.wsum_sumw_group = alloc("wsum_sumw_group", (G,))
    out.wsum_pdf_class = alloc("wsum_pdf_class", (K,))
    out.wsum_pdf_direction = alloc("wsum_pdf_direction", (K, D))
    out.wsum_scalars = alloc("wsum_scalars", (4,))
    out.wsum_prior_offx = alloc("wsum_prior_offx", (K,))
    out.wsum_prior_offy = alloc("wsum_prior_offy", (K,))
    out.bp_data = alloc("bp_data", (K, gridN * 2))
    out.bp_weight = alloc("bp_weight", (K, gridN))
    err = C.create_string_buffer(512)
    rc = lib().or_estep(C.byref(cfg), C.byref(ims), C.byref(out), err, 512)
    if rc != 0:
        raise RuntimeError("oracle E-step failed: " + err.value.decode())
    return r
