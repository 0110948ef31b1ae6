"""FP64 restatement of ROME's image preparation for the E-step (TEST INFRASTRUCTURE: only tests/, the
smoke check and bench.py's CPU-baseline leg may use it).

ParticleModel::setFourierTransforms, src/exp_model.h:397-639, per image:
  * norm correction: image * avg_norm_correction / NORM                         (:430-437)
  * rounded old offsets, translate(DONT_WRAP) = integer shift with zero fill      (:446-449; src/image.h:53-257)
  * CenterFFT (forward), FFT normalised by the number of real samples, window to
    current_size -> the unmasked transform                                       (:461-471; src/image.h:277-343,496-566)
  * softMaskOutsideMap on the translated image: raised cosine over cosine_width
    pixels outside radius = particle_diameter / (2 pixel_size), filled with the
    given noise image or, without one (-zero_mask), with the background average  (:480-545; src/image.h:571-643)
  * CenterFFT, FFT, window -> the masked transform                                (:547-558)
  * power spectrum and highres_Xi2 when current_size < ori_size                    (:569-604)
  * CTF::getFftwImage on the current_size grid                                    (:612-627; src/ctf.cpp:136-210,
                                                                                     src/ctf.h:151-205)
The reference draws its noise image with std::default_random_engine inside the loop; here (as at the
engine's boundary) the noise image is an input, so both sides use the same numbers.  numpy's FFT
(pocketfft) stands in for MKL's FFTW interface.
"""
import math

import numpy as np


def translate_int(img, ox, oy):
    """applyGeometry of translate() for integer offsets without wrapping: out(y, x) = in(y - oy, x - ox),
    zero where that falls outside (src/image.h:124-236 with Aref = [[1, 0, -ox], [0, 1, -oy]])."""
    n = img.shape[0]
    out = np.zeros_like(img)
    ox, oy = int(ox), int(oy)
    ys = slice(max(0, oy), min(n, n + oy))
    xs = slice(max(0, ox), min(n, n + ox))
    yi = slice(max(0, -oy), min(n, n - oy))
    xi = slice(max(0, -ox), min(n, n - ox))
    out[ys, xs] = img[yi, xi]
    return out


def center_fft(img, forward=True):
    """CenterFFT: circular shift by n/2 along each axis (src/image.h:277-343)."""
    n = img.shape[0]
    s = n // 2 if forward else -(n // 2)
    return np.roll(np.roll(img, s, axis=1), s, axis=0)


def fft_half(img):
    """FourierTransform: r2c half-plane [n][n/2+1] divided by n*n (normalize = true)."""
    n = img.shape[0]
    return np.fft.rfft2(img) / float(n * n)


def window_ft(F, in_size, out_size):
    """windowFourierTransform for 2D half-planes (src/image.h:496-566), both branches."""
    in_F, out_F = in_size // 2 + 1, out_size // 2 + 1
    if in_size == out_size:
        return F.copy()
    out = np.zeros((out_size, out_F), dtype=complex)
    if out_F > in_F:
        max_r2 = (in_F - 1) * (in_F - 1)
        for i in range(in_size):
            ip = i if i < in_F else i - in_size
            in_i = ip + in_size if ip < 0 else ip
            out_i = ip + out_size if ip < 0 else ip
            jp2 = max_r2 - ip * ip
            if jp2 < 0:
                continue
            j_max = min(int(math.sqrt(jp2)) + 1, in_F)
            out[out_i, :j_max] = F[in_i, :j_max]
    else:
        for i in range(out_size):
            ip = i if i < out_F else i - out_size
            in_i = ip + in_size if ip < 0 else ip
            out_i = ip + out_size if ip < 0 else ip
            out[out_i, :] = F[in_i, :out_F]
    return out


def soft_mask_outside(img, radius, cosine_width, noise=None):
    """softMaskOutsideMap for 2D images (src/image.h:571-643): origin at -(int)(n/2.0)."""
    n = img.shape[0]
    if radius < 0:
        radius = n / 2.0
    radius_p = radius + cosine_width
    o = -int(float(n) / 2.0)
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    r = np.sqrt(((ii + o) ** 2 + (jj + o) ** 2).astype(np.float64))
    out = img.copy()
    edge = (r >= radius) & (r <= radius_p)
    outer = r > radius_p
    rc = 0.5 + 0.5 * np.cos(math.pi * (radius_p - r) / cosine_width)
    if noise is None:
        sum_bg = 0.0
        cnt = 0.0
        for i in range(n):          # the reference's accumulation order (row by row)
            for j in range(n):
                if r[i, j] < radius:
                    continue
                if r[i, j] > radius_p:
                    cnt += 1.0
                    sum_bg += img[i, j]
                else:
                    cnt += rc[i, j]
                    sum_bg += rc[i, j] * img[i, j]
        fill = np.full_like(img, sum_bg / cnt if cnt else math.nan)  # 0/0 when the mask covers it all: unused
    else:
        fill = noise
    out[edge] = (1.0 - rc[edge]) * img[edge] + rc[edge] * fill[edge]
    out[outer] = fill[outer]
    return out


def ctf_image(dim, ori, angpix, defU, defV, defAng, voltage, Cs, Q0, Bfac, scale=1.0, do_abs=False,
              only_flip_phases=False, intact_first_peak=False, do_damping=True):
    """CTF::initialise + getFftwImage (src/ctf.cpp:136-210, src/ctf.h:151-205) on a dim x (dim/2+1) grid."""
    local_Cs = Cs * 1e7
    local_kV = voltage * 1e3
    rad_az = defAng * math.pi / 180.0  # DEG2RAD
    dfa = -(defU + defV) * 0.5
    dfd = -(defU - defV) * 0.5
    lam = 12.2643247 / math.sqrt(local_kV * (1.0 + local_kV * 0.978466e-6))
    K1 = math.pi / 2 * 2 * lam
    K2 = math.pi / 2 * local_Cs * lam * lam * lam
    K3 = math.sqrt(1 - Q0 * Q0)
    K4 = -Bfac / 4.0
    xs = float(ori) * angpix
    ys = float(ori) * angpix
    h = dim // 2 + 1
    out = np.zeros((dim, h))
    for i in range(dim):
        ip = i if i < h else i - dim
        for j in range(h):
            X = j / xs
            Y = ip / ys
            u2 = X * X + Y * Y
            u4 = u2 * u2
            if abs(X) < 1e-6 and abs(Y) < 1e-6:
                deltaf = 0.0
            else:
                ang = math.atan2(Y, X) - rad_az
                deltaf = dfa + dfd * math.cos(2 * ang)
            arg = K1 * deltaf * u2 + K2 * u4
            if intact_first_peak and abs(arg) < math.pi / 2.0:
                v = 1.0
            else:
                v = -(K3 * math.sin(arg) - Q0 * math.cos(arg))
            if do_damping:
                v *= math.exp(K4 * u2)
            if do_abs:
                v = abs(v)
            elif only_flip_phases:
                v = -1.0 if v < 0.0 else 1.0
            out[i, j] = scale * v
    return out


def set_fourier_transforms(images, norm, xoff, yoff, ctf, p, noise=None):
    """images [n][ori][ori]; norm, xoff, yoff [n]; ctf [n][7] (defU, defV, defAngle, voltage, Cs, Q0, Bfac);
    p: dict ori, current, pixel_size, particle_diameter, width_mask_edge, do_norm_correction,
    avg_norm_correction, do_ctf_correction, ctf_phase_flipped, only_flip_phases, intact_ctf_first_peak;
    noise [n][ori][ori] or None.  Returns the E-step inputs (the fields of rome_images)."""
    n, ori = images.shape[0], images.shape[1]
    cur = p["current"]
    oriF, curF = ori // 2 + 1, cur // 2 + 1
    out = {k: [] for k in ("fnomask", "fmask", "fctf", "power", "highres_Xi2", "old_offx", "old_offy")}
    for i in range(n):
        img = images[i].astype(np.float64)
        if p["do_norm_correction"]:
            img = img * p["avg_norm_correction"] / norm[i]
        ox, oy = _c_round(xoff[i]), _c_round(yoff[i])
        timg = translate_int(img, ox, oy)
        F = fft_half(center_fft(timg, True))
        out["fnomask"].append(window_ft(F, ori, cur))
        radius = p["particle_diameter"] / (2.0 * p["pixel_size"])
        mimg = soft_mask_outside(timg, radius, float(p["width_mask_edge"]), None if noise is None else noise[i])
        Fm = fft_half(center_fft(mimg, True))
        out["fmask"].append(window_ft(Fm, ori, cur))
        spec = np.zeros(oriF)
        hx = 0.0
        if cur < ori:
            for a in range(ori):
                ip = a if a < oriF else a - ori
                for b in range(oriF):
                    ires = _c_round(math.sqrt(ip * ip + b * b))
                    if 0 < ires < oriF and not (b == 0 and ip < 0):
                        v = Fm[a, b].real ** 2 + Fm[a, b].imag ** 2
                        spec[ires] += v
                        if ires >= curF:
                            hx += v
        out["power"].append(spec)
        out["highres_Xi2"].append(hx)
        if p["do_ctf_correction"]:
            out["fctf"].append(ctf_image(cur, ori, p["pixel_size"], *ctf[i], do_abs=p["ctf_phase_flipped"],
                                         only_flip_phases=p["only_flip_phases"],
                                         intact_first_peak=p["intact_ctf_first_peak"], do_damping=True))
        else:
            out["fctf"].append(np.ones((cur, curF)))
        out["old_offx"].append(float(ox))
        out["old_offy"].append(float(oy))
    return {k: np.array(v) for k, v in out.items()}


def _c_round(x):
    """C round(): half away from zero."""
    return int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))
