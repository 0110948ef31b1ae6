"""Seeded synthetic MAP2D / MAP3D E-step problems (numpy).

ROME ships no data generator (src/simulatedata.h is an empty stub), so the inputs of
the E-step are synthesised here in the shapes the reference hands to
expectationSomeParticles: reference volumes (sums of Gaussian blobs) turned into
2x-padded Fourier grids the way MyProjector::computeFourierTransformMap does
(src/reconstruct.cpp:106-223), particles = CTF * projection * shift phase + white
noise at a given SNR, written directly as the prepared Fourier half-planes
(no real-space soft mask), plus the model state (sigma2_noise, pdfs, priors) and the
perturbed sampling tables of the product's own HealpixSampler port.
"""
import numpy as np

from . import PRECISIONS, pose_table2d, pose_table3d, shift_table, translations

PI = 3.14159265358979323846


# ---------------- volumes ----------------
def blob_volume(ori, nblobs, sigma_range, radius, rng, ndim=3):
    """Sum of isotropic Gaussian blobs inside a sphere / disc of `radius` px (centred)."""
    ax = np.arange(ori) - ori // 2
    grids = np.meshgrid(*([ax] * ndim), indexing="ij")
    vol = np.zeros([ori] * ndim)
    for _ in range(nblobs):
        while True:
            c = rng.uniform(-radius, radius, ndim)
            if np.sum(c * c) <= radius * radius:
                break
        s = rng.uniform(*sigma_range)
        r2 = sum((g - ci) ** 2 for g, ci in zip(grids, c))
        vol += rng.uniform(0.5, 1.5) * np.exp(-r2 / (2 * s * s))
    return vol


def fourier_grid(vol, current_size, padding_factor=2, gridding=True):
    """2x-padded, r <= 2*r_max windowed Fourier grid [z][y][x] (x half), complex interleaved.

    Follows computeFourierTransformMap: gridding correction (divide by sinc^2),
    zero padding to 2*ori, centre, forward FFT normalised by 1/N, scale by normfft.
    """
    ndim = vol.ndim
    ori = vol.shape[0]
    r_max = min(current_size // 2, ori // 2)
    pad = 2 * (padding_factor * r_max + 1) + 1
    padF = pad // 2 + 1
    v = vol.astype(np.float64).copy()
    if gridding:
        ax = np.arange(ori) - ori // 2
        g = np.meshgrid(*([ax] * ndim), indexing="ij")
        r = np.sqrt(sum(x * x for x in g))
        rval = r / (ori * padding_factor)
        with np.errstate(invalid="ignore", divide="ignore"):
            sinc = np.sin(PI * rval) / (PI * rval)
        corr = np.where(r > 0, sinc * sinc, 1.0)
        v = v / corr
    N = padding_factor * ori
    P = np.zeros([N] * ndim)
    o = N // 2 - ori // 2
    sl = tuple(slice(o, o + ori) for _ in range(ndim))
    P[sl] = v
    P = np.fft.ifftshift(P)   # origin of the map at index 0
    F = np.fft.rfftn(P) / P.size
    normfft = padding_factor ** 3 * ori if ndim == 3 else padding_factor ** 2
    F = F * normfft
    max_r2 = (r_max * padding_factor) ** 2
    c = pad // 2
    if ndim == 3:
        grid = np.zeros((pad, pad, padF), np.complex128)
        for kp in range(-2 * r_max, 2 * r_max + 1):
            for ip in range(-2 * r_max, 2 * r_max + 1):
                jmax2 = max_r2 - kp * kp - ip * ip
                if jmax2 < 0:
                    continue
                jm = int(np.floor(np.sqrt(jmax2)))
                grid[kp + c, ip + c, :jm + 1] = F[kp % N, ip % N, :jm + 1]
    else:
        grid = np.zeros((pad, padF), np.complex128)
        for ip in range(-2 * r_max, 2 * r_max + 1):
            jmax2 = max_r2 - ip * ip
            if jmax2 < 0:
                continue
            jm = int(np.floor(np.sqrt(jmax2)))
            grid[ip + c, :jm + 1] = F[ip % N, :jm + 1]
    out = np.zeros(grid.shape + (2,))
    out[..., 0] = grid.real
    out[..., 1] = grid.imag
    return out.reshape(-1), pad


def blob_volume_fast(ori, nblobs, sigma_range, radius, rng):
    """3D blob volume evaluated on a +-4 sigma sub-box per blob (same model as blob_volume)."""
    vol = np.zeros((ori, ori, ori))
    c0 = ori // 2
    for _ in range(nblobs):
        while True:
            c = rng.uniform(-radius, radius, 3)
            if np.sum(c * c) <= radius * radius:
                break
        s = rng.uniform(*sigma_range)
        amp = rng.uniform(0.5, 1.5)
        h = int(np.ceil(4 * s))
        lo = [max(0, int(np.floor(ci)) + c0 - h) for ci in c]
        hi = [min(ori, int(np.floor(ci)) + c0 + h + 2) for ci in c]
        ax = [np.arange(l, u) - c0 - ci for l, u, ci in zip(lo, hi, c)]
        r2 = ax[0][:, None, None] ** 2 + ax[1][None, :, None] ** 2 + ax[2][None, None, :] ** 2
        vol[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] += amp * np.exp(-r2 / (2 * s * s))
    return vol


# ---------------- geometry ----------------
def euler_matrix(rot, tilt, psi):
    a, b, g = np.deg2rad(rot), np.deg2rad(tilt), np.deg2rad(psi)
    ca, cb, cg, sa, sb, sg = np.cos(a), np.cos(b), np.cos(g), np.sin(a), np.sin(b), np.sin(g)
    cc, cs, sc, ss = cb * ca, cb * sa, sb * ca, sb * sa
    return np.array([[cg * cc - sg * sa, cg * cs + sg * ca, -cg * sb],
                     [-sg * cc - cg * sa, -sg * cs + cg * ca, sg * sb],
                     [sc, ss, cb]])


def halfplane_xy(size):
    F = size // 2 + 1
    i = np.arange(size)[:, None]
    j = np.arange(F)[None, :]
    y = np.where(i < F, i, i - size) + 0 * j
    x = j + 0 * i
    return x.reshape(-1), y.reshape(-1)


def project_numpy(grid, pad, r_max, A, size, is3d=True):
    """Vectorised trilinear (3D) / bilinear (2D) slice of the padded grid (float64)."""
    padF = pad // 2 + 1
    c = pad // 2
    g = grid.reshape((pad, pad, padF, 2) if is3d else (pad, padF, 2))
    gc = g[..., 0] + 1j * g[..., 1]
    x, y = halfplane_xy(size)
    myr = min(r_max, size // 2)
    F = size // 2 + 1
    i = np.arange(size * F) // F
    okrow = (i <= myr) | (i >= size - myr)
    ok = okrow & (x <= myr) & (x * x + y * y <= myr * myr)
    Ai = 2.0 * A.T
    xp = Ai[0, 0] * x + Ai[0, 1] * y
    yp = Ai[1, 0] * x + Ai[1, 1] * y
    zp = Ai[2, 0] * x + Ai[2, 1] * y
    neg = xp < 0
    xp, yp, zp = np.where(neg, -xp, xp), np.where(neg, -yp, yp), np.where(neg, -zp, zp)
    x0, y0, z0 = np.floor(xp).astype(int), np.floor(yp).astype(int), np.floor(zp).astype(int)
    fx, fy, fz = xp - x0, yp - y0, zp - z0
    x0 = np.where(ok, x0, 0)
    y0 = np.where(ok, y0 + c, c)
    z0 = np.where(ok, z0 + c, c)
    if is3d:
        def at(dz, dy, dx):
            return gc[z0 + dz, y0 + dy, x0 + dx]
        v = ((1 - fz) * ((1 - fy) * ((1 - fx) * at(0, 0, 0) + fx * at(0, 0, 1)) +
                         fy * ((1 - fx) * at(0, 1, 0) + fx * at(0, 1, 1))) +
             fz * ((1 - fy) * ((1 - fx) * at(1, 0, 0) + fx * at(1, 0, 1)) +
                   fy * ((1 - fx) * at(1, 1, 0) + fx * at(1, 1, 1))))
    else:
        def at2(dy, dx):
            return gc[y0 + dy, x0 + dx]
        v = (1 - fy) * ((1 - fx) * at2(0, 0) + fx * at2(0, 1)) + fy * ((1 - fx) * at2(1, 0) + fx * at2(1, 1))
    v = np.where(neg, np.conj(v), v)
    return np.where(ok, v, 0)


def ctf_image(size, ori, angpix, defU, defV, ang_deg, kv=300.0, cs=2.7, q0=0.1, bfac=0.0):
    """CTF::getFftwImage (src/ctf.cpp:173-191) with CTF::getCTF (src/ctf.h:151-205)."""
    local_cs = cs * 1e7
    local_kv = kv * 1e3
    lam = 12.2643247 / np.sqrt(local_kv * (1. + local_kv * 0.978466e-6))
    K1 = PI / 2 * 2 * lam
    K2 = PI / 2 * local_cs * lam ** 3
    K3 = np.sqrt(1 - q0 * q0)
    K4 = -bfac / 4.
    defavg = -(defU + defV) * 0.5
    defdev = -(defU - defV) * 0.5
    rad_az = np.deg2rad(ang_deg)
    xs = ori * angpix
    x, y = halfplane_xy(size)
    X = x / xs
    Y = y / xs
    u2 = X * X + Y * Y
    ell = np.arctan2(Y, X) - rad_az
    deltaf = np.where((np.abs(X) < 1e-6) & (np.abs(Y) < 1e-6), 0.0, defavg + defdev * np.cos(2 * ell))
    arg = K1 * deltaf * u2 + K2 * u2 * u2
    return -(K3 * np.sin(arg) - q0 * np.cos(arg)) * np.exp(K4 * u2)


def mresol(size):
    x, y = halfplane_xy(size)
    F = size // 2 + 1
    ires = np.round(np.sqrt(x * x + y * y)).astype(int)
    return np.where((ires < F) & ~((x == 0) & (y < 0)), ires, -1)


def window(arr, in_size, out_size):
    """windowFourierTransform down-size branch on a [.., in F2] array."""
    Fi, Fo = in_size // 2 + 1, out_size // 2 + 1
    a = arr.reshape(arr.shape[:-1] + (in_size, Fi))
    rows = [ip if ip >= 0 else ip + in_size for ip in [i if i < Fo else i - out_size for i in range(out_size)]]
    return a[..., rows, :Fo].reshape(arr.shape[:-1] + (out_size * Fo,))


# ---------------- problem ----------------
def make_problem(is3d=True, ori=32, current=32, coarse=16, K=2, order=1, psi_step=10.0, offset_range=2.0,
                 offset_step=1.0, oversampling=1, perturbation=0.3, n=4, snr=0.5, nr_groups=1, seed=0,
                 do_ctf=True, angpix=2.0, nblobs=20, sigma_offset=None, true_class=None):
    rng = np.random.default_rng(seed)
    ndim = 3 if is3d else 2
    radius = 0.35 * ori
    if is3d and ori >= 64:
        vols = [blob_volume_fast(ori, nblobs, (2.0, 6.0), radius, rng) for _ in range(K)]
    else:
        vols = [blob_volume(ori, nblobs, (1.0, max(1.5, ori / 16)), radius, rng, ndim) for _ in range(K)]
    grids = []
    pad = None
    for v in vols:
        g, pad = fourier_grid(v, current)
        grids.append(g)
    grids = np.stack(grids)
    r_max = min(current // 2, ori // 2)
    # sampling (product port of HealpixSampler)
    if is3d:
        pose_c = pose_table3d(order, 0, perturbation)
        pose_f = pose_table3d(order, 1, perturbation) if oversampling else None
        D, PSI = pose_c.shape[0], pose_c.shape[1]
    else:
        pose_c = pose_table2d(psi_step, 0, perturbation)
        pose_f = pose_table2d(psi_step, 1, perturbation) if oversampling else None
        D, PSI = 1, pose_c.shape[0]
    tx, ty = translations(offset_range, offset_step)
    T = len(tx)
    scx, scy = shift_table(offset_range, offset_step, 0, perturbation)
    sfx, sfy = shift_table(offset_range, offset_step, 1, perturbation) if oversampling else (None, None)
    oriF = ori // 2 + 1
    F2o = ori * oriF
    # particles at ori size, then windowed
    imgs = np.zeros((n, F2o), np.complex128)
    ctfs = np.zeros((n, F2o))
    ctf_params = np.zeros((n, 7))  # defocus U, V (A), angle (deg), voltage (kV), Cs (mm), Q0, B-factor
    truth = []
    sig_power = []
    for i in range(n):
        k = int(rng.integers(K)) if true_class is None else true_class
        if is3d:
            rot, tilt, psi = rng.uniform(-180, 180), np.rad2deg(np.arccos(rng.uniform(-1, 1))), rng.uniform(0, 360)
        else:
            rot, tilt, psi = 0.0, 0.0, rng.uniform(0, 360)
        sh = rng.integers(-int(offset_range), int(offset_range) + 1, 2).astype(float)
        A = euler_matrix(rot, tilt, psi)
        proj = project_numpy(grids[k], pad, r_max, A, ori if current == ori else current, is3d)
        if current != ori:
            full = np.zeros(F2o, np.complex128)
            xo, yo = halfplane_xy(ori)
            xc, yc = halfplane_xy(current)
            idx = {(a, b): q for q, (a, b) in enumerate(zip(xo, yo))}
            for q, (a, b) in enumerate(zip(xc, yc)):
                full[idx[(a, b)]] = proj[q]
            proj = full
        if do_ctf:
            dU, dV, da = rng.uniform(1e4, 3e4), rng.uniform(1e4, 3e4), rng.uniform(0, 180)
            ctf = ctf_image(ori, ori, angpix, dU, dV, da)
            ctf_params[i] = (dU, dV, da, 300.0, 2.7, 0.1, 0.0)
        else:
            ctf = np.ones(F2o)
        xo, yo = halfplane_xy(ori)
        # particle translated so that the reference's estimate XOFF/YOFF equals `sh`
        phase = np.exp(2j * PI * (xo * sh[0] + yo * sh[1]) / ori)
        sig = ctf * proj * phase
        imgs[i] = sig
        ctfs[i] = ctf
        sig_power.append(np.mean(np.abs(sig) ** 2))
        truth.append((k, rot, tilt, psi, sh[0], sh[1]))
    sigma2 = 0.5 * float(np.mean(sig_power)) / snr if snr > 0 else 1.0
    noise = rng.normal(0, np.sqrt(sigma2), (n, F2o)) + 1j * rng.normal(0, np.sqrt(sigma2), (n, F2o))
    imgs = imgs + noise
    # x = 0 column: keep Hermitian-consistent values (y<0 mirrors y>0)
    xo, yo = halfplane_xy(ori)
    mres_o = mresol(ori)
    power = np.zeros((n, oriF))
    hxi2 = np.zeros(n)
    curF = current // 2 + 1
    for i in range(n):
        a2 = np.abs(imgs[i]) ** 2
        sel = mres_o > 0
        np.add.at(power[i], mres_o[sel], a2[sel])
        if current < ori:
            hxi2[i] = a2[sel & (mres_o >= curF)].sum()
    fimg = window(imgs, ori, current)
    fctf = window(ctfs, ori, current)
    groups = rng.integers(0, nr_groups, n).astype(np.int32)
    sigma2_noise = np.full((nr_groups, oriF), sigma2) * rng.uniform(0.9, 1.1, (nr_groups, oriF))
    prob = dict(
        is_3d=int(is3d), ori_size=ori, current_size=current, coarse_size=coarse if oversampling else current,
        nr_classes=K, nr_groups=nr_groups, adaptive_oversampling=oversampling, nr_dir=D, nr_psi=PSI, nr_trans=T,
        adaptive_fraction=0.999, sigma2_fudge=1.0,
        sigma2_offset=(offset_range / 2.0) ** 2 if sigma_offset is None else sigma_offset,
        avg_norm_correction=1.0, do_ctf_correction=int(do_ctf), refs_are_ctf_corrected=1, do_norm_correction=1,
        do_map=1, pad=pad, sigma2_noise=sigma2_noise, pdf_class=np.full(K, 1.0 / K),
        pdf_direction=np.full((K, D), 1.0 / D), prior_offx=np.zeros(K), prior_offy=np.zeros(K), grids=grids,
        pose_coarse=pose_c, pose_fine=pose_f, shift_coarse_x=scx[:, 0], shift_coarse_y=scy[:, 0],
        shift_fine_x=sfx, shift_fine_y=sfy, trans_x=tx, trans_y=ty)
    batch = dict(n=n, fmask_re=fimg.real.copy(), fmask_im=fimg.imag.copy(), fnomask_re=fimg.real.copy(),
                 fnomask_im=fimg.imag.copy(), fctf=fctf, highres_Xi2=hxi2, power_img=power,
                 old_offx=np.zeros(n), old_offy=np.zeros(n), group=groups, norm=np.ones(n),
                 ctf_params=ctf_params, angpix=angpix)
    return prob, batch, truth


def make_particles_device(engine, nr_classes, n, size, ori, angpix, offset_range, snr, seed, chunk=2000,
                          is3d=True):
    """Particles generated on the GPU: projector of the engine (HIP) + CTF / shift / noise in torch.

    Returns device tensors: F (complex64 [n][F2]) used as both masked and unmasked image, CTF (float32
    [n][F2]), and the host-side truth (class, rot, tilt, psi, sx, sy) plus the noise sigma2 per component.
    """
    import torch
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(seed)
    F2 = size * (size // 2 + 1)
    cls = rng.integers(0, nr_classes, n)
    if is3d:
        rot = rng.uniform(-180, 180, n)
        tilt = np.rad2deg(np.arccos(rng.uniform(-1, 1, n)))
    else:
        rot = np.zeros(n)
        tilt = np.zeros(n)
    psi = rng.uniform(0, 360, n)
    sh = rng.uniform(-offset_range, offset_range, (n, 2))
    defU = rng.uniform(1e4, 3e4, n)
    defV = rng.uniform(1e4, 3e4, n)
    ang = rng.uniform(0, 180, n)
    F = torch.empty((n, F2), dtype=torch.complex64, device=dev)
    x, y = halfplane_xy(size)
    xt = torch.tensor(x, dtype=torch.float64, device=dev)
    yt = torch.tensor(y, dtype=torch.float64, device=dev)
    # CTF (src/ctf.h:151-205) evaluated in FP64 on the device
    lam = 12.2643247 / np.sqrt(300e3 * (1. + 300e3 * 0.978466e-6))
    K1 = PI / 2 * 2 * lam
    K2 = PI / 2 * 2.7e7 * lam ** 3
    q0 = 0.1
    K3 = np.sqrt(1 - q0 * q0)
    xs = ori * angpix
    X = xt / xs
    Y = yt / xs
    u2 = X * X + Y * Y
    atan = torch.atan2(Y, X)
    origin = (X.abs() < 1e-6) & (Y.abs() < 1e-6)
    ctf = torch.empty((n, F2), dtype=torch.float32, device=dev)
    for c0 in range(0, n, chunk):
        c1 = min(n, c0 + chunk)
        idx = np.arange(c0, c1)
        for k in range(nr_classes):
            sel = idx[cls[idx] == k]
            if len(sel) == 0:
                continue
            A = np.stack([euler_matrix(rot[i], tilt[i], psi[i]) for i in sel])
            buf = torch.empty((len(sel), F2), dtype=torch.complex64, device=dev)
            engine.project_device(k, size, A, buf.data_ptr())
            F[torch.tensor(sel, device=dev)] = buf
        dU = torch.tensor(defU[c0:c1], device=dev)[:, None]
        dV = torch.tensor(defV[c0:c1], device=dev)[:, None]
        az = torch.tensor(np.deg2rad(ang[c0:c1]), device=dev)[:, None]
        davg = -(dU + dV) * 0.5
        ddev = -(dU - dV) * 0.5
        deltaf = torch.where(origin[None, :], torch.zeros_like(davg), davg + ddev * torch.cos(2 * (atan[None, :] - az)))
        arg = K1 * deltaf * u2[None, :] + K2 * u2[None, :] ** 2
        c = -(K3 * torch.sin(arg) - q0 * torch.cos(arg))
        ctf[c0:c1] = c.float()
        sx = torch.tensor(sh[c0:c1, 0], device=dev)[:, None]
        sy = torch.tensor(sh[c0:c1, 1], device=dev)[:, None]
        ph = 2 * PI * (xt[None, :] * sx + yt[None, :] * sy) / ori
        F[c0:c1] = F[c0:c1] * c.to(torch.complex64) * torch.polar(torch.ones_like(ph), ph).to(torch.complex64)
    inside = torch.tensor((x * x + y * y <= (size // 2) ** 2).astype(np.float32), device=dev)
    sig_pow = float(((F.abs() ** 2) * inside).sum() / (inside.sum() * n))
    sigma2 = 0.5 * sig_pow / snr
    g = torch.Generator(device=dev)
    g.manual_seed(int(seed) + 2)
    noise = torch.randn((n, F2, 2), generator=g, device=dev, dtype=torch.float32) * float(np.sqrt(sigma2))
    F += torch.view_as_complex(noise)
    truth = np.column_stack([cls, rot, tilt, psi, sh])
    return F, ctf, truth, sigma2


def lognormal_counts(n, median, sigma, lo, hi, rng):
    """Significant coarse samples per image: lognormal with the given median (log-space sigma), rounded and
    clipped to [lo, hi]."""
    return np.clip(np.rint(median * np.exp(sigma * rng.standard_normal(n))), lo, hi).astype(np.int32)


def clustered_significance(prob, truth, counts, cap=None, chunk=256):
    """Coarse significance around each image's true pose, as an adaptive coarse pass leaves it: the
    counts[i] (class, direction, psi, translation) samples of the true class nearest the true pose, by
    (angle / psi step)^2 + (offset distance / offset step)^2.  Returns int32 keys [n][cap] (key = m * T + t,
    m = (class * D + dir) * PSI + psi; unused slots -1) and counts [n]."""
    K, D, PSI, T = prob["nr_classes"], prob["nr_dir"], prob["nr_psi"], prob["nr_trans"]
    n = len(counts)
    cap = int(cap or max(1, int(np.max(counts))))
    counts = np.minimum(np.asarray(counts, np.int64), min(cap, D * PSI * T))
    pc = np.asarray(prob["pose_coarse"], float)
    if prob["is_3d"]:
        Rc = np.stack([euler_matrix(*a) for a in pc.reshape(-1, 3)])  # [D*PSI][3][3]
    else:
        Rc = np.stack([euler_matrix(0.0, 0.0, a) for a in pc.reshape(-1)])
    tx, ty = np.asarray(prob["trans_x"], float), np.asarray(prob["trans_y"], float)
    ang_step = 360.0 / PSI
    off_step = max(float(np.min(np.abs(tx[tx != 0]))) if np.any(tx != 0) else 1.0, 1e-6)
    keys = np.full((n, cap), -1, np.int32)
    R = D * PSI
    for c0 in range(0, n, chunk):
        c1 = min(n, c0 + chunk)
        Rt = np.stack([euler_matrix(*truth[i, 1:4]) if prob["is_3d"] else euler_matrix(0.0, 0.0, truth[i, 3])
                       for i in range(c0, c1)])
        cosang = np.clip((np.einsum("nij,rij->nr", Rt, Rc) - 1.0) / 2.0, -1.0, 1.0)
        ang = np.degrees(np.arccos(cosang)) / ang_step  # [chunk][R]
        for j, i in enumerate(range(c0, c1)):
            c = int(counts[i])
            d_off = np.hypot(tx - truth[i, 4], ty - truth[i, 5]) / off_step
            nr = min(R, max(16, 4 * c))
            rsel = np.argpartition(ang[j], nr - 1)[:nr]
            score = ang[j][rsel][:, None] ** 2 + d_off[None, :] ** 2  # [nr][T]
            flat = np.argpartition(score.ravel(), c - 1)[:c]
            r_idx, t_idx = rsel[flat // T], flat % T
            k = int(truth[i, 0])
            keys[i, :c] = np.sort(((k * R + r_idx) * T + t_idx).astype(np.int64)).astype(np.int32)
    return keys, counts.astype(np.int32)


def clustered_significance_device(prob, truth, counts, cap, chunk=64):
    """Device (torch) form of clustered_significance for large pools: int32 keys [n][cap] (unused -1) and
    counts [n] as cuda tensors.  Ties in the score may pick other samples than the host form."""
    import torch
    dev = torch.device("cuda:0")
    K, D, PSI, T = prob["nr_classes"], prob["nr_dir"], prob["nr_psi"], prob["nr_trans"]
    R = D * PSI
    pc = np.asarray(prob["pose_coarse"], float).reshape(-1, 3)
    Rc = torch.tensor(np.stack([euler_matrix(*a) for a in pc]).reshape(R, 9), dtype=torch.float32, device=dev)
    tx = torch.tensor(np.asarray(prob["trans_x"], float), dtype=torch.float32, device=dev)
    ty = torch.tensor(np.asarray(prob["trans_y"], float), dtype=torch.float32, device=dev)
    txn = np.asarray(prob["trans_x"], float)
    off_step = float(np.min(np.abs(txn[txn != 0]))) if np.any(txn != 0) else 1.0
    ang_step = 360.0 / PSI
    n = len(counts)
    counts = np.minimum(np.asarray(counts, np.int64), min(cap, R * T))
    keys = torch.full((n, cap), -1, dtype=torch.int32, device=dev)
    tr = np.asarray(truth, float)
    for c0 in range(0, n, chunk):
        c1 = min(n, c0 + chunk)
        Rt = torch.tensor(np.stack([euler_matrix(*tr[i, 1:4]) for i in range(c0, c1)]).reshape(-1, 9),
                          dtype=torch.float32, device=dev)
        cosang = ((Rt @ Rc.T - 1.0) / 2.0).clamp(-1.0, 1.0)
        ang = torch.rad2deg(torch.arccos(cosang)) / ang_step  # [c][R]
        cmax = int(counts[c0:c1].max())
        nr = min(R, max(16, 4 * cmax))
        rang, rsel = torch.topk(ang, nr, dim=1, largest=False)
        sx = torch.tensor(tr[c0:c1, 4], dtype=torch.float32, device=dev)[:, None]
        sy = torch.tensor(tr[c0:c1, 5], dtype=torch.float32, device=dev)[:, None]
        doff = torch.hypot(tx[None, :] - sx, ty[None, :] - sy) / off_step  # [c][T]
        score = (rang[:, :, None] ** 2 + doff[:, None, :] ** 2).reshape(c1 - c0, -1)
        _, idx = torch.topk(score, cmax, dim=1, largest=False, sorted=True)
        r = torch.gather(rsel, 1, idx // T)
        t = idx % T
        k = torch.tensor(tr[c0:c1, 0].astype(np.int64), device=dev)[:, None]
        kk = ((k * R + r) * T + t).to(torch.int32)
        keep = torch.arange(cmax, device=dev)[None, :] < torch.tensor(counts[c0:c1], device=dev)[:, None]
        kk = torch.where(keep, kk, torch.full_like(kk, -1))
        keys[c0:c1, :cmax] = kk
    return keys, torch.tensor(counts.astype(np.int32), device=dev)


def significance_mask(prob, keys, counts):
    """The dense coarse significance [n][K][D][PSI][T] (uint8) of per-image keys."""
    K, D, PSI, T = prob["nr_classes"], prob["nr_dir"], prob["nr_psi"], prob["nr_trans"]
    n = len(counts)
    mask = np.zeros((n, K * D * PSI * T), np.uint8)
    for i in range(n):
        mask[i, keys[i, :counts[i]]] = 1
    return mask.reshape(n, K, D, PSI, T)


def engine_config(prob, precision="fp32"):
    keys = ("ori_size", "current_size", "coarse_size", "nr_classes", "nr_groups", "adaptive_oversampling",
            "nr_dir", "nr_psi", "nr_trans", "adaptive_fraction", "sigma2_fudge", "sigma2_offset",
            "avg_norm_correction", "do_ctf_correction", "refs_are_ctf_corrected", "do_norm_correction", "do_map")
    cfg = {k: prob[k] for k in keys}
    cfg["ref_dim"] = 3 if prob["is_3d"] else 2
    cfg["precision"] = PRECISIONS[precision]
    return cfg


def load_engine_tables(engine, prob, precision="fp32"):
    """configure + sampling tables + reference grids (everything but the model)."""
    engine.configure(**engine_config(prob, precision))
    engine.set_sampling(prob["pose_coarse"], prob["pose_fine"], prob["shift_coarse_x"], prob["shift_coarse_y"],
                        prob["shift_fine_x"], prob["shift_fine_y"], prob["trans_x"], prob["trans_y"])
    for k in range(prob["nr_classes"]):
        engine.set_reference(k, prob["grids"][k])


def load_engine(engine, prob, precision="fp32"):
    """Configure an Engine with a problem dict (same tables the oracle receives); precision is the
    arithmetic of the coarse cross term ("fp32", "bf16" or "fp16")."""
    load_engine_tables(engine, prob, precision)
    engine.set_model(prob["sigma2_noise"], prob["pdf_class"], prob["pdf_direction"], prob["prior_offx"],
                     prob["prior_offy"])
    engine.reset_sums()


def real_space_images(F, ori):
    """Real-space images [n][ori][ori] whose centred, 1/ori^2-normalised transform is the half-plane F
    [n][ori * (ori/2 + 1)] (the convention of ParticleModel::setFourierTransforms: CenterFFT, then the r2c
    transform divided by ori^2); the x = 0 column takes its Hermitian-symmetric part."""
    F = np.asarray(F).reshape(-1, ori, ori // 2 + 1)
    img = np.fft.irfft2(F * float(ori * ori), s=(ori, ori))
    s_ = -(ori // 2)
    return np.roll(np.roll(img, s_, axis=2), s_, axis=1)


def write_dataset(directory, prefix, batch, truth, ori, angpix=None, stack_size=10000):
    """A synthetic pool as the reference reads it: prefix.star (image names, micrograph, group, CTF
    parameters, the true angles, origins and class) and prefix_NNN.mrcs stacks of real-space float32
    images.  batch / truth as make_problem returns them (current_size = ori_size)."""
    import os
    from . import starfile
    os.makedirs(directory, exist_ok=True)
    n = int(batch["n"])
    apx = float(angpix if angpix is not None else batch.get("angpix", 1.0))
    F = np.asarray(batch["fnomask_re"]) + 1j * np.asarray(batch["fnomask_im"])
    names = []
    for s0 in range(0, n, stack_size):
        s1 = min(n, s0 + stack_size)
        stack = f"{prefix}_{s0 // stack_size:03d}.mrcs"
        starfile.write_mrcs(os.path.join(directory, stack), real_space_images(F[s0:s1], ori), apx)
        names += [starfile.image_name(i - s0 + 1, stack) for i in range(s0, s1)]
    tr = np.asarray(truth, dtype=float)
    cp = np.asarray(batch.get("ctf_params", np.zeros((n, 7))))
    cols = {"rlnImageName": names, "rlnMicrographName": ["mic1"] * n,
            "rlnGroupNumber": [int(g) + 1 for g in batch["group"]],
            "rlnDefocusU": cp[:, 0], "rlnDefocusV": cp[:, 1], "rlnDefocusAngle": cp[:, 2], "rlnVoltage": cp[:, 3],
            "rlnSphericalAberration": cp[:, 4], "rlnAmplitudeContrast": cp[:, 5], "rlnBfactor": cp[:, 6],
            "rlnAngleRot": tr[:, 1], "rlnAngleTilt": tr[:, 2], "rlnAnglePsi": tr[:, 3],
            "rlnOriginX": tr[:, 4], "rlnOriginY": tr[:, 5], "rlnClassNumber": [int(k) + 1 for k in tr[:, 0]]}
    path = os.path.join(directory, prefix + ".star")
    starfile.write_star(path, cols)
    return path
