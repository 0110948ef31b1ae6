"""STAR metadata and MRC / MRCS image files in the layouts the reference reads and writes.

* STAR (MetaDataTable::writeToStar / printTable, src/metadata.cpp:343-395): ``data_`` then ``loop_`` with
  one ``_rlnLabel #column`` line per column, then one whitespace-separated row per particle.  Image names
  are ``NNNNNNN@stack.mrcs`` (7-digit 1-based index, src/metadata.h:51).
* MRC / MRCS (Mrcs::writeMrcsData / writeMrcData, src/mrcs.cpp:127-236; header src/mrcs.h:53-120): a
  1024-byte header of 256 32-bit words (NC, NR, NS, MODE = 2 for float32, ..., cell lengths, MAPC/R/S,
  AMIN/AMAX/AMEAN, NSYMBT, 'MAP ' stamp and machine stamp) followed by little-endian float32 data, images
  (or z sections) one after the other, x fastest.
"""
import os

import numpy as np

# The STAR labels of the per-particle fields the E-step reads and writes (MetaDataLabel names,
# src/metadata.h:52-108), with their numeric kind.
STAR_LABELS = {
    "rlnImageName": str, "rlnMicrographName": str, "rlnGroupNumber": int, "rlnGroupName": str,
    "rlnVoltage": float, "rlnDefocusU": float, "rlnDefocusV": float, "rlnDefocusAngle": float,
    "rlnSphericalAberration": float, "rlnAmplitudeContrast": float, "rlnBfactor": float,
    "rlnAngleRot": float, "rlnAngleTilt": float, "rlnAnglePsi": float, "rlnOriginX": float, "rlnOriginY": float,
    "rlnClassNumber": int, "rlnNormCorrection": float, "rlnLogLikeliContribution": float,
    "rlnMaxValueProbDistribution": float, "rlnNrOfSignificantSamples": int,
}


def image_name(index, stack):
    """STAR image reference of the 1-based index in a stack file."""
    return f"{int(index):07d}@{stack}"


def parse_image_name(name):
    idx, stack = name.split("@", 1)
    return int(idx), stack


def write_star(path, columns):
    """columns: ordered mapping label (without the leading underscore) -> sequence, all of one length."""
    labels = list(columns)
    n = len(columns[labels[0]]) if labels else 0
    for k in labels:
        if len(columns[k]) != n:
            raise ValueError(f"column {k} has {len(columns[k])} rows, expected {n}")
    with open(path, "w") as f:
        f.write("\ndata_\n\nloop_\n")
        for i, k in enumerate(labels):
            f.write(f"_{k} #{i + 1}\n")
        for r in range(n):
            cells = []
            for k in labels:
                v = columns[k][r]
                if isinstance(v, (float, np.floating)):
                    cells.append(f"{float(v):12.6f}")
                else:
                    cells.append(f"{v!s:>12}")
            f.write("\t".join(cells) + "\n")
        f.write("\n")


def read_star(path):
    """The first loop of a STAR file as label -> numpy array (numbers) or list (strings)."""
    labels, rows = [], []
    in_loop = False
    with open(path) as f:
        for raw in f:
            line = raw.strip()
            if not line or line.startswith("#"):
                if in_loop and rows and not line:
                    break
                continue
            if line.startswith("data_"):
                continue
            if line == "loop_":
                in_loop = True
                continue
            if line.startswith("_"):
                labels.append(line[1:].split()[0])
                continue
            if in_loop and labels:
                rows.append(line.split())
    out = {}
    for i, k in enumerate(labels):
        col = [r[i] for r in rows]
        kind = STAR_LABELS.get(k)
        if kind is str:
            out[k] = col
            continue
        try:
            arr = np.array(col, dtype=np.float64)
            out[k] = arr.astype(np.int64) if kind is int else arr
        except ValueError:
            out[k] = col
    return out


def _header(nx, ny, nz, pixel_size, data, is_volume):
    h = np.zeros(256, dtype=np.int32)
    hf = h.view(np.float32)
    h[0], h[1], h[2] = nx, ny, nz     # NC, NR, NS
    h[3] = 2                           # MODE: float32
    h[7], h[8], h[9] = nx, ny, nz     # NX, NY, NZ (sampling intervals)
    hf[10] = nx * pixel_size           # cell lengths (A)
    hf[11] = ny * pixel_size
    hf[12] = nz * pixel_size
    hf[13] = hf[14] = hf[15] = 90.0   # cell angles
    h[16], h[17], h[18] = 1, 2, 3     # MAPC, MAPR, MAPS
    if data.size:
        hf[19], hf[20], hf[21] = float(data.min()), float(data.max()), float(data.mean())
    h[22] = 1 if is_volume else 0     # ISPG
    h[23] = 0                          # NSYMBT
    h[52] = int.from_bytes(b"MAP ", "little")
    h[53] = 0x00004144                 # machine stamp: little-endian
    return h


def write_mrcs(path, stack, pixel_size=1.0):
    """Image stack [n][ny][nx] as float32 (MODE 2)."""
    a = np.ascontiguousarray(stack, dtype="<f4")
    n, ny, nx = a.shape
    with open(path, "wb") as f:
        f.write(_header(nx, ny, n, pixel_size, a, False).tobytes())
        f.write(a.tobytes())


def write_mrc(path, vol, pixel_size=1.0):
    """Volume [nz][ny][nx] as float32 (MODE 2)."""
    a = np.ascontiguousarray(vol, dtype="<f4")
    nz, ny, nx = a.shape
    with open(path, "wb") as f:
        f.write(_header(nx, ny, nz, pixel_size, a, True).tobytes())
        f.write(a.tobytes())


def read_mrc_header(path):
    h = np.fromfile(path, dtype="<i4", count=256)
    if h.size < 256:
        raise ValueError(f"{path}: short MRC header")
    if h[3] != 2:
        raise ValueError(f"{path}: MODE {h[3]} (only float32, MODE 2, is read)")
    return {"nx": int(h[0]), "ny": int(h[1]), "nz": int(h[2]), "nsymbt": int(h[23]),
            "pixel_size": float(h.view(np.float32)[10] / max(int(h[7]), 1))}


def read_mrcs(path, indices=None):
    """Images [n][ny][nx] (all, or the 1-based indices given) as float64, memory-mapped reads."""
    hd = read_mrc_header(path)
    nx, ny, nz = hd["nx"], hd["ny"], hd["nz"]
    m = np.memmap(path, dtype="<f4", mode="r", offset=1024 + hd["nsymbt"], shape=(nz, ny, nx))
    if indices is None:
        return np.array(m, dtype=np.float64)
    return np.array(m[np.asarray(indices, dtype=np.int64) - 1], dtype=np.float64)


def read_mrc(path):
    return read_mrcs(path)


def stack_images(star, base_dir="."):
    """The images a STAR table names, in table order: [n][ny][nx] float64."""
    names = star["rlnImageName"]
    by_stack = {}
    for r, nm in enumerate(names):
        idx, stack = parse_image_name(nm)
        by_stack.setdefault(stack, []).append((r, idx))
    out = None
    for stack, items in by_stack.items():
        p = stack if os.path.isabs(stack) else os.path.join(base_dir, stack)
        imgs = read_mrcs(p, [i for _, i in items])
        if out is None:
            out = np.zeros((len(names),) + imgs.shape[1:])
        out[[r for r, _ in items]] = imgs
    return out
