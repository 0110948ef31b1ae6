"""The C-ABI library loads without a GPU and exports every symbol include/rome_gpu.h declares."""
import os
import re

import rome_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "rome_gpu.h")).read()
    return sorted(set(re.findall(r"\b(rome_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported():
    L = rome_amd.lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), f"{s} declared in rome_gpu.h but not exported"
    assert set(syms) == set(rome_amd.EXPORTED)


def test_host_sampling_without_gpu():
    tx, ty = rome_amd.translations(5, 1)
    assert len(tx) == 81
    z, phi = rome_amd.healpix_pix2ang(2, 17)
    assert -1 < z < 1 and 0 <= phi < 6.3
