"""A compiled C99 caller of include/rome_gpu.h (tests/capi_driver.c): the header compiles as C with
-Wall -Wextra -Werror -pedantic, the program links against librome_gpu.so, and it runs, host-only here
and with one MAP3D pool, the Fourier map, the reconstruction and the model update on the GPU."""
import os
import shutil
import subprocess

import pytest

import rome_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "rome_amd")


def _build(tmp_path):
    rome_amd.lib()  # the library is built
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    exe = str(tmp_path / "capi_driver")
    cmd = [cc, "-std=c99", "-O1", "-Wall", "-Wextra", "-Werror", "-pedantic", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "capi_driver.c"), "-o", exe, "-L", LIBDIR, "-lrome_gpu",
           "-Wl,-rpath," + LIBDIR, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_c_caller_host(tmp_path):
    exe = _build(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "capi_driver ok" in r.stdout


@pytest.mark.gpu
def test_c_caller_gpu(tmp_path):
    exe = _build(tmp_path)
    print("[capi_driver gpu]", flush=True)
    r = subprocess.run([exe, "gpu"], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "gpu calls ok" in r.stdout
