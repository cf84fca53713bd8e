"""The C ABI from C: tests/abi_client.c, a C99 program that includes only include/cutcell_hip.h and
links libcutcell_hip.so, as a Julia ccall binding (a HIPBackend tag) binds it.  The CPU
test builds it (warnings are errors) and runs its device-free mode; the GPU test runs an
allocate_cache / update_cache! / readers / error-path session and compares the columns it copied
out with the oracle bit for bit."""
import os
import subprocess

import numpy as np
import pytest

import oracle as orc
from cutcell import SDFSphere

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "cutcellmethods.jl_amd")


def build_client(tmp_path):
    exe = str(tmp_path / "abi_client")
    cmd = ["gcc", "-std=c99", "-O1", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "abi_client.c"), "-o", exe, "-L", PKG, "-l:libcutcell_hip.so",
           f"-Wl,-rpath,{PKG}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_a_c_caller_builds_against_the_header_alone(tmp_path):
    exe = build_client(tmp_path)
    r = subprocess.run([exe, "version"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "gfx950" in r.stdout


@pytest.mark.gpu
def test_a_c_caller_updates_a_cache_bitwise_the_oracle(tmp_path):
    exe = build_client(tmp_path)
    out = tmp_path / "cols.bin"
    print("abi_client run", flush=True)
    r = subprocess.run([exe, "run", str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    n, nn = 24 ** 3, 25 ** 3
    raw = out.read_bytes()
    assert len(raw) == n + 8 * n + 48 * n + 8 * nn
    kind = np.frombuffer(raw, np.int8, n, 0)
    vf = np.frombuffer(raw, np.float64, n, n)
    ff = np.frombuffer(raw, np.float64, 6 * n, 9 * n).reshape(n, 6)
    phi = np.frombuffer(raw, np.float64, nn, 57 * n)
    geo = SDFSphere((0.01, 0.02, -0.03), 1.0)
    x0, d, ncell = (-1.5,) * 3, (0.125,) * 3, (24, 24, 24)
    phi_ref = orc.sample_nodes(geo, 3, x0, d, ncell)
    ref = orc.mc_cells(phi_ref, x0, d, ncell)
    assert np.array_equal(phi.view(np.uint64), phi_ref.view(np.uint64))
    assert np.array_equal(kind, ref["kind"])
    assert np.array_equal(vf.view(np.uint64), ref["volume_fraction"].view(np.uint64))
    assert np.array_equal(ff.view(np.uint64), ref["face_fraction"].view(np.uint64))
