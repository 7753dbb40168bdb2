"""The in-process boundary: cosmo::GpuMatrix, the adapter a CoSMo build hands to
MyMatrixDiagonal (include/kdav/MandD.cxx:7-42), and the Davidson defaults of the
reference driver."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _compile(tmp_path):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    exe = str(tmp_path / "matrixbase_check")
    lib = os.path.join(ROOT, "cosmo_amd")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    os.path.join(HERE, "cpp", "matrixbase_check.cpp"), "-L", lib, "-lcosmo_hip",
                    "-Wl,-rpath," + lib, "-o", exe], check=True)
    return exe


def test_adapter_compiles_against_the_contract(tmp_path):
    """The header-only adapter and the test program that exercises it build with the host
    compiler alone (no HIP toolchain needed by a CoSMo build)."""
    if not os.path.exists(os.path.join(ROOT, "cosmo_amd", "libcosmo_hip.so")):
        pytest.skip("library not built")
    _compile(tmp_path)


def test_davidson_defaults_are_the_reference_driver_values():
    """cosmo_dav_defaults(nume) gives DavidsonDiagonalization's settings
    (include/kdav/davidson.cxx:186-202): crite 1e-15, critc 1e-12, critr 1e-8, ortho 1e-9,
    numemax = mblock = nstates, lim = numemax + 20, maxiter = max(40 numemax, 200)."""
    import cosmo_amd as ca
    for nume in (1, 4, 5, 10, 16, 40):
        p = ca.dav_params(nume)
        assert (p.nume, p.mblock, p.lim, p.maxiter) == (nume, nume, nume + 20, max(40 * nume, 200))
        assert (p.crite, p.critc, p.critr, p.ortho) == (1e-15, 1e-12, 1e-8, 1e-9)
        assert p.precision == ca.PREC_FP64


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["sd2p2n_s11", "fp2p2n_s5"])
def test_adapter_products_match_the_oracle(gpu, golden, name, tmp_path):
    """Through the adapter, both TimesMatrix forms (pointer arrays, and contiguous vectors on
    a copy of the matrix) equal the oracle's CSB product to 1e-12, and ReadDiagonal gives
    the diagonal of the same H."""
    g = golden(name)
    exe = _compile(tmp_path)
    out = str(tmp_path / "mb.bin")
    r = subprocess.run([exe, g.xml, name, "3", out], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    raw = open(out, "rb").read()
    n, m = np.frombuffer(raw, np.uint64, 2)
    n, m = int(n), int(m)
    a = np.frombuffer(raw, np.float64, offset=16)
    diag, X, Y1, Y2 = a[:n], a[n:n + n * m], a[n + n * m:n + 2 * n * m], a[n + 2 * n * m:n + 3 * n * m]
    X, Y1, Y2 = (z.reshape(m, n).T for z in (X, Y1, Y2))
    o = oracle.System(g.xml)
    rp, c, v = o.h_upper(drop_small=True)
    ref = oracle.csb_times(n, rp, c, v, X)
    scale = np.abs(ref).max()
    assert np.max(np.abs(Y1 - ref)) < 1e-12 * scale
    assert np.array_equal(Y1, Y2)
    H = oracle.dense_from_upper(n, rp, c, v) if n <= 5000 else None
    if H is not None:
        assert np.max(np.abs(diag - np.diag(H))) < 1e-12 * np.abs(np.diag(H)).max()
