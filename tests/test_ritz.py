"""The Rayleigh-Ritz eigensolver of the host Davidson driver (ritz.cpp), checked against
numpy's symmetric eigensolver: eigenvalues, residuals and orthogonality of the wanted
lowest pairs, on random, clustered, exactly degenerate, diagonal and zero matrices."""
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "..", "cosmo_amd", "csrc", "ritz.cpp")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("ritz")
    exe = str(d / "ritz_check")
    subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(HERE, "cpp", "ritz_check.cpp"), SRC, "-o", exe],
                   check=True)
    return exe, d


def _cases(rng):
    out = []
    for k in (1, 2, 3, 5, 17, 36, 60, 100, 128):
        A = rng.standard_normal((k, k))
        out.append(A + A.T)
    for k in (12, 40, 90):
        Q, _ = np.linalg.qr(rng.standard_normal((k, k)))
        lam = np.repeat(np.arange(k // 3 + 1, dtype=float), 3)[:k]         # triple degeneracies
        out.append((Q * lam) @ Q.T)
        lam2 = np.sort(np.concatenate([1.0 + 1e-9 * np.arange(k // 2), 2.0 + rng.random(k - k // 2)]))
        out.append((Q * lam2) @ Q.T)                                        # a tight cluster
    out.append(np.diag(rng.permutation(np.repeat([3.0, -1.0, 0.5], 7))))   # diagonal, repeated
    out.append(np.zeros((6, 6)))
    out.append(np.diag(np.full(9, 2.5)) + 1e-300 * np.ones((9, 9)))
    return [0.5 * (A + A.T) for A in out]


def test_ritz_lowest_against_eigh(driver):
    exe, d = driver
    rng = np.random.default_rng(11)
    cases = _cases(rng)
    wants = [min(A.shape[0], w) for A, w in zip(cases, rng.integers(1, 17, len(cases)))]
    with open(d / "in.bin", "wb") as f:
        f.write(np.int32(len(cases)).tobytes())
        for A, w in zip(cases, wants):
            f.write(np.int32(A.shape[0]).tobytes() + np.int32(w).tobytes() + np.ascontiguousarray(A).tobytes())
    subprocess.run([exe, str(d / "in.bin"), str(d / "out.bin")], check=True)
    raw = open(d / "out.bin", "rb").read()
    pos = 0
    for A, w in zip(cases, wants):
        k = A.shape[0]
        ok = np.frombuffer(raw, np.int32, 1, pos)[0]
        pos += 4
        ev = np.frombuffer(raw, np.float64, w, pos)
        pos += 8 * w
        V = np.frombuffer(raw, np.float64, w * k, pos).reshape(w, k)
        pos += 8 * w * k
        assert ok == 1
        scale = max(np.abs(A).max(), 1e-12)
        ref = np.linalg.eigvalsh(A)[:w]
        assert np.max(np.abs(ev - ref)) <= 1e-12 * scale, (k, ev, ref)
        res = A @ V.T - V.T * ev
        assert np.max(np.abs(res)) <= 1e-12 * scale
        assert np.max(np.abs(V @ V.T - np.eye(w))) <= 1e-12
