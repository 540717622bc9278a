"""The host-side Hermitian eigen-decomposition behind the min-max
time-segmentation interpolator (powergrid_amd/csrc/pg_herm.h), compiled with
g++ and checked against numpy's eigh: eigenvalues and the reconstruction
V diag(w) V^H.  CPU only."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "herm_eig_test.cpp")


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("herm") / "herm_eig_test")
    subprocess.check_call(["g++", "-std=c++17", "-O2", "-Wall", SRC, "-o", out])
    return out


def run(exe, A, tmp_path):
    n = A.shape[0]
    fi, fo = tmp_path / "a.bin", tmp_path / "o.bin"
    with open(fi, "wb") as f:
        np.array([n], np.int32).tofile(f)
        np.ascontiguousarray(A, np.complex128).tofile(f)
    subprocess.check_call([exe, str(fi), str(fo)])
    w = np.fromfile(fo, np.float64, count=n)
    V = np.fromfile(fo, np.complex128, offset=8 * n).reshape(n, n)
    return w, V


@pytest.mark.parametrize("n,seed", [(2, 0), (5, 1), (8, 2), (16, 3)])
def test_herm_eig_matches_eigh(exe, tmp_path, n, seed):
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = B @ B.conj().T + np.diag(rng.uniform(0, 1, n))
    w, V = run(exe, A, tmp_path)
    assert np.allclose(np.sort(w), np.linalg.eigvalsh(A), rtol=1e-12, atol=1e-12 * np.abs(w).max())
    assert np.abs(V @ np.diag(w) @ V.conj().T - A).max() < 1e-12 * np.abs(A).max()
    assert np.abs(V.conj().T @ V - np.eye(n)).max() < 1e-12


def test_herm_eig_gram_of_unit_circle_nodes(exe, tmp_path):
    # the interpolator's Gram matrix G^H G (Vandermonde on unit-circle nodes),
    # ill-conditioned: relative accuracy of the largest eigenvalues, and the
    # smallest within eps * max
    rng = np.random.default_rng(7)
    g = np.exp(1j * rng.uniform(-2.4, 2.4, 400))
    G = np.stack([g ** c for c in range(8)], axis=1)
    A = G.conj().T @ G
    w, V = run(exe, A, tmp_path)
    ref = np.linalg.eigvalsh(A)
    ws = np.sort(w)
    assert np.all(np.abs(ws - ref) <= 1e-13 * ref.max())
    assert np.abs(V @ np.diag(w) @ V.conj().T - A).max() < 1e-12 * np.abs(A).max()
