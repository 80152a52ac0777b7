"""GMRES cores (FP64) on the CPU: the restarted matrix-free core (C-ABI
bit_gmres_dense) and the reference-interface class solver.h, against LU."""
import os
import subprocess

import numpy as np
import pytest

import bit_amd
from conftest import ROOT

BIN = os.path.join(ROOT, "tests", "cpp", "bin", "solver_dense")


def _system(n, seed, shift=4.0):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) / np.sqrt(n) + shift * np.eye(n)
    b = rng.standard_normal(n)
    return A, b


@pytest.mark.parametrize("n,restart", [(40, 5), (120, 20), (200, 200)])
def test_restarted_gmres_matches_lu(n, restart):
    A, b = _system(n, n)
    x, it, rr = bit_amd.gmres_dense(A, b, restart=restart, tol=1e-12, maxit=20 * n)
    ref = np.linalg.solve(A, b)
    assert rr <= 1e-12
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-10
    true_rr = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    assert true_rr < 1e-11  # the Givens estimate is the true residual


def test_gmres_exact_in_n_steps_and_zero_rhs():
    A, b = _system(30, 3, shift=0.3)  # not diagonally dominant
    x, it, rr = bit_amd.gmres_dense(A, b, restart=30, tol=1e-13, maxit=30)
    assert it <= 30 and np.allclose(A @ x, b, atol=1e-10)
    x0, it0, rr0 = bit_amd.gmres_dense(A, np.zeros(30))
    assert it0 == 0 and not x0.any()


def test_gmres_respects_maxit():
    A, b = _system(100, 5, shift=0.05)
    x, it, rr = bit_amd.gmres_dense(A, b, restart=5, tol=1e-14, maxit=7)
    assert it == 7 and rr > 1e-14


def test_solver_h_reference_signature(tmp_path):
    """solver::gmres(a, b, x, num) with the reference's stopping rule (max
    relative change of x < 1e-5) and the matrix-free overload."""
    n = 60
    A, b = _system(n, 9)
    inp = tmp_path / "in.bin"
    np.concatenate([[n], A.ravel(), b]).astype(np.float64).tofile(inp)
    out = tmp_path / "out.bin"
    r = subprocess.run([BIN, str(inp), str(out)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    x = np.fromfile(out).reshape(2, n)
    ref = np.linalg.solve(A, b)
    assert np.linalg.norm(x[0] - ref) / np.linalg.norm(ref) < 1e-4
    assert np.linalg.norm(x[1] - ref) / np.linalg.norm(ref) < 1e-10
