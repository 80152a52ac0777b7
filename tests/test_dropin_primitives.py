"""The drop-in's host primitives (no device): one cluster's Taylor sum
(shielded_treetaylorpot, treetaylorpotential; tree3d.cpp:3004-3227, 4468-4640)
against the oracle's Taylor coefficients, which the independent power series
pins (tests/test_oracle_math.py), and the one-pair gradients
(eval_grad_kernel_noreg/_reg; tree3d.cpp:40-97) against their closed form."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rel_max

BIN = os.path.join(ROOT, "tests", "cpp", "bin", "dropin_treecode")


@pytest.mark.parametrize("nterms,kappa", [(1, 0.5), (5, 0.0), (9, 0.7), (12, 2.0)])
def test_cluster_primitives_match_oracle(tmp_path, oracle, nterms, kappa):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} not built: run python -m bit_amd.build")
    rng = np.random.default_rng(nterms)
    cases = []
    for c in range(12):
        d = rng.uniform(-1.0, 1.0, 3)
        d *= rng.uniform(0.5, 2.0) / np.linalg.norm(d)
        rs = float(d @ d) + (0.0 if c % 2 else 1e-4)
        mom = rng.uniform(-1.0, 1.0, (nterms, nterms, nterms)) * 0.3 ** np.add.outer(
            np.add.outer(np.arange(nterms), np.arange(nterms)), np.arange(nterms))
        cases.append((d, rs, mom))
    inp = tmp_path / "in.bin"
    np.concatenate([np.array([nterms, kappa, len(cases)], dtype=np.float64)] +
                   [np.concatenate([d, [rs], mom.ravel()]) for d, rs, mom in cases]).tofile(inp)
    out = tmp_path / "out.bin"
    r = subprocess.run([BIN, "primitives", str(inp), str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    res = np.fromfile(out).reshape(len(cases), 8)
    for (d, rs, mom), row in zip(cases, res):
        tri = np.add.outer(np.add.outer(np.arange(nterms), np.arange(nterms)),
                           np.arange(nterms)) < nterms
        a = oracle.taylor_coeffs(nterms, *d, rs, kappa)[:nterms, :nterms, :nterms]
        a0 = oracle.taylor_coeffs(nterms, *d, rs, 0.0)[:nterms, :nterms, :nterms]
        scale = np.sum(np.abs(mom * a)[tri])
        assert abs(row[0] - (-np.sum((mom * a)[tri]))) <= 1e-12 * scale
        scale0 = np.sum(np.abs(mom * a0)[tri])
        assert abs(row[1] - np.sum((mom * a0)[tri])) <= 1e-12 * scale0
        g = -d / (rs * np.sqrt(rs))  # KERNEL_ORDER = 0: both forms are the Coulomb gradient
        assert rel_max(row[2:5], g) < 1e-14 and rel_max(row[5:8], g) < 1e-14
