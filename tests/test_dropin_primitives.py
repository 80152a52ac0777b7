"""The drop-in's host primitives (no device): one cluster's Taylor sum
(shielded_treetaylorpot, treetaylorpotential; tree3d.cpp:3004-3227, 4468-4640)
and field (treetaylorforce; tree3d.cpp:456-810) against the oracle's Taylor
coefficients, which the independent power series
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
    res = np.fromfile(out).reshape(len(cases), 11)
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
        # treetaylorforce: E_e = sum_{|k| < nterms} (k_e + 1) M_k a_{k+e} with the
        # oracle's signed coefficients a at degree nterms (Coulomb, rs includes DEL)
        a1 = oracle.taylor_coeffs(nterms + 1, *d, rs, 0.0)
        for e in range(3):
            sh = [slice(0, nterms)] * 3
            sh[e] = slice(1, nterms + 1)
            kk = np.arange(nterms).reshape([-1 if x == e else 1 for x in range(3)]) + 1.0
            terms = (mom * kk * a1[tuple(sh)])[tri]
            assert abs(row[8 + e] - np.sum(terms)) <= 1e-12 * np.sum(np.abs(terms))


DRIVER = r"""
#include <cstdio>
#include <vector>
#include "tree3d.h"
// IN: nterms, ncase, then per case dx, dy, dz, rs and the moment cube; OUT: the
// field of treetaylorforce (3 per case, from 0)
int main(int argc, char** argv) {
  FILE* f = fopen(argv[1], "rb");
  double hdr[2];
  if (fread(hdr, sizeof(double), 2, f) != 2) return 2;
  const int nt = (int)hdr[0], ncase = (int)hdr[1];
  std::vector<double> cube((size_t)nt * nt * nt), res;
  std::vector<double*> rows((size_t)nt * nt);
  std::vector<double**> planes(nt);
  for (int i = 0; i < nt; i++) {
    planes[i] = &rows[(size_t)i * nt];
    for (int j = 0; j < nt; j++) rows[(size_t)i * nt + j] = &cube[((size_t)i * nt + j) * nt];
  }
  for (int c = 0; c < ncase; c++) {
    double d[4];
    if (fread(d, sizeof(double), 4, f) != 4) return 2;
    if (fread(cube.data(), sizeof(double), cube.size(), f) != cube.size()) return 2;
    double e[3] = {0.0, 0.0, 0.0};
    treetaylorforce(nt, d[0], d[1], d[2], d[3], &e[0], &e[1], &e[2], planes.data());
    res.insert(res.end(), e, e + 3);
  }
  fclose(f);
  FILE* o = fopen(argv[2], "wb");
  fwrite(res.data(), sizeof(double), res.size(), o);
  fclose(o);
  return 0;
}
"""


@pytest.mark.parametrize("korder,nterms", [(1, 6), (2, 9), (3, 12)])
def test_treetaylorforce_kernel_family(tmp_path, oracle, korder, nterms):
    """treetaylorforce with KERNEL_ORDER = L >= 1 (a compile-time constant of
    problem.h, as in the reference): the drop-in's sources are built in a
    scratch directory with that value, and the field of one cluster equals
    sum_{|k| < nterms} (k_e + 1) M_k a_{k+e} with a = sum_l w_l DEL^l times the
    oracle's coefficients of R^-(2l+1) (pinned by tests/test_oracle_math.py)."""
    import shutil
    src = os.path.join(ROOT, "bit_amd", "src")
    pkg = os.path.join(ROOT, "bit_amd")
    if not os.path.exists(os.path.join(pkg, "libbit_amd.so")):
        pytest.fail("bit_amd/libbit_amd.so not built: run python -m bit_amd.build")
    for name in ("tree3d.cpp", "tree3d.h", "constants.h", "solver.h"):
        shutil.copy(os.path.join(src, name), tmp_path / name)
    prob = open(os.path.join(src, "problem.h")).read()
    assert "const int KERNEL_ORDER = 0;" in prob
    (tmp_path / "problem.h").write_text(
        prob.replace("const int KERNEL_ORDER = 0;", f"const int KERNEL_ORDER = {korder};"))
    (tmp_path / "drv.cpp").write_text(DRIVER)
    exe = tmp_path / "drv"
    cmd = ["g++", "-O1", "-std=c++17", "-fopenmp", "-I" + str(tmp_path),
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(pkg, "csrc"),
           str(tmp_path / "drv.cpp"), str(tmp_path / "tree3d.cpp"), "-o", str(exe),
           "-L" + pkg, "-lbit_amd", "-Wl,-rpath," + pkg]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    DEL = 0.01
    rng = np.random.default_rng(korder)
    cases = []
    for c in range(6):
        d = rng.uniform(-1.0, 1.0, 3)
        d *= rng.uniform(0.3, 1.5) / np.linalg.norm(d)
        mom = rng.uniform(-1.0, 1.0, (nterms, nterms, nterms)) * 0.3 ** np.add.outer(
            np.add.outer(np.arange(nterms), np.arange(nterms)), np.arange(nterms))
        cases.append((d, float(d @ d) + DEL, mom))
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    np.concatenate([np.array([nterms, len(cases)], dtype=np.float64)] +
                   [np.concatenate([d, [rs], mom.ravel()]) for d, rs, mom in cases]).tofile(inp)
    r = subprocess.run([str(exe), str(inp), str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    res = np.fromfile(out).reshape(len(cases), 3)
    tri = np.add.outer(np.add.outer(np.arange(nterms), np.arange(nterms)),
                       np.arange(nterms)) < nterms
    for (d, rs, mom), row in zip(cases, res):
        a = 0.0
        w = 1.0
        for l in range(korder + 1):
            if l > 0:
                w *= (l - 0.5) / l * DEL
            a = a + w * oracle.powinv_coeffs(nterms + 1, l, *d, rs)
        for e in range(3):
            sh = [slice(0, nterms)] * 3
            sh[e] = slice(1, nterms + 1)
            kk = np.arange(nterms).reshape([-1 if x == e else 1 for x in range(3)]) + 1.0
            terms = (mom * kk * a[tuple(sh)])[tri]
            assert abs(row[e] - np.sum(terms)) <= 1e-12 * np.sum(np.abs(terms)), (korder, e)
