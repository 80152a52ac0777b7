"""The drop-in reference interface (bit_amd/libtree3d.so, bit_amd/src/tree3d.h)
driven by a C++ caller written against the reference headers
(tests/cpp/dropin_treecode.cpp), compared with the reference's own outputs."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_fixture, rel_max, subtree_expected

pytestmark = pytest.mark.gpu
BIN = os.path.join(ROOT, "tests", "cpp", "bin", "dropin_treecode")
INTG = -0.0000483811229725
CHARGE_CONSTANT = 0.5 / 3.14159265359 / 8.85418781762e-12  # constants.h


def _run(*args):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} not built: run python -m bit_amd.build")
    r = subprocess.run([BIN, *args], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "Test PASSED" in r.stdout, r.stdout + r.stderr
    return r.stdout


def test_treecode_mp_field(tmp_path):
    """tests/treecode_mp.cpp's recipe: boxparticles(16000), maxmemb 100,
    nterms 5, serial == OpenMP treeforce; values equal the reference's."""
    out = tmp_path / "f.bin"
    _run("field", str(out))
    E = np.fromfile(out).reshape(-1, 3)
    fx = load_fixture("box16k")
    assert rel_max(E[fx["tidx"]], fx["field_tree"]) < 1e-10


def test_treecode_convergence(tmp_path):
    """tests/treecode.cpp's criterion: the max-abs field error against the
    direct sum does not grow with nterms (1..4)."""
    out = tmp_path / "c.bin"
    _run("conv", str(out))
    E = np.fromfile(out).reshape(4, -1, 3)
    fx = load_fixture("box16k")
    ds = fx["field_ds"]
    errs = [float(np.abs(E[k][fx["tidx"]] - ds).max()) for k in range(4)]
    assert all(b <= a for a, b in zip(errs, errs[1:])), errs


def test_shielded_treepot_and_batch_driver(tmp_path):
    fx = load_fixture("blobs12k_screened")
    n = fx["x"].size
    inp = tmp_path / "in.bin"
    hdr = np.array([n, 200, 9, 0.7, 0.5], dtype=np.float64)
    np.concatenate([hdr, fx["x"], fx["y"], fx["z"], fx["q"]]).tofile(inp)
    out = tmp_path / "p.bin"
    _run("screened", str(inp), str(out))
    raw = np.fromfile(out)
    want = -fx["phi_screened"] + fx["q"][fx["tidx"]] * INTG  # the reference's raw value
    assert rel_max(raw[fx["tidx"]], want) < 1e-10


def test_point_calls_p_minus_1(tmp_path, oracle):
    """treepotential / treeforce / shielded_treepot with p = -1 from OpenMP
    threads: at particle positions treepotential equals the reference's own
    values (which include the target's q_i/sqrt(DEL)); at arbitrary points all
    three equal the oracle's point traversal (shielded: -phi, unregularized)."""
    fx = load_fixture("uniform12k_reg")
    n, t = fx["x"].size, fx["tidx"][:600]
    rng = np.random.default_rng(8)
    extra = rng.uniform(-0.7, 0.7, size=(200, 3))
    px = np.concatenate([fx["x"][t], extra[:, 0]])
    py = np.concatenate([fx["y"][t], extra[:, 1]])
    pz = np.concatenate([fx["z"][t], extra[:, 2]])
    m = px.size
    inp = tmp_path / "in.bin"
    hdr = np.array([n, 1000, 9, 0.5, 0.8, m], dtype=np.float64)
    np.concatenate([hdr, fx["x"], fx["y"], fx["z"], fx["q"], px, py, pz]).tofile(inp)
    out = tmp_path / "p.bin"
    _run("points", str(inp), str(out))
    r = np.fromfile(out)
    pot, E, sh = r[:m], r[m:4 * m].reshape(m, 3), r[4 * m:]
    ref = fx["phi_regcoulomb"][:600] + fx["q"][t] / np.sqrt(0.01)
    assert rel_max(pot[:600], ref) < 1e-10
    o = oracle.tree(fx["x"], fx["y"], fx["z"], 1000)
    mom = o.moments(fx["q"], 9)
    assert rel_max(pot, o.potential_points(fx["q"], mom, 9, 0.5, 0.0, 0.01, 0.01, px, py, pz)) < 1e-10
    assert rel_max(E, o.treeforce_points(fx["q"], mom, 9, 0.5, 0.01, 0.01, px, py, pz)) < 1e-10
    want = -o.potential_points(fx["q"], mom, 9, 0.5, 0.8, 0.0, 0.01, px, py, pz)
    assert rel_max(sh, want) < 1e-10


def test_directsum_mp_recipe(tmp_path, oracle):
    """tests/directsum_mp.cpp's recipe through the drop-in: serial and OpenMP
    force_particle_ds_noreg agree to 1e-15 (the reference test's bar), and the
    values are the unregularized direct-sum field of the oracle."""
    out = tmp_path / "d.bin"
    _run("directsum", str(out))
    E = np.fromfile(out).reshape(-1, 3)
    fx = load_fixture("box16k")
    t = fx["tidx"]
    want = oracle.direct_force(fx["x"], fx["y"], fx["z"], fx["q"], 0.0, t)
    assert rel_max(E[t], want) < 1e-12


def test_treecode_recipe(tmp_path):
    """tests/treecode.cpp's recipe: force_particle_ds_reg from OpenMP threads
    equals the reference's own direct sum (fixture), and treeforce's max-abs
    error against it does not grow over nterms 1..4 (the test's criterion,
    checked by the driver and again here)."""
    out = tmp_path / "t.bin"
    _run("treecode", str(out))
    r = np.fromfile(out)
    fx = load_fixture("box16k")
    n = fx["x"].size
    ds = r[:3 * n].reshape(n, 3) / (CHARGE_CONSTANT * fx["q"][:, None])
    assert rel_max(ds[fx["tidx"]], fx["field_ds"]) < 1e-12
    errs = r[3 * n:].reshape(4, 3).sum(axis=1)
    assert all(b <= a for a, b in zip(errs, errs[1:])), errs


def test_recharge_treemoments(tmp_path, oracle):
    """treemake once, then treemoments + shielded_treepot for two charge sets:
    the second pass equals the oracle with the second charges (treemoments
    re-reads part[].tot_charge, and the cached potentials are dropped)."""
    fx = load_fixture("blobs12k_screened")
    n = fx["x"].size
    q2 = np.random.default_rng(21).uniform(-1.0, 1.0, n)
    inp = tmp_path / "in.bin"
    hdr = np.array([n, 200, 9, 0.7, 0.5], dtype=np.float64)
    np.concatenate([hdr, fx["x"], fx["y"], fx["z"], fx["q"], q2]).tofile(inp)
    out = tmp_path / "p.bin"
    _run("recharge", str(inp), str(out))
    raw = np.fromfile(out).reshape(2, n)
    t = fx["tidx"]
    want1 = -fx["phi_screened"] + fx["q"][t] * INTG
    assert rel_max(raw[0][t], want1) < 1e-10
    o = oracle.tree(fx["x"], fx["y"], fx["z"], 200)
    phi2 = o.potential(q2, o.moments(q2, 9), 9, 0.7, 0.5, 0.0, 0.01, t)
    want2 = -phi2 + q2[t] * INTG
    assert rel_max(raw[1][t], want2) < 1e-10


def test_direct_sum_functions(tmp_path, oracle):
    """force_point_ds_noreg/_reg at points, shielded_pot_particle_ds and
    freespace_particles_ds_noreg/_reg from the drop-in against the oracle's
    direct sums (points: numpy)."""
    P_n, m, alpha = 3000, 64, 0.7
    rng = np.random.default_rng(5)
    x, y, z = rng.uniform(-0.5, 0.5, (3, P_n))
    q = rng.uniform(-1.0, 1.0, P_n)
    pts = rng.uniform(-0.6, 0.6, (3, m))
    inp = tmp_path / "in.bin"
    np.concatenate([np.array([P_n, m, alpha], dtype=np.float64), x, y, z, q,
                    pts.ravel()]).tofile(inp)
    out = tmp_path / "o.bin"
    _run("points_ds", str(inp), str(out))
    r = np.fromfile(out)
    e_noreg = r[:3 * m].reshape(m, 3)
    e_reg = r[3 * m:6 * m].reshape(m, 3)
    pot = r[6 * m:6 * m + P_n]
    f_noreg = r[6 * m + P_n:6 * m + 4 * P_n].reshape(P_n, 3)
    f_reg = r[6 * m + 4 * P_n:].reshape(P_n, 3)
    for d2, got in ((0.0, e_noreg), (0.01, e_reg)):
        d = np.stack([x, y, z], 1)[None, :, :] - pts.T[:, None, :]  # source - point
        rs = (d * d).sum(-1) + d2
        want = (q[None, :, None] * (-d) / (np.sqrt(rs) * rs)[:, :, None]).sum(1)
        assert rel_max(got, want) < 1e-12
    assert rel_max(pot, -oracle.direct(x, y, z, q, alpha, 0.0)) < 1e-12
    scale = CHARGE_CONSTANT * q[:, None]
    assert rel_max(f_noreg, oracle.direct_force(x, y, z, q, 0.0) * scale) < 1e-12
    assert rel_max(f_reg, oracle.direct_force(x, y, z, q, 0.01) * scale) < 1e-12


def test_subtree_calls(tmp_path):
    """Calls with ind != 0 (the traversal from a subtree, the reference's own
    recursion step) from OpenMP threads: shielded_treepot(ind, i) with its self
    term only for the subtree's members, treepotential(ind, -1, x_i) and
    treeforce(ind, i) equal the reference's values; a node outside the tree
    adds nothing."""
    fx = load_fixture("subtree_blobs3k")
    n, m = fx["x"].size, fx["inds"].size
    inp = tmp_path / "in.bin"
    hdr = np.array([n, fx["maxmemb"], fx["nterms"], fx["acc"], fx["kappa"], m], dtype=np.float64)
    np.concatenate([hdr, fx["x"], fx["y"], fx["z"], fx["q"], fx["inds"].astype(np.float64),
                    fx["tidx"].astype(np.float64)]).tofile(inp)
    out = tmp_path / "s.bin"
    _run("subtree", str(inp), str(out))
    r = np.fromfile(out)
    sh, pot, E = r[:m], r[m:2 * m], r[2 * m:].reshape(m, 3)
    want = subtree_expected(fx)
    assert want["inside"].any() and not want["inside"].all()
    for nd in np.unique(fx["inds"]):
        s = fx["inds"] == nd
        assert rel_max(sh[s], fx["sub_shielded_raw"][s]) < 1e-10, nd
        assert rel_max(pot[s], want["pot_pt"][s]) < 1e-10, nd
        assert rel_max(E[s], want["field"][s]) < 1e-10, nd
