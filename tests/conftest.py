"""Test configuration: the `gpu` marker, fixture loading and shared helpers.

CPU tests (-m "not gpu") check the oracle against the reference's fixtures and
the mathematics, the product's host tree build and the C-ABI library's
exports.  GPU tests (-m gpu) call the HIP path through the C-ABI and compare it
with the oracle and the fixtures.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

FIXTURES = ["box16k", "uniform10k", "uniform12k_reg", "blobs12k_screened", "aniso6k", "tiny1",
            "tiny5", "tiny300", "subtree_blobs3k"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def load_fixture(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def fixture_members(fx):
    """Member lists of every node from children + leaf_of (ascending indices)."""
    ch, leaf_of = fx["children"], fx["leaf_of"]
    nt = ch.shape[0]
    by_leaf = {}
    for p, lf in enumerate(leaf_of):
        by_leaf.setdefault(int(lf), []).append(p)
    mem = [None] * nt
    for i in range(nt - 1, -1, -1):
        if ch[i, 0] < 0:
            mem[i] = np.asarray(by_leaf.get(i, []), dtype=np.int64)
        else:
            kids = [mem[c] for c in ch[i] if c >= 0]
            mem[i] = np.sort(np.concatenate(kids))
    return mem


def subtree_expected(fx):
    """The subtree fixture's reference values in this library's conventions,
    per (inds[t], tidx[t]) pair:
      inside   particle tidx[t] is a member of node inds[t]
      phi      +phi of the subtree at the particle (delta 0, MAC shift 0.01):
               shielded_treepot's raw value with its sign and self term undone
               (the self term is added when the target's own leaf is summed
               directly, i.e. when it is a member of the subtree)
      pot_pt   treepotential(ind, -1, x_i): regularized (delta^2 = 0.01), the
               target's own charge included as q_i/sqrt(0.01) when inside
      field    treeforce(ind, i)"""
    mem = fixture_members(fx)
    inds, tidx = fx["inds"], fx["tidx"]
    inside = np.array([np.isin(i, mem[nd]) for nd, i in zip(inds, tidx)])
    q = fx["q"][tidx]
    phi = -(fx["sub_shielded_raw"] - np.where(inside, q * float(fx["intG"]), 0.0))
    return dict(inside=inside, phi=phi, pot_pt=fx["sub_treepotential_point"],
                field=fx["sub_treeforce"])


def rel_l2(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def rel_max(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="session")
def oracle():
    from oracle.pyoracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def ctx():
    import bit_amd
    try:
        c = bit_amd.Context(0)
    except bit_amd.BitError as e:
        if e.code == bit_amd.BIT_E_HIP or e.code == bit_amd.BIT_E_INVALID:
            pytest.skip(f"no HIP device: {e}")
        raise
    yield c
    c.close()


def pytest_runtest_logstart(nodeid, location):
    # name the case before it touches the GPU, so a run that dies names it
    sys.stderr.write(f"[start] {nodeid}\n")
    sys.stderr.flush()
