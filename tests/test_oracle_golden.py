"""The oracle against the reference's own outputs (fixtures written by
tests/golden/gen_golden.py from the compiled reference)."""
import numpy as np
import pytest

from conftest import FIXTURES, fixture_members, load_fixture, rel_max, subtree_expected

INTG = -0.0000483811229725  # shielded_treepot self term (tree3d.cpp:3575-3596)


@pytest.mark.parametrize("name", FIXTURES)
def test_tree_bit_exact(oracle, name):
    fx = load_fixture(name)
    t = oracle.tree(fx["x"], fx["y"], fx["z"], int(fx["maxmemb"])).export()
    assert t.numtree == fx["children"].shape[0]
    assert np.array_equal(t.children, fx["children"])
    assert np.array_equal(t.nummemb, fx["nummemb"])
    # every TREE double, bit for bit (aspect ratios of flat boxes are NaN: compare bits)
    assert np.array_equal(t.boxes.view(np.uint64), fx["boxes"].view(np.uint64))
    mem = fixture_members(fx)
    for i in range(t.numtree):
        assert np.array_equal(t.node_members(i), mem[i])


def test_quirk_exercised(oracle):
    """aniso6k hits treemake's z-split index quirk (tree3d.cpp:1458-1477)."""
    fx = load_fixture("aniso6k")
    t = oracle.tree(fx["x"], fx["y"], fx["z"], int(fx["maxmemb"]))
    assert t.quirk_hits > 0


@pytest.mark.parametrize("name", ["box16k", "uniform10k"])
def test_moments_match_reference(oracle, name):
    fx = load_fixture(name)
    nterms = int(fx["nterms"])
    t = oracle.tree(fx["x"], fx["y"], fx["z"], int(fx["maxmemb"]))
    m = t.moments(fx["q"], nterms)
    assert np.array_equal(m[fx["mom_nodes"]], fx["moments"])


def test_coulomb_potential(oracle):
    fx = load_fixture("uniform10k")
    t = oracle.tree(fx["x"], fx["y"], fx["z"], 500)
    m = t.moments(fx["q"], 7)
    phi = t.potential(fx["q"], m, 7, 0.5, 0.0, 0.0, 0.01, fx["tidx"])
    assert rel_max(phi, fx["phi_coulomb"]) < 1e-12
    ds = oracle.direct(fx["x"], fx["y"], fx["z"], fx["q"], 0.0, 0.0, fx["tidx"])
    assert rel_max(ds, fx["phi_coulomb_ds"]) < 1e-12


def test_regularized_coulomb_potential(oracle):
    fx = load_fixture("uniform12k_reg")
    t = oracle.tree(fx["x"], fx["y"], fx["z"], 1000)
    m = t.moments(fx["q"], 9)
    phi = t.potential(fx["q"], m, 9, 0.5, 0.0, 0.01, 0.01, fx["tidx"])
    assert rel_max(phi, fx["phi_regcoulomb"]) < 1e-12


@pytest.mark.parametrize("name", ["blobs12k_screened", "aniso6k", "tiny1", "tiny5", "tiny300"])
def test_screened_potential(oracle, name):
    fx = load_fixture(name)
    nt, acc, kap = int(fx["nterms"]), float(fx["acc"]), float(fx["kappa"])
    t = oracle.tree(fx["x"], fx["y"], fx["z"], int(fx["maxmemb"]))
    m = t.moments(fx["q"], nt)
    phi = t.potential(fx["q"], m, nt, acc, kap, 0.0, 0.01, fx["tidx"])
    assert rel_max(phi, fx["phi_screened"]) < 1e-12
    if "phi_screened_ds" in fx:
        ds = oracle.direct(fx["x"], fx["y"], fx["z"], fx["q"], kap, 0.0, fx["tidx"])
        assert rel_max(ds, fx["phi_screened_ds"]) < 1e-12


def test_field_matches_treeforce(oracle):
    fx = load_fixture("box16k")
    t = oracle.tree(fx["x"], fx["y"], fx["z"], 100)
    m = t.moments(fx["q"], 5)
    E = t.treeforce(fx["q"], m, 5, 0.5, 0.01, 0.01, fx["tidx"])
    assert rel_max(E, fx["field_tree"]) < 1e-12
    Ed = oracle.direct_force(fx["x"], fx["y"], fx["z"], fx["q"], 0.01, fx["tidx"])
    assert rel_max(Ed, fx["field_ds"]) < 1e-12


def test_subtree_calls_match_reference(oracle):
    """Traversals that start at a node other than the root (the reference's
    ind != 0) for the potential, the point potential and the field: every
    depth of the tree, leaves included, targets inside and outside."""
    fx = load_fixture("subtree_blobs3k")
    nt, acc, kap = int(fx["nterms"]), float(fx["acc"]), float(fx["kappa"])
    t = oracle.tree(fx["x"], fx["y"], fx["z"], int(fx["maxmemb"]))
    m = t.moments(fx["q"], nt)
    want = subtree_expected(fx)
    inds, tidx = fx["inds"], fx["tidx"]
    assert 0 < want["inside"].sum() < len(inds) and len(np.unique(inds)) > 20
    for nd in np.unique(inds):
        s = inds == nd
        ti = tidx[s]
        phi = t.potential(fx["q"], m, nt, acc, kap, 0.0, 0.01, ti, root=int(nd))
        assert rel_max(phi, want["phi"][s]) < 1e-12, nd
        pp = t.potential_points(fx["q"], m, nt, acc, 0.0, 0.01, 0.01, fx["x"][ti], fx["y"][ti],
                                fx["z"][ti], root=int(nd))
        assert rel_max(pp, want["pot_pt"][s]) < 1e-12, nd
        E = t.treeforce(fx["q"], m, nt, acc, 0.01, 0.01, ti, root=int(nd))
        assert rel_max(E, want["field"][s]) < 1e-12, nd
    with pytest.raises(ValueError):
        t.potential(fx["q"], m, nt, acc, kap, 0.0, 0.01, tidx[:1], root=t.numtree)
