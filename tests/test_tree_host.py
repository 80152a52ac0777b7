"""The product's host tree build (bit_amd/csrc/tree_build.cpp, C-ABI
bit_tree_build) against the reference fixtures and the oracle: node numbering,
children order, member lists and every TREE double must be bit-identical."""
import numpy as np
import pytest

import bit_amd
from bit_amd import workloads
from conftest import FIXTURES, fixture_members, load_fixture


def _same_tree(t, children, boxes, members):
    assert t.numtree == children.shape[0]
    assert np.array_equal(t.children, children)
    assert np.array_equal(t.boxes.view(np.uint64), boxes.view(np.uint64))
    for i in range(t.numtree):
        assert np.array_equal(t.member_list(i), members[i]), i


@pytest.mark.parametrize("name", FIXTURES)
def test_host_tree_matches_reference(name):
    fx = load_fixture(name)
    t = bit_amd.build_tree(fx["x"], fx["y"], fx["z"], int(fx["maxmemb"]))
    _same_tree(t, fx["children"], fx["boxes"], fixture_members(fx))


@pytest.mark.parametrize("kind,n,seed,mm", [("blobs", 30000, 5, 64), ("aniso", 20000, 6, 25),
                                            ("uniform", 25000, 7, 17), ("aniso", 5000, 8, 3)])
def test_host_tree_matches_oracle(oracle, kind, n, seed, mm):
    P = workloads.make(kind, n, seed)
    o = oracle.tree(P.x, P.y, P.z, mm)
    e = o.export()
    t = bit_amd.build_tree(P.x, P.y, P.z, mm)
    _same_tree(t, e.children, e.boxes, [e.node_members(i) for i in range(e.numtree)])


def test_host_tree_parallel_partition_path(oracle):
    """Big enough that the top levels use the multi-threaded stable partition."""
    P = workloads.gaussian_blobs(300_000, seed=12)
    o = oracle.tree(P.x, P.y, P.z, 2000)
    e = o.export()
    t = bit_amd.build_tree(P.x, P.y, P.z, 2000)
    _same_tree(t, e.children, e.boxes, [e.node_members(i) for i in range(e.numtree)])


def test_leaf_ranges_partition_slots():
    P = workloads.gaussian_blobs(20000, seed=13)
    t = bit_amd.build_tree(P.x, P.y, P.z, 100)
    assert np.array_equal(np.sort(t.perm), np.arange(P.n))
    leaves = [i for i in range(t.numtree) if t.children[i, 0] < 0]
    cover = np.zeros(P.n, dtype=int)
    for lf in leaves:
        cover[t.begin[lf]:t.end[lf]] += 1
        assert t.end[lf] - t.begin[lf] <= 100
    assert (cover == 1).all()
    for i in range(t.numtree):  # children partition their parent's range
        kids = [c for c in t.children[i] if c >= 0]
        if kids:
            assert sum(t.end[c] - t.begin[c] for c in kids) == t.end[i] - t.begin[i]
            assert min(t.begin[c] for c in kids) == t.begin[i]


def test_invalid_inputs():
    x = np.array([0.0, 1.0, np.nan])
    with pytest.raises(bit_amd.BitError) as e:
        bit_amd.build_tree(x, x, x, 1)
    assert e.value.code == bit_amd.BIT_E_INVALID
    with pytest.raises(bit_amd.BitError):
        bit_amd.build_tree(np.zeros(0), np.zeros(0), np.zeros(0), 4)


def test_coincident_particles_rejected():
    """All-coincident points can never be split; the reference grows its array
    without bound, here the build stops with BIT_E_TREE."""
    x = np.full(50, 0.25)
    with pytest.raises(bit_amd.BitError) as e:
        bit_amd.build_tree(x, x, x, 10)
    assert e.value.code == bit_amd.BIT_E_TREE
