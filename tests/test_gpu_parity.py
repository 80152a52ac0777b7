"""HIP path (through the C-ABI) against the reference fixtures and the oracle.

Bars: trees and interaction lists bit-exact; FP64 potentials and fields within
1e-10 relative (max norm) of the reference; FP32 (Kahan) within the stated
tolerance of the FP64 result.
"""
import numpy as np
import pytest

import bit_amd
from bit_amd import workloads
from conftest import FIXTURES, fixture_members, load_fixture, rel_l2, rel_max, subtree_expected

pytestmark = pytest.mark.gpu

# Reference entry point of each fixture's phi -> (kappa, delta2, mac_sub).
PHI_KEYS = {
    "phi_coulomb": lambda fx: (0.0, 0.0, 0.01),       # shielded_treepot, alpha 0
    "phi_regcoulomb": lambda fx: (0.0, 0.01, 0.01),   # treepotential, p=-1
    "phi_screened": lambda fx: (float(fx["kappa"]), 0.0, 0.01),  # shielded_treepot
}
FP32_TOL = 2e-6  # relative L2 of the FP32 form against the FP64 form


def _system(ctx, fx):
    return bit_amd.System(ctx, fx["x"], fx["y"], fx["z"], fx["q"], int(fx["maxmemb"]))


@pytest.mark.parametrize("src", ["host", "device"])
@pytest.mark.parametrize("name", FIXTURES)
def test_system_tree_bit_exact(ctx, name, src):
    """Host inputs of this size take the host build, device inputs the device
    build: both give the reference's tree."""
    fx = load_fixture(name)
    if src == "host":
        s = _system(ctx, fx)
    else:
        dev = [bit_amd.DeviceArray.from_numpy(ctx, fx[k]) for k in ("x", "y", "z", "q")]
        s = bit_amd.System(ctx, *dev, int(fx["maxmemb"]))
    t = s.export_tree()
    assert np.array_equal(t.children, fx["children"])
    assert np.array_equal(t.boxes.view(np.uint64), fx["boxes"].view(np.uint64))
    mem = fixture_members(fx)
    for i in range(t.numtree):
        assert np.array_equal(t.member_list(i), mem[i])


@pytest.mark.parametrize("name", ["box16k", "uniform10k"])
def test_moments_match_reference(ctx, name):
    fx = load_fixture(name)
    s = _system(ctx, fx)
    nt = int(fx["nterms"])
    s.moments(nt)
    m = s.moments_cube()[fx["mom_nodes"]]
    ref = fx["moments"]
    for a, r in zip(m, ref):
        assert np.max(np.abs(a - r)) <= 1e-12 * max(np.max(np.abs(r)), 1e-300)


def _per_target_lists(lst, n):
    far = [[] for _ in range(n)]
    near = [[] for _ in range(n)]
    for key, out in (("far", far), ("near", near)):
        off, masks, nodes = lst[key + "_off"], lst[key + "_mask"], lst[key + "_node"]
        for b in range(len(off) - 1):
            for it in range(off[b], off[b + 1]):
                m = int(masks[it])
                for lane in range(64):
                    if (m >> lane) & 1:
                        out[b * 64 + lane].append(int(nodes[it]))
    return far, near


@pytest.mark.parametrize("name,acc,d2,sub", [("uniform10k", 0.5, 0.0, 0.01),
                                             ("blobs12k_screened", 0.7, 0.0, 0.01),
                                             ("uniform12k_reg", 0.5, 0.01, 0.01),
                                             ("aniso6k", 0.5, 0.0, 0.01),
                                             ("blobs12k_screened", 0.3, 2.5e-5, 2.5e-5),
                                             ("tiny300", 0.5, 0.0, 0.01)])
def test_lists_bit_exact(ctx, oracle, name, acc, d2, sub):
    """Every target's far clusters (as a set: a batch's far items are ordered
    for the far kernel's pairing) and near leaves (in traversal order) equal
    the oracle's per-target recursion (whose tree equals the reference's)."""
    fx = load_fixture(name)
    s = _system(ctx, fx)
    lst = s.lists(4, acc, d2, sub)
    far, near = _per_target_lists(lst, s.n)
    perm = s.export_tree().perm
    o = oracle.tree(fx["x"], fx["y"], fx["z"], int(fx["maxmemb"]))
    for slot in range(s.n):
        f, nr = o.lists(int(perm[slot]), 4, acc, d2, sub)
        assert sorted(far[slot]) == sorted(f), slot
        assert near[slot] == list(nr), slot


def _tree_depth(children):
    depth = np.zeros(len(children), dtype=np.int64)
    for i in range(len(children)):
        for c in children[i]:
            if c >= 0:
                depth[c] = depth[i] + 1
    return int(depth.max())


def test_lists_deep_tree_bit_exact(ctx, oracle):
    """About 100 levels (points at geometric distances along a line, maxmemb 1):
    the traversal's stack, sized 8 entries per level, holds every path and the
    lists equal the oracle's."""
    rng = np.random.default_rng(77)
    k = np.arange(1, 101)
    x = np.concatenate([0.5 * 2.0 ** (-k), rng.uniform(-0.5, 0.5, 60)])
    y = np.concatenate([np.zeros(100), rng.uniform(-0.5, 0.5, 60)])
    z = np.concatenate([np.zeros(100), rng.uniform(-0.5, 0.5, 60)])
    q = rng.uniform(-1, 1, x.size)
    s = bit_amd.System(ctx, x, y, z, q, 1)
    t = s.export_tree()
    assert _tree_depth(t.children) > 75
    lst = s.lists(4, 0.5, 0.0, 0.0)
    far, near = _per_target_lists(lst, s.n)
    o = oracle.tree(x, y, z, 1)
    for slot in range(s.n):
        f, nr = o.lists(int(t.perm[slot]), 4, 0.5, 0.0, 0.0)
        assert sorted(far[slot]) == sorted(f), slot
        assert near[slot] == list(nr), slot


def test_lists_single_traversal_equals_two(ctx):
    """A rebuild stages items into per-batch regions sized by the context's
    previous build (one traversal) and falls back to count-then-fill when a
    batch outgrows them; either way the lists equal a first, two-traversal
    build on a fresh context.  The pair statistics (counted by the count
    traversal, or from the items by the compaction or after a fallback fill)
    agree too."""
    fx = load_fixture("blobs12k_screened")
    own = bit_amd.Context(0)  # no hint from earlier builds
    keys = ("far_pairs", "near_pairs", "lane_far", "lane_near_pairs")
    try:
        s = _system(own, fx)
        seen = []
        for acc in (0.5, 0.7, 0.3, 0.3001, 0.9, 0.5):
            got = s.lists(4, acc, 2.5e-5, 2.5e-5)
            st = s.stats()
            passes = st["list_passes"]
            seen.append(passes)
            fresh = bit_amd.Context(0)
            try:
                r = _system(fresh, fx)
                ref = r.lists(4, acc, 2.5e-5, 2.5e-5)
                st_ref = r.stats()
                assert st_ref["list_passes"] == 2
            finally:
                fresh.close()
            for k in ref:
                assert np.array_equal(got[k], ref[k]), (acc, k, passes)
            for k in keys:
                assert st[k] == st_ref[k], (acc, k, passes)
    finally:
        own.close()
    assert seen[0] == 2 and set(seen) == {1, 2}, seen


@pytest.mark.parametrize("name", ["uniform10k", "uniform12k_reg", "blobs12k_screened", "aniso6k",
                                  "tiny1", "tiny5", "tiny300"])
def test_potential_fp64_matches_reference(ctx, name):
    fx = load_fixture(name)
    s = _system(ctx, fx)
    key = next(k for k in PHI_KEYS if k in fx)
    kappa, d2, sub = PHI_KEYS[key](fx)
    phi = s.potential(int(fx["nterms"]), float(fx["acc"]), kappa, d2, sub, "fp64")
    ref = fx[key]
    got = phi[fx["tidx"]]
    if fx["x"].size == 1:
        assert got[0] == 0.0 and ref[0] == 0.0
        return
    assert rel_max(got, ref) < 1e-10
    assert rel_l2(got, ref) < 1e-11


@pytest.mark.parametrize("name", ["uniform10k", "uniform12k_reg", "blobs12k_screened", "aniso6k",
                                  "tiny300"])
def test_potential_fp32_within_tolerance(ctx, name):
    fx = load_fixture(name)
    s = _system(ctx, fx)
    key = next(k for k in PHI_KEYS if k in fx)
    kappa, d2, sub = PHI_KEYS[key](fx)
    args = (int(fx["nterms"]), float(fx["acc"]), kappa, d2, sub)
    p64 = s.potential(*args, precision="fp64")
    p32 = s.potential(*args, precision="fp32")
    assert rel_l2(p32, p64) < FP32_TOL


@pytest.mark.parametrize("kappa", [1e-7, 0.5, 40.0])
def test_potential_fp32_near_scaled_and_unscaled(ctx, kappa):
    """The screened FP32 near field runs in coordinates scaled by kappa*log2(e)
    when that factor lies in [2^-20, 2^20] (0.5, 40) and unscaled below it
    (1e-7); all three stay within the FP32 tolerance of the FP64 result."""
    fx = load_fixture("blobs12k_screened")
    s = _system(ctx, fx)
    args = (int(fx["nterms"]), float(fx["acc"]), kappa, 0.0, 0.01)
    p64 = s.potential(*args, precision="fp64")
    p32 = s.potential(*args, precision="fp32")
    assert np.all(np.isfinite(p32))
    assert rel_l2(p32, p64) < FP32_TOL


def test_points_fp32_on_particles_scaled_domain(ctx):
    """Points exactly on particles in a domain 20 times the blobs' extent,
    screened (kappa 0.5, so the FP32 near field runs in scaled coordinates)
    and unregularized (delta2 0): targets and sources round their offsets
    from the leaf centre by the same operations, so a coincident pair has a
    zero offset and is skipped as in FP64 (a missed pair would add about
    q / 1e-7).  FP32 stays within the FP32 tolerance of FP64, and every point
    within 1e-5 of the largest |phi| (cancelling charges leave some points
    with small potentials)."""
    P = workloads.gaussian_blobs(30000, seed=43)
    sc = 20.0
    x, y, z = P.x * sc, P.y * sc, P.z * sc
    s = bit_amd.System(ctx, x, y, z, P.q, 300)
    rng = np.random.default_rng(8)
    k = rng.choice(P.n, 600, replace=False)
    px, py, pz = x[k].copy(), y[k].copy(), z[k].copy()
    p64 = s.potential_at(px, py, pz, 9, 0.5, 0.5, 0.0, 0.01, "fp64")
    p32 = s.potential_at(px, py, pz, 9, 0.5, 0.5, 0.0, 0.01, "fp32")
    assert np.all(np.isfinite(p32)) and np.all(np.isfinite(p64))
    assert rel_l2(p32, p64) < FP32_TOL
    assert np.max(np.abs(p32 - p64)) < 1e-5 * np.max(np.abs(p64))
    s.close()


def test_field_fp64_matches_treeforce(ctx):
    fx = load_fixture("box16k")
    s = _system(ctx, fx)
    E = s.field(5, 0.5, 0.01, 0.01)[fx["tidx"]]
    assert rel_max(E, fx["field_tree"]) < 1e-10


def test_gpu_direct_sum_matches_oracle(ctx, oracle):
    fx = load_fixture("blobs12k_screened")
    s = _system(ctx, fx)
    got = s.direct(0.5, 0.0, fx["tidx"])
    assert rel_max(got, fx["phi_screened_ds"]) < 1e-12
    o = oracle.direct(fx["x"], fx["y"], fx["z"], fx["q"], 1.3, 1e-4, fx["tidx"])
    assert rel_max(s.direct(1.3, 1e-4, fx["tidx"]), o) < 1e-12


@pytest.mark.parametrize("kappa,d2,nterms,acc", [(0.5, 2.5e-5, 9, 0.7), (0.0, 1e-4, 12, 0.4),
                                                 (2.0, 4e-4, 5, 0.5), (1.0, 0.0, 13, 0.3)])
def test_clean_mac_matches_oracle(ctx, oracle, kappa, d2, nterms, acc):
    """The default MAC (mac_sub = delta2, the reference's treepotential rule)
    with regularized screened kernels the reference cannot be compiled for."""
    P = workloads.gaussian_blobs(20000, seed=21)
    s = bit_amd.System(ctx, P.x, P.y, P.z, P.q, 150)
    tidx = workloads.sample_targets(P.n, 2000, seed=3)
    o = oracle.tree(P.x, P.y, P.z, 150)
    m = o.moments(P.q, nterms)
    ref = o.potential(P.q, m, nterms, acc, kappa, d2, d2, tidx)
    got = s.potential(nterms, acc, kappa, d2, None, "fp64")[tidx]
    assert rel_max(got, ref) < 1e-10
    got32 = s.potential(nterms, acc, kappa, d2, None, "fp32")[tidx]
    assert rel_l2(got32, ref) < FP32_TOL


def test_device_resident_inputs(ctx):
    P = workloads.gaussian_blobs(50000, seed=22)
    dev = [bit_amd.DeviceArray.from_numpy(ctx, a) for a in (P.x, P.y, P.z, P.q)]
    out = bit_amd.DeviceArray(ctx, P.n)
    s1 = bit_amd.System(ctx, *dev, 300)
    s2 = bit_amd.System(ctx, P.x, P.y, P.z, P.q, 300)
    assert np.array_equal(s1.export_tree().perm, s2.export_tree().perm)
    s1.potential(8, 0.5, 0.5, 1e-4, None, "fp32", out=out)
    a = out.numpy()
    b = s2.potential(8, 0.5, 0.5, 1e-4, None, "fp32")
    assert np.array_equal(a, b)
    for d in dev + [out]:
        d.free()


TORCH_INTEROP = """
import numpy as np, torch
import bit_amd
from bit_amd import workloads
P = workloads.uniform_cube(20000, seed=26)
dev = [torch.tensor(a, dtype=torch.float64, device="cuda") for a in (P.x, P.y, P.z, P.q)]
out = torch.empty(P.n, dtype=torch.float64, device="cuda")
ctx = bit_amd.Context(0)
ctx.set_stream(torch.cuda.current_stream().cuda_stream)
s = bit_amd.System(ctx, *dev, 200)
s.potential(7, 0.5, 0.3, 0.0, None, "fp64", out=out)
torch.cuda.synchronize()
a = out.cpu().numpy()
b = bit_amd.System(ctx, P.x, P.y, P.z, P.q, 200).potential(7, 0.5, 0.3, 0.0, None, "fp64")
assert np.array_equal(a, b), float(np.abs(a - b).max())
print("torch interop ok")
"""


def test_torch_interop_subprocess():
    """torch tensors in and out, torch's stream; torch imported first so both
    share one HIP runtime (own process: this session already loaded one)."""
    import subprocess
    import sys
    pytest.importorskip("torch")
    r = subprocess.run([sys.executable, "-c", TORCH_INTEROP], capture_output=True, text=True,
                       timeout=300, cwd=__import__("conftest").ROOT)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "torch interop ok" in r.stdout


def test_repeat_is_deterministic(ctx):
    P = workloads.gaussian_blobs(40000, seed=23)
    s = bit_amd.System(ctx, P.x, P.y, P.z, P.q, 200)
    a = s.potential(9, 0.6, 0.5, 1e-4, None, "fp32")
    b = s.potential(9, 0.6, 0.5, 1e-4, None, "fp32")
    assert np.array_equal(a, b)


def test_stats_count_interactions(ctx, oracle):
    fx = load_fixture("blobs12k_screened")
    s = _system(ctx, fx)
    s.potential(9, 0.7, 0.5, 0.0, 0.01, "fp64")
    st = s.stats()
    o = oracle.tree(fx["x"], fx["y"], fx["z"], int(fx["maxmemb"]))
    m = o.moments(fx["q"], 2)
    _, nn, nf = o.potential(fx["q"], m, 2, 0.7, 0.5, 0.0, 0.01, None, counts=True)
    # the oracle's near count excludes the self pair
    assert st["near_pairs"] == int(nn.sum()) + s.n
    assert st["far_pairs"] == int(nf.sum())


def test_oom_and_budget_errors_leave_context_usable(ctx):
    P = workloads.uniform_cube(5000, seed=24)
    small = bit_amd.Context(0, hbm_budget=1 << 16)
    with pytest.raises(bit_amd.BitError) as e:
        bit_amd.System(small, P.x, P.y, P.z, P.q, 100)
    assert e.value.code == bit_amd.BIT_E_OOM
    small.close()
    s = bit_amd.System(ctx, P.x, P.y, P.z, P.q, 100)
    ctx.inject_alloc_failures(1)
    with pytest.raises(bit_amd.BitError) as e:
        s.potential(6, 0.5, 0.0, 0.0, None, "fp64")
    assert e.value.code == bit_amd.BIT_E_OOM
    ok = s.potential(6, 0.5, 0.0, 0.0, None, "fp64")
    assert np.isfinite(ok).all()


def test_subtree_root_matches_reference(ctx, oracle):
    """Evaluations that start at node `root` (the reference's ind != 0 calls),
    for every node of the subtree fixture (each depth, leaves included): FP64
    potential, point potential and field within 1e-10 of the reference's
    values, FP32 within FP32_TOL of FP64, the lists of every third node equal
    to the oracle's traversal from the same node, a mixed launch equal to the
    single evaluation, and a root outside the tree rejected."""
    fx = load_fixture("subtree_blobs3k")
    nt, acc, kap = int(fx["nterms"]), float(fx["acc"]), float(fx["kappa"])
    s = _system(ctx, fx)
    want = subtree_expected(fx)
    inds, tidx = fx["inds"], fx["tidx"]
    perm = s.export_tree().perm
    o = oracle.tree(fx["x"], fx["y"], fx["z"], int(fx["maxmemb"]))
    nodes = np.unique(inds)
    for k, nd in enumerate(nodes):
        nd = int(nd)
        sel = inds == nd
        ti = tidx[sel]
        p64 = s.potential(nt, acc, kap, 0.0, 0.01, "fp64", root=nd)
        assert rel_max(p64[ti], want["phi"][sel]) < 1e-10, nd
        p32 = s.potential(nt, acc, kap, 0.0, 0.01, "fp32", root=nd)
        assert rel_l2(p32, p64) < FP32_TOL, nd
        pp = s.potential_at(fx["x"][ti], fx["y"][ti], fx["z"][ti], nt, acc, 0.0, 0.01, 0.01,
                            "fp64", root=nd)
        assert rel_max(pp, want["pot_pt"][sel]) < 1e-10, nd
        E = s.field(nt, acc, 0.01, 0.01, 0, "fp64", root=nd)
        assert rel_max(E[ti], want["field"][sel]) < 1e-10, nd
        E32 = s.field(nt, acc, 0.01, 0.01, 0, "fp32", root=nd)
        assert rel_l2(E32, E) < FP32_TOL, nd
        if k % 3 == 0:
            far, near = _per_target_lists(s.lists(nt, acc, 0.0, 0.01, root=nd), s.n)
            for slot in range(s.n):
                f, nr = o.lists(int(perm[slot]), nt, acc, 0.0, 0.01, root=nd)
                assert sorted(far[slot]) == sorted(f) and near[slot] == list(nr), (nd, slot)
        if k % 5 == 0:
            prm = bit_amd.make_params(nt, acc, kap, 0.0, 0.01, "fp32", root=nd)
            outs, _ = bit_amd.eval_multi(ctx, [s], [prm])
            assert np.array_equal(outs[0], p32), nd
    for bad in (-1, s.numtree):
        with pytest.raises(bit_amd.BitError) as e:
            s.potential(nt, acc, kap, 0.0, 0.01, "fp64", root=bad)
        assert e.value.code == bit_amd.BIT_E_INVALID
    s.close()


def test_invalid_params_rejected(ctx):
    P = workloads.uniform_cube(1000, seed=25)
    s = bit_amd.System(ctx, P.x, P.y, P.z, P.q, 100)
    for bad in (dict(nterms=0), dict(nterms=14), dict(kappa=float("nan")), dict(delta2=-1.0)):
        kw = dict(nterms=6, acc=0.5, kappa=0.0, delta2=0.0)
        kw.update(bad)
        with pytest.raises(bit_amd.BitError) as e:
            s.potential(**kw)
        assert e.value.code == bit_amd.BIT_E_INVALID


@pytest.mark.parametrize("kind,n,seed,mm", [("blobs", 300_000, 31, 2000), ("aniso", 60_000, 32, 25),
                                            ("uniform", 100_000, 33, 7)])
def test_device_tree_equals_host_and_oracle(ctx, oracle, kind, n, seed, mm):
    """The device tree build (inputs in HBM: split rounds across a generation,
    then per-node programs of 1024 or 256 threads) gives the host build's tree
    and slot order, and the oracle's (= reference's) tree."""
    P = workloads.make(kind, n, seed)
    dev = [bit_amd.DeviceArray.from_numpy(ctx, a) for a in (P.x, P.y, P.z, P.q)]
    s = bit_amd.System(ctx, *dev, mm)
    d = s.export_tree()
    s.close()
    for a in dev:
        a.free()
    h = bit_amd.build_tree(P.x, P.y, P.z, mm)
    assert d.numtree == h.numtree
    assert np.array_equal(d.children, h.children)
    assert np.array_equal(d.boxes.view(np.uint64), h.boxes.view(np.uint64))
    assert np.array_equal(d.begin, h.begin) and np.array_equal(d.end, h.end)
    assert np.array_equal(d.perm, h.perm)
    e = oracle.tree(P.x, P.y, P.z, mm).export()
    assert np.array_equal(d.children, e.children)
    assert np.array_equal(d.boxes.view(np.uint64), e.boxes.view(np.uint64))
    for i in range(0, d.numtree, max(1, d.numtree // 50)):
        assert np.array_equal(d.member_list(i), e.node_members(i))


def test_device_build_rejects_inseparable_particles(ctx):
    """Host inputs above 8,192 particles are built on the device.  Two points
    holding 10,000 coincident particles each can never be split into leaves of
    100: the build stops with BIT_E_TREE at its generation cap (the reference
    grows its tree without bound), and the context stays usable."""
    n = 20_000
    x = np.where(np.arange(n) % 2 == 0, 0.25, -0.5)
    y = np.full(n, 0.125)
    z = np.where(np.arange(n) % 2 == 0, -0.375, 0.75)
    q = np.ones(n)
    with pytest.raises(bit_amd.BitError) as e:
        bit_amd.System(ctx, x, y, z, q, 100)
    assert e.value.code == bit_amd.BIT_E_TREE
    P = workloads.uniform_cube(n, seed=5)
    s = bit_amd.System(ctx, P.x, P.y, P.z, P.q, 100)
    phi = s.potential(5, 0.5, 0.0, 0.0, None, "fp64")
    assert s.numtree > 1 and np.all(np.isfinite(phi))
    s.close()


def test_device_inputs_rejected_when_nonfinite(ctx):
    x = np.random.default_rng(1).uniform(size=1000)
    x[17] = np.nan
    dev = [bit_amd.DeviceArray.from_numpy(ctx, a) for a in (x, x, x, x)]
    with pytest.raises(bit_amd.BitError) as e:
        bit_amd.System(ctx, *dev, 10)
    assert e.value.code == bit_amd.BIT_E_INVALID
    for d in dev:
        d.free()


def test_points_at_particles_match_reference_p_minus_1(ctx):
    """treepotential(0, -1, x_i, ...) at the particle positions includes the
    target's own regularized term q_i/sqrt(DEL); the fixture stores it removed."""
    fx = load_fixture("uniform12k_reg")
    s = _system(ctx, fx)
    t = fx["tidx"]
    got = s.potential_at(fx["x"][t], fx["y"][t], fx["z"][t], int(fx["nterms"]),
                         float(fx["acc"]), 0.0, 0.01, 0.01, "fp64")
    ref = fx["phi_regcoulomb"] + fx["q"][t] / np.sqrt(0.01)
    assert rel_max(got, ref) < 1e-10


def test_field_at_particles_matches_treeforce(ctx):
    """treeforce(0, -1) at a particle: its own term has zero gradient, so the
    point form equals the per-particle form."""
    fx = load_fixture("box16k")
    s = _system(ctx, fx)
    t = fx["tidx"]
    E = s.field_at(fx["x"][t], fx["y"][t], fx["z"][t], 5, 0.5, 0.01, 0.01)
    assert rel_max(E, fx["field_tree"]) < 1e-10


def _points(P, m, seed):
    """Points inside and around the particles' bounding box, plus a few
    exactly on particles (coincident pairs) and far outside."""
    rng = np.random.default_rng(seed)
    lo = np.array([P.x.min(), P.y.min(), P.z.min()])
    hi = np.array([P.x.max(), P.y.max(), P.z.max()])
    pts = lo + (hi - lo) * rng.uniform(-0.2, 1.2, size=(m, 3))
    k = rng.choice(P.n, 37, replace=False)
    pts[:37] = np.stack([P.x[k], P.y[k], P.z[k]], axis=1)
    pts[37:40] = hi + 5.0 * (hi - lo)
    return pts[:, 0].copy(), pts[:, 1].copy(), pts[:, 2].copy()


@pytest.mark.parametrize("kappa,d2,sub,nterms,acc", [(0.5, 0.0, 0.01, 9, 0.5),
                                                     (0.5, 2.5e-5, 2.5e-5, 10, 0.3),
                                                     (0.0, 0.01, 0.01, 7, 0.6)])
def test_points_match_oracle(ctx, oracle, kappa, d2, sub, nterms, acc):
    P = workloads.gaussian_blobs(30000, seed=41)
    s = bit_amd.System(ctx, P.x, P.y, P.z, P.q, 300)
    px, py, pz = _points(P, 3001, seed=5)
    o = oracle.tree(P.x, P.y, P.z, 300)
    m = o.moments(P.q, nterms)
    ref = o.potential_points(P.q, m, nterms, acc, kappa, d2, sub, px, py, pz)
    got = s.potential_at(px, py, pz, nterms, acc, kappa, d2, sub, "fp64")
    assert rel_max(got, ref) < 1e-10
    got32 = s.potential_at(px, py, pz, nterms, acc, kappa, d2, sub, "fp32")
    assert rel_l2(got32, ref) < FP32_TOL
    # the particle-target evaluation is untouched by the point evaluation
    p64 = s.potential(nterms, acc, kappa, d2, sub, "fp64")
    tidx = workloads.sample_targets(P.n, 500, seed=6)
    assert rel_max(p64[tidx], o.potential(P.q, m, nterms, acc, kappa, d2, sub, tidx)) < 1e-10


def test_field_points_match_oracle(ctx, oracle):
    P = workloads.gaussian_blobs(30000, seed=42)
    s = bit_amd.System(ctx, P.x, P.y, P.z, P.q, 300)
    px, py, pz = _points(P, 2000, seed=7)
    o = oracle.tree(P.x, P.y, P.z, 300)
    m = o.moments(P.q, 6)
    ref = o.treeforce_points(P.q, m, 6, 0.5, 0.01, 0.01, px, py, pz)
    got = s.field_at(px, py, pz, 6, 0.5, 0.01, 0.01)
    assert rel_max(got, ref) < 1e-10


def test_points_empty_and_invalid(ctx):
    P = workloads.uniform_cube(2000, seed=3)
    s = bit_amd.System(ctx, P.x, P.y, P.z, P.q, 100)
    e = np.zeros(0)
    assert s.potential_at(e, e, e, 5, 0.5).shape == (0,)
    assert s.field_at(e, e, e, 5, 0.5).shape == (0, 3)
    with pytest.raises(bit_amd.BitError):
        s.potential_at(np.array([np.nan]), np.zeros(1), np.zeros(1), 5, 0.5)
    assert np.isfinite(s.potential(5, 0.5)).all()


@pytest.mark.parametrize("L,nterms", [(1, 6), (2, 5), (3, 8)])
def test_field_family_matches_oracle(ctx, oracle, L, nterms):
    """KERNEL_ORDER = L regularized field family (tree3d.cpp:73-97, 614-791),
    particle targets and arbitrary points, FP64 within 1e-10 of the oracle."""
    P = workloads.gaussian_blobs(20000, seed=43)
    s = bit_amd.System(ctx, P.x, P.y, P.z, P.q, 200)
    o = oracle.tree(P.x, P.y, P.z, 200)
    m = o.moments(P.q, nterms)
    tidx = workloads.sample_targets(P.n, 1500, seed=9)
    ref = o.treeforce(P.q, m, nterms, 0.5, 0.01, 0.01, tidx, kernel_order=L)
    got = s.field(nterms, 0.5, 0.01, 0.01, kernel_order=L)[tidx]
    assert rel_max(got, ref) < 1e-10
    E0 = s.field(nterms, 0.5, 0.01, 0.01)[tidx]
    assert rel_l2(got, E0) > 1e-6  # the higher terms are present
    px, py, pz = _points(P, 700, seed=10)
    refp = o.treeforce_points(P.q, m, nterms, 0.5, 0.01, 0.01, px, py, pz, kernel_order=L)
    assert rel_max(s.field_at(px, py, pz, nterms, 0.5, 0.01, 0.01, kernel_order=L), refp) < 1e-10


def test_kernel_order_rejected_for_potential(ctx):
    """The family is a field option (KERNEL_ORDER is read by the field paths
    only); potentials and out-of-range orders are refused, the system stays usable."""
    import ctypes as C
    P = workloads.uniform_cube(3000, seed=4)
    s = bit_amd.System(ctx, P.x, P.y, P.z, P.q, 100)
    out = np.zeros(P.n)
    prm = bit_amd.make_params(5, 0.5, 0.0, 0.01, None, "fp64", kernel_order=1)
    rc = bit_amd.lib.bit_eval_potential(s.h, C.byref(prm), out.ctypes.data_as(C.c_void_p), 0)
    assert rc != 0 and b"field only" in bit_amd.lib.bit_last_error()
    with pytest.raises(bit_amd.BitError):
        s.field(5, 0.5, 0.01, kernel_order=99)
    assert np.isfinite(s.field(5, 0.5, 0.01, kernel_order=2)).all()


def test_recycled_device_blocks(ctx):
    """Buffers released by a system are reused by the next one (stale contents
    included): results stay bit-identical across rebuilds of different sizes,
    and the live byte count returns to where it was."""
    P = workloads.gaussian_blobs(30000, seed=44)
    Q = workloads.uniform_cube(7000, seed=45)
    base = ctx.bytes_in_use
    ref = None
    for rep in range(3):
        for X in (P, Q):
            s = bit_amd.System(ctx, X.x, X.y, X.z, X.q, 150)
            phi = s.potential(8, 0.4, 0.5, 1e-4, None, "fp32")
            s.close()
            if X is P:
                if ref is None:
                    ref = phi
                assert np.array_equal(phi, ref), rep
    assert ctx.bytes_in_use == base


def _popcount64(m):
    m = m.astype(np.uint64)
    m = m - ((m >> np.uint64(1)) & np.uint64(0x5555555555555555))
    m = (m & np.uint64(0x3333333333333333)) + ((m >> np.uint64(2)) & np.uint64(0x3333333333333333))
    m = (m + (m >> np.uint64(4))) & np.uint64(0x0F0F0F0F0F0F0F0F)
    return ((m * np.uint64(0x0101010101010101)) >> np.uint64(56)).astype(np.int64)


def test_lane_stats_count_the_kernels_slots(ctx):
    """lane_far and lane_near_pairs count what the FP32 kernels run: k_farq's
    sweeps (an item shares one with the next when the two have at most 64
    targets together) and k_near32's pair slots (the leaf's length per lane
    for a target-parallel item; 4 per 64 sources per four targets for items
    of at most 40 targets), restated here from the lists."""
    P = workloads.gaussian_blobs(60000, seed=13)
    s = bit_amd.System(ctx, P.x, P.y, P.z, P.q, 200)
    s.potential(9, 0.4, 0.5, 1e-5, None, "fp32")
    st = s.stats()
    lst = s.lists(9, 0.4, 1e-5, None)
    t = s.export_tree()
    size = (t.end - t.begin).astype(np.int64)
    fo, fpc = lst["far_off"], _popcount64(lst["far_mask"])
    sweeps = sum(_kernel_sweeps(fpc[fo[b]:fo[b + 1]]) for b in range(len(fo) - 1))
    npc, nlen = _popcount64(lst["near_mask"]), size[lst["near_node"]]
    sparse = npc <= 40
    slots = int(np.sum(nlen[~sparse]))
    slots += int(np.sum((npc[sparse] + 3) // 4 * ((nlen[sparse] + 63) // 64) * 4))
    assert st["lane_far"] == 64 * sweeps
    assert st["lane_near_pairs"] == 64 * slots
    assert st["near_pairs"] == int(np.sum(npc * nlen))
    assert sparse.any() and (~sparse).any() and sweeps < len(fpc)
    s.close()


# Items per window of the far pairing order (kFarOrderWindow in kernels.hip;
# tests/test_far_sweep_asm.py checks the two agree).
FAR_ORDER_WINDOW = 128


def _kernel_sweeps(pcs):
    """k_farq's units over a batch's far items of target counts pcs: an item
    shares a sweep with the next when the two have at most 64 targets
    together, never across a window boundary."""
    sweeps, pend = 0, -1
    for k, pc in enumerate(pcs):
        if pend >= 0 and pend + pc <= 64 and k % FAR_ORDER_WINDOW != 0:
            pend = -1
        else:
            sweeps += 1
            pend = pc
    return sweeps


def _pairing_order(items):
    """A batch's far items (traversal order) in the far kernel's pairing order,
    window by window: buckets by target count, the largest remaining bucket j
    against the smallest i: i == j -> all of j; i + j <= 64 -> min(|i|, |j|)
    pairs (j, i); else all of j alone."""
    if len(items) > FAR_ORDER_WINDOW:
        return (_pairing_order(items[:FAR_ORDER_WINDOW]) +
                _pairing_order(items[FAR_ORDER_WINDOW:]))
    buckets = {}
    for it in items:
        buckets.setdefault(bin(int(it[0])).count("1"), []).append(it)
    out = []
    while any(buckets.values()):
        full = sorted(c for c, v in buckets.items() if v)
        i, j = full[0], full[-1]
        if i == j:
            out += buckets[j]
            buckets[j] = []
        elif i + j <= 64:
            k = min(len(buckets[i]), len(buckets[j]))
            for a, b in zip(buckets[j][:k], buckets[i][:k]):
                out += [a, b]
            buckets[i], buckets[j] = buckets[i][k:], buckets[j][k:]
        else:
            out += buckets[j]
            buckets[j] = []
    return out


def _min_sweeps(pcs):
    """Fewest sweeps of at most two items of at most 64 targets together
    (greedy matching of sorted counts, largest with smallest)."""
    a = sorted(pcs)
    i, j, n = 0, len(a) - 1, 0
    while i <= j:
        if i < j and a[i] + a[j] <= 64:
            i += 1
        j -= 1
        n += 1
    return n


def test_far_items_in_pairing_order(ctx):
    """Each batch's far items are its traversal-order items (depth-first
    preorder of their clusters) rearranged as the pairing order states, the
    far kernel's rule (an item shares a sweep with the next when the two have
    at most 64 targets together, within a window) then forms the fewest sweeps
    possible in each window, and lane_far counts them (200k blobs, acc 0.25:
    most batches hold more than one window)."""
    P = workloads.gaussian_blobs(200000, seed=13)
    s = bit_amd.System(ctx, P.x, P.y, P.z, P.q, 100)
    s.potential(5, 0.25, 0.5, 1e-5, None, "fp32")
    st = s.stats()
    lst = s.lists(5, 0.25, 1e-5, None)
    ch = s.export_tree().children
    pre = np.zeros(len(ch), dtype=np.int64)
    stack, k = [0], 0
    while stack:
        v = stack.pop()
        pre[v] = k
        k += 1
        stack += [c for c in ch[v][::-1] if c >= 0]
    fo, fm, fn = lst["far_off"], lst["far_mask"], lst["far_node"]
    sweeps, total, windows = 0, 0, 0
    for b in range(len(fo) - 1):
        got = list(zip(fm[fo[b]:fo[b + 1]].tolist(), fn[fo[b]:fo[b + 1]].tolist()))
        trav = sorted(got, key=lambda it: pre[it[1]])
        assert got == _pairing_order(trav), b
        pcs = [bin(m).count("1") for m, _ in got]
        sweeps += _kernel_sweeps(pcs)
        total += sum(_min_sweeps(pcs[w:w + FAR_ORDER_WINDOW])
                     for w in range(0, len(pcs), FAR_ORDER_WINDOW))
        windows = max(windows, (len(pcs) + FAR_ORDER_WINDOW - 1) // FAR_ORDER_WINDOW)
    assert sweeps == total and st["lane_far"] == 64 * sweeps
    assert windows >= 2  # batches of more than one window occur
    s.close()


def test_full_size_blobs4m(ctx):
    """The 4M-particle, 16-blob workload at its tuned point (kappa 0.5, delta
    0.005, p 11, acc 0.255, N0 500): the FP32 evaluation stays within 1e-6
    relative L2 of the FP64 direct sum on 2,000 seeded targets, and the lists
    partition the sources for every target in total: the far clusters' and
    near leaves' sizes, times their targets, sum to N^2 (each target's
    traversal accepts or visits every particle exactly once)."""
    P = workloads.gaussian_blobs(4_000_000, seed=3)
    dev = [bit_amd.DeviceArray.from_numpy(ctx, a) for a in (P.x, P.y, P.z, P.q)]
    s = bit_amd.System(ctx, *dev, 500)
    d2 = 0.005 ** 2
    phi = s.potential(12, 0.255, 0.5, d2, None, "fp32")
    tidx = workloads.sample_targets(P.n, 2000, seed=9)
    ref = s.direct(0.5, d2, tidx)
    assert rel_l2(phi[tidx], ref) <= 1e-6
    lst = s.lists(12, 0.255, d2, None)
    t = s.export_tree()
    size = (t.end - t.begin).astype(np.int64)
    total = int(np.sum(_popcount64(lst["far_mask"]) * size[lst["far_node"]]))
    total += int(np.sum(_popcount64(lst["near_mask"]) * size[lst["near_node"]]))
    assert total == P.n * P.n
    s.close()
    for a in dev:
        a.free()


@pytest.mark.parametrize("L,nterms", [(0, 5), (0, 12), (2, 9), (1, 12)])
def test_field_fp32_and_high_order(ctx, oracle, L, nterms):
    """The field in the radial-derivative form: FP64 within 1e-10 of the oracle's
    treeforce up to nterms 12 (with and without the family), FP32 within
    FP32_TOL (relative L2) of FP64, at particles and at points."""
    P = workloads.gaussian_blobs(30000, seed=47)
    s = bit_amd.System(ctx, P.x, P.y, P.z, P.q, 200)
    o = oracle.tree(P.x, P.y, P.z, 200)
    m = o.moments(P.q, nterms)
    tidx = workloads.sample_targets(P.n, 1500, seed=19)
    ref = o.treeforce(P.q, m, nterms, 0.5, 0.01, 0.01, tidx, kernel_order=L)
    E64 = s.field(nterms, 0.5, 0.01, 0.01, kernel_order=L)
    assert rel_max(E64[tidx], ref) < 1e-10
    E32 = s.field(nterms, 0.5, 0.01, 0.01, kernel_order=L, precision="fp32")
    assert rel_l2(E32, E64) < FP32_TOL
    px, py, pz = _points(P, 500, seed=20)
    p64 = s.field_at(px, py, pz, nterms, 0.5, 0.01, 0.01, kernel_order=L)
    p32 = s.field_at(px, py, pz, nterms, 0.5, 0.01, 0.01, kernel_order=L, precision="fp32")
    assert rel_l2(p32, p64) < FP32_TOL


def test_field_fp32_reference_fixture(ctx):
    """tests/treecode_mp.cpp's set (boxparticles(16000), maxmemb 100, nterms 5):
    the FP32 field within FP32_TOL of the reference's treeforce values."""
    fx = load_fixture("box16k")
    s = _system(ctx, fx)
    E = s.field(5, 0.5, 0.01, 0.01, precision="fp32")[fx["tidx"]]
    assert rel_l2(E, fx["field_tree"]) < FP32_TOL
