"""Writes the golden fixtures of tests/golden/*.npz from the reference itself.

Runs in a checkout where oracle/_ref/libbitref.so was built from the
reference's own src/tree3d.cpp (``make -C oracle ref``).  Every expected value
below comes from calling the reference functions; inputs are either the
reference's own generator (boxparticles, tree3d.cpp:1698) or seeded sets from
bit_amd.workloads, stored with the outputs.

    python tests/golden/gen_golden.py

Fixture contents (all arrays):
  x, y, z, q              inputs (q = PARTICLE::tot_charge, weight 1)
  maxmemb, nterms, acc    treecode parameters
  children, nummemb, boxes, leaf_of
                          treemake output: node children [nt,8], member counts,
                          TREE box fields [nt,13] and, for every particle, the
                          leaf holding it (member lists are the ascending
                          indices of the leaves below a node)
  moments                 treemoments cube of a few nodes (mom_nodes)
  tidx                    target particle indices
  phi_*                   +phi at tidx from the reference (see keys)
  inds, sub_*             (subtree fixture) raw values of the calls that start
                          the traversal at node inds[t] for particle tidx[t]
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from bit_amd import workloads  # noqa: E402
from oracle.pyoracle import Reference  # noqa: E402


def tree_arrays(t):
    leaf_of = np.full(int(t.nummemb[0]), -1, dtype=np.int32)
    for i in range(t.numtree):
        if t.children[i, 0] < 0:
            leaf_of[t.node_members(i)] = i
    assert (leaf_of >= 0).all()
    return dict(children=t.children.astype(np.int32), nummemb=t.nummemb.astype(np.int32),
                boxes=t.boxes, leaf_of=leaf_of)


def save(name, **arrs):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **arrs)
    print(f"wrote {path} ({os.path.getsize(path) / 1024:.0f} KiB)")


def main():
    R = Reference()

    # 1. The reference's own test set (tests/treecode_mp.cpp): boxparticles(16000),
    #    maxmemb 100, nterms 5, ACC 0.5 -> tree and treeforce fields.
    x, y, z, q = R.boxparticles(16000)
    t, mom = R.tree(x, y, z, q, 100, 5)
    tidx = workloads.sample_targets(len(x), 2000, seed=11)
    E = R.treeforce(x, y, z, q, 100, 5, 0.5, tidx)
    Eds = R.force_ds_reg(x, y, z, q, tidx)
    sel = np.array([0, 1, t.numtree // 2, t.numtree - 1])
    save("box16k", x=x, y=y, z=z, q=q, maxmemb=100, nterms=5, acc=0.5, **tree_arrays(t),
         mom_nodes=sel, moments=mom[sel], tidx=tidx, field_tree=E, field_ds=Eds)

    # 2. Config-1 shape: 10k uniform, Coulomb delta=0 (shielded_treepot, kappa 0),
    #    nterms 7, acc 0.5, maxmemb 500; all targets; reference direct sum.
    P = workloads.uniform_cube(10_000, seed=1)
    t, mom = R.tree(P.x, P.y, P.z, P.q, 500, 7)
    tidx = np.arange(P.n, dtype=np.int32)
    phi = R.shielded_phi(P.x, P.y, P.z, P.q, 500, 7, 0.5, 0.0, tidx)
    phi_ds = R.shielded_phi_ds(P.x, P.y, P.z, P.q, 0.0, tidx)
    sel = np.arange(t.numtree)
    save("uniform10k", x=P.x, y=P.y, z=P.z, q=P.q, maxmemb=500, nterms=7, acc=0.5,
         **tree_arrays(t), mom_nodes=sel, moments=mom, tidx=tidx, phi_coulomb=phi,
         phi_coulomb_ds=phi_ds)

    # 3. Config-2 shape: 12k uniform, regularized Coulomb delta^2 = DEL = 0.01
    #    (treepotential, p=-1 at the particles), nterms 9, acc 0.5, maxmemb 1000.
    P = workloads.uniform_cube(12_000, seed=2)
    t = R.tree(P.x, P.y, P.z, P.q, 1000, 9, with_moments=False)
    tidx = workloads.sample_targets(P.n, 3000, seed=12)
    phi = R.coulomb_reg_phi(P.x, P.y, P.z, P.q, 1000, 9, 0.5, tidx)
    save("uniform12k_reg", x=P.x, y=P.y, z=P.z, q=P.q, maxmemb=1000, nterms=9, acc=0.5,
         **tree_arrays(t), tidx=tidx, phi_regcoulomb=phi)

    # 4. Config-3 shape: 16 blobs, screened kappa 0.5 (delta 0), nterms 9, acc 0.7,
    #    maxmemb 200.
    P = workloads.gaussian_blobs(12_000, seed=3)
    t = R.tree(P.x, P.y, P.z, P.q, 200, 9, with_moments=False)
    tidx = workloads.sample_targets(P.n, 3000, seed=13)
    phi = R.shielded_phi(P.x, P.y, P.z, P.q, 200, 9, 0.7, 0.5, tidx)
    phi_ds = R.shielded_phi_ds(P.x, P.y, P.z, P.q, 0.5, tidx)
    save("blobs12k_screened", x=P.x, y=P.y, z=P.z, q=P.q, maxmemb=200, nterms=9, acc=0.7,
         kappa=0.5, **tree_arrays(t), tidx=tidx, phi_screened=phi, phi_screened_ds=phi_ds)

    # 5. Anisotropic box: y/z splits after x splits (treemake index quirk),
    #    screened kappa 1, nterms 5, acc 0.5, maxmemb 40.
    P = workloads.anisotropic_box(6000, seed=4)
    t = R.tree(P.x, P.y, P.z, P.q, 40, 5, with_moments=False)
    tidx = workloads.sample_targets(P.n, 2000, seed=14)
    phi = R.shielded_phi(P.x, P.y, P.z, P.q, 40, 5, 0.5, 1.0, tidx)
    save("aniso6k", x=P.x, y=P.y, z=P.z, q=P.q, maxmemb=40, nterms=5, acc=0.5, kappa=1.0,
         **tree_arrays(t), tidx=tidx, phi_screened=phi)

    # 6. Tiny / edge inputs: one particle, maxmemb+1 particles, a deep narrow tree.
    for name, n, mm in (("tiny1", 1, 4), ("tiny5", 5, 4), ("tiny300", 300, 3)):
        P = workloads.uniform_cube(n, seed=20 + n)
        t = R.tree(P.x, P.y, P.z, P.q, mm, 4, with_moments=False)
        tidx = np.arange(n, dtype=np.int32)
        phi = R.shielded_phi(P.x, P.y, P.z, P.q, mm, 4, 0.5, 0.3, tidx)
        save(name, x=P.x, y=P.y, z=P.z, q=P.q, maxmemb=mm, nterms=4, acc=0.5, kappa=0.3,
             **tree_arrays(t), tidx=tidx, phi_screened=phi)

    # 7. Subtree calls (ind != 0): 3k blobs, maxmemb 40, nterms 6, acc 0.5,
    #    kappa 0.7.  The root and up to five nodes per depth (leaves included)
    #    against 40 sampled particles plus one member of each chosen node:
    #    shielded_treepot(ind, i), treepotential(ind, -1, x_i), treeforce(ind, i).
    P = workloads.gaussian_blobs(3000, seed=5)
    mm, nt, acc, kap = 40, 6, 0.5, 0.7
    t = R.tree(P.x, P.y, P.z, P.q, mm, nt, with_moments=False)
    depth = np.zeros(t.numtree, dtype=np.int64)
    for i in range(t.numtree):
        depth[t.children[i][t.children[i] >= 0]] = depth[i] + 1
    nodes = [0]
    for d in range(1, int(depth.max()) + 1):
        cand = np.flatnonzero(depth == d)
        nodes += cand[np.linspace(0, len(cand) - 1, min(5, len(cand))).astype(int)].tolist()
    nodes = np.array(sorted(set(nodes)), dtype=np.int32)
    rng = np.random.default_rng(15)
    targets = set(workloads.sample_targets(P.n, 40, seed=15).tolist())
    for nd in nodes:
        mem = t.node_members(int(nd))
        targets.add(int(mem[rng.integers(len(mem))]))
    targets = np.array(sorted(targets), dtype=np.int32)
    inds = np.repeat(nodes, len(targets)).astype(np.int32)
    tidx = np.tile(targets, len(nodes)).astype(np.int32)
    sh, pp, E = R.subtree_calls(P.x, P.y, P.z, P.q, mm, nt, acc, kap, inds, tidx)
    save("subtree_blobs3k", x=P.x, y=P.y, z=P.z, q=P.q, maxmemb=mm, nterms=nt, acc=acc,
         kappa=kap, **tree_arrays(t), inds=inds, tidx=tidx, sub_shielded_raw=sh,
         sub_treepotential_point=pp, sub_treeforce=E, intG=R.intG)


if __name__ == "__main__":
    main()
