"""bit_eval_potential_noalloc (the evaluation entry point that allocates
nothing) against bit_eval_potential on the same prepared systems: two calls in
a row, the second with a smaller system, bit-identical results; size and state
errors are reported."""
import numpy as np
import pytest

import bit_amd
from bit_amd import workloads

pytestmark = pytest.mark.gpu


def _pair(ctx, s, n, prm_kw):
    """(noalloc result, bit_eval_potential result), both through device outputs."""
    s.prepare(**prm_kw)
    a = bit_amd.DeviceArray(ctx, n)
    b = bit_amd.DeviceArray(ctx, n)
    try:
        s.potential_noalloc(a, **prm_kw)
        s.potential(out=b, **prm_kw)
        return a.numpy(), b.numpy()
    finally:
        a.free()
        b.free()


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_noalloc_equals_eval_two_calls(ctx, precision):
    P1 = workloads.gaussian_blobs(120_000, seed=41)
    P2 = workloads.uniform_cube(30_000, seed=42)
    kw1 = dict(nterms=9, acc=0.5, kappa=0.5, delta2=2.5e-5, precision=precision)
    kw2 = dict(nterms=7, acc=0.6, kappa=0.0, delta2=1e-4, precision=precision)
    s1 = bit_amd.System(ctx, P1.x, P1.y, P1.z, P1.q, 300)
    s2 = bit_amd.System(ctx, P2.x, P2.y, P2.z, P2.q, 200)
    try:
        got1, want1 = _pair(ctx, s1, P1.n, kw1)
        got2, want2 = _pair(ctx, s2, P2.n, kw2)  # second call, smaller input
        assert np.array_equal(got1, want1)
        assert np.array_equal(got2, want2)
        assert np.isfinite(got2).all() and np.abs(got2).max() > 0
        # the first system again after the smaller one: nothing stale in the scratch
        got1b, _ = _pair(ctx, s1, P1.n, kw1)
        assert np.array_equal(got1b, want1)
    finally:
        s1.close()
        s2.close()


def test_noalloc_needs_prepared_system(ctx):
    P = workloads.uniform_cube(5_000, seed=43)
    s = bit_amd.System(ctx, P.x, P.y, P.z, P.q, 100)
    out = bit_amd.DeviceArray(ctx, P.n)
    try:
        with pytest.raises(bit_amd.BitError) as e:
            s.potential_noalloc(out, nterms=6, acc=0.5)
        assert e.value.code == bit_amd.BIT_E_STATE
        s.prepare(nterms=6, acc=0.5)
        with pytest.raises(bit_amd.BitError) as e:  # other MAC parameters
            s.potential_noalloc(out, nterms=6, acc=0.4)
        assert e.value.code == bit_amd.BIT_E_STATE
        s.potential_noalloc(out, nterms=6, acc=0.5)
        assert np.array_equal(out.numpy(), s.potential(nterms=6, acc=0.5))
    finally:
        out.free()
        s.close()


def test_noalloc_rejects_systems_above_the_limit(ctx):
    n = (1 << 22) + 64
    rng = np.random.default_rng(44)
    dev = [bit_amd.DeviceArray.from_numpy(ctx, a) for a in
           (*rng.uniform(-0.5, 0.5, (3, n)), rng.uniform(-1.0, 1.0, n))]
    out = bit_amd.DeviceArray(ctx, n)
    s = bit_amd.System(ctx, *dev, 2000)
    try:
        s.prepare(nterms=4, acc=0.5)
        with pytest.raises(bit_amd.BitError) as e:
            s.potential_noalloc(out, nterms=4, acc=0.5)
        assert e.value.code == bit_amd.BIT_E_INVALID
    finally:
        s.close()
        out.free()
        for d in dev:
            d.free()
