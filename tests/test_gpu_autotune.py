"""The autotuner returns the cheapest measured candidate meeting the target."""
import pytest

from bit_amd import autotune, workloads

pytestmark = pytest.mark.gpu


def test_autotune_picks_cheapest_passing(ctx):
    P = workloads.gaussian_blobs(60_000, seed=41)
    best, table = autotune.autotune(ctx, P, kappa=0.5, delta=0.005, target=1e-5,
                                    n0s=(100, 300), nterms_list=(6, 8, 10),
                                    accs=(0.5, 0.35, 0.2), sample=1000)
    assert best is not None and best["err"] <= 0.8e-5
    passing = [r for r in table if r["err"] <= 0.8e-5]
    assert best["ms"] == min(r["ms"] for r in passing)
