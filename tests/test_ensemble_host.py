"""Host side of the batched ensemble: the seeded draws follow the stated
distributions, grouping covers every pulse once, and members that do not
share a grid are refused before any device work."""
import numpy as np
import pytest


def test_ensemble_draws_and_groups(pkg):
    W = pkg.workloads
    ps = W.rt_ensemble(64, seed=0)
    assert len(ps) == 64
    for p in ps:
        d = p["draw"]
        assert d["Nr"] in (256, 512, 1024) and d["Nt"] in (1024, 2048, 4096)
        assert 1.0 <= d["power_pcr"] <= 10.0 and 0.5e-3 <= d["a0"] <= 2e-3 and 25e-15 <= d["tau0"] <= 100e-15
        assert (p["Nr"], p["Nt"]) == (d["Nr"], d["Nt"]) and p["power_pcr"] == d["power_pcr"]
    groups = W.group_by_grid(ps)
    ix = sorted(k for _, g in groups for k in g)
    assert ix == list(range(64))
    assert len(groups) == 9   # every (Nr, Nt) pair is drawn at this size
    for (Nr, Nt), g in groups:
        assert all((ps[k]["Nr"], ps[k]["Nt"]) == (Nr, Nt) for k in g)
    # seeded: the same draw twice
    assert [p["draw"] for p in W.rt_ensemble(64, seed=0)] == [p["draw"] for p in ps]
    # the carrier is resolved on the coarsest t grid (analytic signal: w0 dt < pi)
    g = pkg.physics.GridRT(10.0, 256, -500.0, 500.0, 1024)
    assert 2 * np.pi * 299792458.0 / 800e-9 * g.dt * 1e-15 < np.pi


def test_ensemble_refuses_mixed_grids(pkg):
    W = pkg.workloads
    a = W.rt_air(128, 1024, 2.0)
    b = W.rt_air(256, 1024, 2.0)
    with pytest.raises(ValueError, match="Nr"):
        pkg.RTEnsemble([a, b])
