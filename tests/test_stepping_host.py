"""CoarseStepper's sequence of z-step calls, on a recording stand-in for the
model (no device): the Lie order takes one zstep(h) per coarse step; the
Strang order opens every run of coarse steps with a half linear step,
advances by N(h) L(h) and closes the run with N(h) L(h/2) before a check and
at the end of each advance call, so the state is a z-step state there; the
verification takes one whole symmetric step from the saved state."""
import numpy as np
import pytest

import _helpers as H


class Recorder:
    """Records the calls CoarseStepper makes; the fluence never differs, so
    every check passes."""

    def __init__(self):
        self.calls = []

    def zstep(self, z, dz):
        self.calls.append(("N", dz, "L", dz))

    def zstep_parts(self, z, dz_nl, dz_lin):
        self.calls.append(("N", dz_nl, "L", dz_lin))

    def state_save(self, slot):
        self.calls.append(("save",))

    def state_restore(self, slot):
        self.calls.append(("restore",))

    def fluence(self):
        return np.array([1.0, 2.0])


def totals(calls):
    """Total nonlinear and linear distance of a call list."""
    n = sum(c[1] for c in calls if c[0] == "N")
    lin = sum(c[3] for c in calls if c[0] == "N")
    return n, lin


def run(pkg, scheme, factor, check_every, chunks, dz=1.0):
    m = Recorder()
    st = pkg.CoarseStepper(m, factor=factor, check_every=check_every, tol=1.0, scheme=scheme)
    z = 0.0
    for n in chunks:
        z = st.advance(z, dz, n)
    return m.calls, st.stats(), z


@pytest.fixture(scope="module")
def pkg():
    return H.load_pkg()


def test_lie_is_one_zstep_per_coarse_step(pkg):
    calls, stats, z = run(pkg, "lie", 4, 1000, [32])
    assert calls == [("N", 4.0, "L", 4.0)] * 8
    assert stats["coarse_steps"] == 8 and z == 32.0


def test_strang_run_shares_its_half_linear_steps(pkg):
    calls, stats, z = run(pkg, "strang", 4, 1000, [16])
    assert calls == [("N", 0.0, "L", 2.0), ("N", 4.0, "L", 4.0), ("N", 4.0, "L", 4.0), ("N", 4.0, "L", 4.0),
                     ("N", 4.0, "L", 2.0)]
    assert totals(calls) == (16.0, 16.0)
    assert stats["coarse_steps"] == 4


def test_strang_runs_close_at_each_advance_call(pkg):
    calls, _, z = run(pkg, "strang", 4, 1000, [8, 4, 8])
    half, whole, close = ("N", 0.0, "L", 2.0), ("N", 4.0, "L", 4.0), ("N", 4.0, "L", 2.0)
    assert calls == [half, whole, close, half, close, half, whole, close]
    assert totals(calls) == (20.0, 20.0) and z == 20.0


def test_strang_runs_close_before_a_check(pkg):
    # check_every = 3: slots 0, 1 coarse, slot 2 the check (a whole symmetric step
    # from the saved state, then the factor fine steps), slots 3, 4 coarse, slot 5 the check
    calls, stats, z = run(pkg, "strang", 2, 3, [12])
    half, whole, close = ("N", 0.0, "L", 1.0), ("N", 2.0, "L", 2.0), ("N", 2.0, "L", 1.0)
    fine = ("N", 1.0, "L", 1.0)
    check = [("save",), half, close, ("restore",), fine, fine]
    assert calls == [half, whole, close] + check + [half, whole, close] + check
    assert stats["checks"] == 2 and stats["coarse_steps"] == 6 and stats["fine_steps"] == 4
    # what the kept state went through: the two runs and the two kept fine-step pairs
    kept = [half, whole, close, fine, fine, half, whole, close, fine, fine]
    assert totals(kept) == (12.0, 12.0)


def test_scheme_is_checked(pkg):
    with pytest.raises(ValueError):
        pkg.CoarseStepper(Recorder(), factor=2, scheme="midpoint")
