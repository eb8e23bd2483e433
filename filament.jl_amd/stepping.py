"""Coarse z-stepping with fine-step verification and fallback.

The reference advances the field by fine steps dz chosen by
AdaptiveSteps.AStepKerrPlasma (src/modules/AdaptiveSteps.jl:96-107) inside
Filament.main_loop (src/Filament.jl:111-184).  ``CoarseStepper`` advances the
same model by ``factor * dz`` instead, and every ``check_every`` coarse steps
it verifies one of them: from the same saved state it runs the coarse step and
the ``factor`` fine steps it replaces, compares the two fluence profiles
(Parseval on the device, no field download) and keeps the fine result.  When
the relative fluence distance exceeds ``tol`` the stepper falls back to fine
steps until the next check passes.  An optional learned corrector
(``filament_amd.corrector``) is applied after every coarse step.

Two orders of the coarse step's two parts, the nonlinear step N and the
linear step L (Models.jl:97-107):
  "lie"     N(h) L(h), the reference's zstep at h = factor * dz: its splitting
            error grows linearly with h;
  "strang"  L(h/2) N(h) L(h/2), symmetric, second order in h.  Consecutive
            coarse steps share their half linear steps, so a run of n coarse
            steps is L(h/2) [N(h) L(h)]^(n-1) N(h) L(h/2): the same launches as
            n Lie steps plus one half linear step.  The state is a z-step state
            (the reference's) at every check and when ``advance`` returns; in
            between it is offset by half a linear step.
The fine steps (the verification and the fallback) are always the
reference's zstep at dz.

Throughput is counted in fine-step equivalents: one coarse step advances
``factor`` fine steps, and the verification steps are charged to the run.
"""
import numpy as np


SCHEMES = ("lie", "strang")


def coarse_step(model, z, dz, factor, scheme="lie"):
    """One whole coarse step of h = factor * dz from a z-step state at z."""
    h = factor * dz
    if scheme == "lie":
        model.zstep(z + h, h)
    elif scheme == "strang":
        model.zstep_parts(z, 0.0, 0.5 * h)
        model.zstep_parts(z + h, h, 0.5 * h)
    else:
        raise ValueError(f"scheme must be one of {SCHEMES}")


def fluence_distance(F, Fref):
    """Relative L2 distance of two fluence profiles."""
    n = np.linalg.norm(Fref)
    return float(np.linalg.norm(F - Fref) / n) if n > 0 else float(np.linalg.norm(F))


class CoarseStepper:
    # device state slots used by the verification
    SLOT_START = 0

    def __init__(self, model, factor=4, check_every=16, tol=2e-6, corrector=None, scheme="lie"):
        if factor < 1 or int(factor) != factor:
            raise ValueError("factor must be a positive integer")
        if check_every < 1:
            raise ValueError("check_every must be >= 1")
        if scheme not in SCHEMES:
            raise ValueError(f"scheme must be one of {SCHEMES}")
        self.scheme = scheme
        self.m = model
        self.factor = int(factor)
        self.check_every = int(check_every)
        self.tol = float(tol)
        self.corrector = corrector
        self.coarse = True            # current mode
        self.n_coarse = 0             # coarse steps taken (verified ones included)
        self.n_fine = 0               # fine steps taken (verification and fallback)
        self.n_checks = 0
        self.n_fail = 0
        self.last_distance = None
        self._since_check = 0

    def _coarse_step(self, z, dz, first=True, last=True):
        """One coarse step; with the Strang order ``first`` opens a run of
        coarse steps with a half linear step and ``last`` closes it with one
        (a step that is both is a whole symmetric step)."""
        h = self.factor * dz
        if self.scheme == "strang":
            if first:
                self.m.zstep_parts(z, 0.0, 0.5 * h)
            self.m.zstep_parts(z + h, h, 0.5 * h if last else h)
        else:
            self.m.zstep(z + h, h)
        if self.corrector is not None:
            self.corrector.apply()
        self.n_coarse += 1

    def _fine_steps(self, z, dz):
        for k in range(self.factor):
            self.m.zstep(z + (k + 1) * dz, dz)
        self.n_fine += self.factor

    def _verify(self, z, dz):
        """Coarse step and its fine steps from the same state; keeps the fine result."""
        m = self.m
        m.state_save(self.SLOT_START)
        self._coarse_step(z, dz)
        Fc = m.fluence()
        m.state_restore(self.SLOT_START)
        self._fine_steps(z, dz)
        Ff = m.fluence()
        d = fluence_distance(Fc, Ff)
        self.n_checks += 1
        self.last_distance = d
        ok = d <= self.tol
        if not ok:
            self.n_fail += 1
        return ok

    def advance(self, z, dz, nfine):
        """Advance by ``nfine`` fine steps of ``dz`` (a multiple of ``factor``);
        returns the new z."""
        if nfine % self.factor:
            raise ValueError("nfine must be a multiple of factor")
        n = nfine // self.factor
        in_run = False          # a run of coarse steps is open (Strang: offset state)
        for i in range(n):
            if self._since_check + 1 >= self.check_every:
                self.coarse = self._verify(z, dz)
                self._since_check = 0
            elif self.coarse:
                # the run closes at the end of this call and before a check
                last = i == n - 1 or self._since_check + 2 >= self.check_every
                self._coarse_step(z, dz, first=not in_run, last=last)
                in_run = not last
                self._since_check += 1
            else:
                self._fine_steps(z, dz)
                self._since_check += 1
            z += self.factor * dz
        return z

    def stats(self):
        return dict(factor=self.factor, scheme=self.scheme, check_every=self.check_every, tol=self.tol,
                    coarse_steps=self.n_coarse,
                    fine_steps=self.n_fine, checks=self.n_checks, failed_checks=self.n_fail,
                    last_distance=self.last_distance)
