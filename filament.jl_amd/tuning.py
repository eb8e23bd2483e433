"""Autotuner of the Hankel-product tile shapes: the tile table of an ensemble.

The two MFMA products of the RT linear step (the Hankel transform and its
inverse around the propagator) run with 128-row or 64-row tiles of the QDHT
matrix (``fil_set_tiling``).  Which is faster depends on the grid (Nr, and
Nt/2 columns per pulse), on the number of pulses a handle carries and on the
form (FP32 or split-BF16 products).  ``autotune`` times every candidate on the
device with the handles the run will use, from one saved state, keeps the
fastest for each handle and returns the table; ``apply`` sets a table on other
handles of the same grids.  Every tile shape sums each output over k in the
same order, so the table changes the speed of a run, never its result.

No counterpart in the reference: its Hankel transforms are CUBLAS calls,
which pick their kernels themselves (HankelTransforms `P * E`, Fields.jl:108).
"""
import json

from . import lib as L

CANDIDATES = (L.TILING_128, L.TILING_64)
NAMES = {L.TILING_RULE: "rule", L.TILING_128: "128", L.TILING_64: "64"}
# device state slot the tuner uses to restart every candidate from one state
TUNE_SLOT = 3


def tile_key(model):
    """Table key of a handle: grid, pulses per handle and form."""
    return f"Nr{model.grid.Nr}_Nt{model.grid.Nt}_B{getattr(model, 'B', 1)}_form{model.form}_fp{model.precision}"


def _time_linear(model, dz, steps):
    """Mean time (ms) of the two Hankel products over ``steps`` z-steps."""
    model.timing(True)
    for k in range(steps):
        model.zstep((k + 1) * dz, dz)
    t = model.timing_read()
    model.timing(False)
    return (t["dht1_ms"] + t["dht2_ms"]) / max(1, t["count"])


def autotune(models, dz, steps=4, log=None):
    """Time every tile shape for each handle (one warm-up z-step each), leave
    each handle on its fastest shape and its state as it was; returns the tile
    table {key: {"tiling": t, "ms": {shape: ms}}}."""
    table = {}
    for m in models:
        key = tile_key(m)
        m.state_save(TUNE_SLOT)
        times = {}
        for t in CANDIDATES:
            try:
                m.set_tiling(t)
            except L.FilamentError:
                continue          # 128-row tiles need Nr divisible by 128
            m.state_restore(TUNE_SLOT)
            m.zstep(dz, dz)       # warm-up (propagator, code load)
            m.state_restore(TUNE_SLOT)
            times[t] = _time_linear(m, dz, steps)
        best = min(times, key=times.get)
        m.set_tiling(best)
        m.state_restore(TUNE_SLOT)
        table[key] = {"tiling": int(best), "ms": {NAMES[t]: ms for t, ms in times.items()}}
        if log:
            log(f"tiles {key}: {NAMES[best]} ({', '.join(f'{NAMES[t]} {ms * 1e3:.1f} us' for t, ms in times.items())})")
    return table


def apply(models, table):
    """Set the tile shape of every handle whose key the table holds (the rule
    stays on the others); returns how many were set."""
    n = 0
    for m in models:
        e = table.get(tile_key(m))
        if e is not None:
            m.set_tiling(e["tiling"])
            n += 1
    return n


def save(table, path):
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def load(path):
    with open(path) as f:
        return json.load(f)
