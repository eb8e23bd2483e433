"""Learned coarse-step corrector on the device (fil_corrector_* of
libfilament_hip.so).  No counterpart in the reference.

A coarse step of ``factor * dz`` differs from the ``factor`` fine steps it
replaces by the splitting error of Models.zstep (src/modules/Models.jl:75-121,
nonlinear step then linear step).  The corrector learns that difference as a
function of the coarse result: a 4-layer 3x3 convolutional network on the
block-averaged envelope of the window (Re A, Im A, |E|^2, Ne) predicts the
block-averaged envelope difference (fine - coarse), which is interpolated and
added to the field.  Weights are BF16 with FP32 accumulation; training runs on
the device (forward, backward and Adam kernels of the library).
"""
import ctypes
import math

import numpy as np

from .stepping import coarse_step

from . import lib as L

# device state slot used while capturing a training pair
CAPTURE_SLOT = 1


def window_geometry(grid, unit, w0, r_window=5.0e-3, t_window=(-256e-15, 256e-15), pooled=(64, 128)):
    """Window rows r <= r_window and samples t in t_window (SI units), sized so
    the pooled grid is ``pooled`` (H, W) blocks whatever the resolution."""
    H, W = pooled
    r = np.asarray(grid.r) * unit.r
    nr_raw = int(np.searchsorted(r, r_window))
    pr = max(1, int(round(nr_raw / H)))
    nr = H * pr
    if nr > grid.Nr:
        raise ValueError("window larger than the r grid")
    t = np.asarray(grid.t) * unit.t
    j0 = int(np.searchsorted(t, t_window[0]))
    j1 = int(np.searchsorted(t, t_window[1]))
    pt = max(1, int(round((j1 - j0) / W)))
    nt = W * pt
    jc = (j0 + j1) // 2
    t0 = max(0, min(grid.Nt - nt, jc - nt // 2))
    return dict(nr=nr, t0=t0, nt=nt, pool_r=pr, pool_t=pt, w0=w0 * unit.t, tmin=float(grid.t[0]),
                dt=float(grid.dt), H=H, W=W)


class Corrector:
    def __init__(self, model, geom=None, capacity=512, batch_max=16, a_scale=1.0, rho_scale=1e-6, seed=0):
        if model.precision != L.FP32:
            raise L.FilamentError("the corrector runs on FP32 models")
        self.m = model
        if geom is None:
            geom = window_geometry(model.grid, model.unit, model.w0)
        self.geom = dict(geom, a_scale=float(a_scale), rho_scale=float(rho_scale))
        g = self.geom
        self.H, self.W = g["nr"] // g["pool_r"], g["nt"] // g["pool_t"]
        d = L.FilCorrDesc()
        d.nr, d.t0, d.nt, d.pool_r, d.pool_t = g["nr"], g["t0"], g["nt"], g["pool_r"], g["pool_t"]
        d.capacity, d.batch_max = capacity, batch_max
        d.w0, d.tmin, d.dt, d.a_scale, d.rho_scale = g["w0"], g["tmin"], g["dt"], a_scale, rho_scale
        self._lib = L.load()
        L.check(self._lib.fil_corrector_init(model.h, ctypes.byref(d), ctypes.c_uint64(seed)))
        self.capacity = capacity
        self.stored = 0

    # -- parameters -----------------------------------------------------------
    def params(self):
        w = np.zeros(L.CORR_NPARAM, dtype=np.float32)
        L.check(self._lib.fil_corrector_params(self.m.h, L.fptr(w), 0))
        return w

    def set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        if w.size != L.CORR_NPARAM:
            raise ValueError(f"expected {L.CORR_NPARAM} parameters")
        L.check(self._lib.fil_corrector_params(self.m.h, L.fptr(w), 1))

    @property
    def y_scale(self):
        y = ctypes.c_double()
        L.check(self._lib.fil_corrector_scale(self.m.h, ctypes.byref(y), 0))
        return y.value

    @y_scale.setter
    def y_scale(self, v):
        y = ctypes.c_double(float(v))
        L.check(self._lib.fil_corrector_scale(self.m.h, ctypes.byref(y), 1))

    def state(self):
        """Trained state as plain arrays (for np.savez)."""
        return dict(params=self.params(), y_scale=np.float64(self.y_scale))

    def load_state(self, st):
        self.set_params(st["params"])
        self.y_scale = float(st["y_scale"])

    # -- data -----------------------------------------------------------------
    def features(self):
        X = np.zeros((4, self.H, self.W), dtype=np.float32)
        L.check(self._lib.fil_corrector_features(self.m.h, L.fptr(X)))
        return X

    def store_pair(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        if X.shape != (4, self.H, self.W) or T.shape != (2, self.H, self.W):
            raise ValueError("pair shapes must be [4][H][W] and [2][H][W]")
        L.check(self._lib.fil_corrector_store_pair(self.m.h, L.fptr(X), L.fptr(T)))
        self.stored = min(self.stored + 1, self.capacity)

    def capture(self, z, dz, factor, scheme="lie"):
        """One training pair at the model's current state: the coarse step (in
        the given order of its parts, stepping.coarse_step) and its fine steps
        from the same state.  Leaves the model at the fine result,
        z + factor * dz.  Returns (X_coarse, T)."""
        m = self.m
        m.state_save(CAPTURE_SLOT)
        coarse_step(m, z, dz, factor, scheme)
        Xc = self.features()
        m.state_restore(CAPTURE_SLOT)
        for k in range(factor):
            m.zstep(z + (k + 1) * dz, dz)
        Xf = self.features()
        return Xc, (Xf[:2] - Xc[:2]).astype(np.float32)

    # -- training -------------------------------------------------------------
    def train(self, iters, batch=8, lr=1e-3, seed=0):
        loss = ctypes.c_double()
        L.check(self._lib.fil_corrector_train(self.m.h, int(iters), int(batch), float(lr), ctypes.c_uint64(seed),
                                              ctypes.byref(loss)))
        return loss.value

    def eval(self, first, count, grad=False):
        loss = ctypes.c_double()
        g = np.zeros(L.CORR_NPARAM, dtype=np.float32) if grad else None
        L.check(self._lib.fil_corrector_eval(self.m.h, int(first), int(count), ctypes.byref(loss), L.fptr(g)))
        return (loss.value, g) if grad else loss.value

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((2, self.H, self.W), dtype=np.float32)
        L.check(self._lib.fil_corrector_forward(self.m.h, L.fptr(X), L.fptr(Y)))
        return Y

    def apply(self):
        L.check(self._lib.fil_corrector_apply(self.m.h))


def ensemble_draws(n, seed=0):
    """Pulse parameters of the ensemble workload: P ~ U[1,10] Pcr,
    a0 ~ U[0.5, 2] mm, tau0 ~ U[25, 100] fs, noise seeds 0..n-1."""
    rng = np.random.default_rng(seed)
    return [dict(power_pcr=float(rng.uniform(1, 10)), a0=float(rng.uniform(0.5e-3, 2e-3)),
                 tau0=float(rng.uniform(25e-15, 100e-15)), seed=k) for k in range(n)]


def capture_vs_fp64(m32, c32, m64, probe, z, dz, factor, scheme="lie"):
    """One training pair whose target is the FP64 build's fine steps: the FP32
    model takes the coarse step, the FP64 model the ``factor`` fine steps from
    the same field, and ``probe`` (a Corrector on a second FP32 model with the
    same window) reads the features of the FP64 result.  The FP32 model is left
    at the FP64 result.  States move between the models on the device
    (fil_state_copy), pending field filter included."""
    m64.copy_state_from(m32)
    coarse_step(m32, z, dz, factor, scheme)
    Xc = c32.features()
    for k in range(factor):
        m64.zstep(z + (k + 1) * dz, dz)
    probe.m.copy_state_from(m64)
    Xf = probe.features()
    m32.copy_state_from(m64)
    return Xc, (Xf[:2] - Xc[:2]).astype(np.float32)


def gather_pairs_fp64(RTModel, p, geom, dz, factor, npairs, spacing, form, scheme="lie"):
    """Training pairs with FP64 fine-step targets along one trajectory of the
    problem ``p`` (FP32 model of the given form between captures)."""
    from . import lib as _L
    m32 = RTModel(p, precision=_L.FP32, form=form)
    m64 = RTModel(p, precision=_L.FP64)
    mp = RTModel(p, precision=_L.FP32, form=form)
    c32 = Corrector(m32, geom=geom, capacity=1, batch_max=1)
    probe = Corrector(mp, geom=geom, capacity=1, batch_max=1)
    z, out = 0.0, []
    try:
        for _ in range(npairs):
            out.append(capture_vs_fp64(m32, c32, m64, probe, z, dz, factor, scheme))
            z += factor * dz
            for _k in range(spacing - factor):
                m32.zstep(z + dz, dz)
                z += dz
    finally:
        for m in (m32, m64, mp):
            m.close()
    return out, z


def gather_pairs(model, corr, dz, factor, npairs, spacing, z0=0.0, scheme="lie"):
    """Training pairs along one fine-step trajectory: every ``spacing`` fine
    steps a coarse-vs-fine capture (which also advances the model)."""
    if spacing < factor:
        raise ValueError("spacing must be >= factor")
    z = z0
    out = []
    for _ in range(npairs):
        X, T = corr.capture(z, dz, factor, scheme)
        z += factor * dz
        out.append((X, T))
        for _k in range(spacing - factor):
            model.zstep(z + dz, dz)
            z += dz
    return out, z


def rms(a):
    return float(math.sqrt(np.mean(np.square(a))))


def fit(RTModel, make_params, draws, geom, dz, factor, form, npairs=32, spacing=None, iters=2000, batch=16,
        lr=1e-3, seed=0, log=None, max_pairs=None, stop_growth=None, scheme="lie"):
    """Gathers FP64-target pairs along the trajectory of every draw and trains
    a corrector on the device.  Returns (state, info); ``state`` loads into any
    Corrector with the same window geometry.

    The draws share the grid and medium of the first one (a draw changes the
    pulse: power, radius, duration, noise), so one set of models serves all of
    them: each draw uploads its initial field (``RTModel.reset_field``).  Pairs
    go to the device as they are made, every ``spacing`` fine steps, up to
    ``npairs`` per draw and ``max_pairs`` in all (default: npairs per draw);
    ``stop_growth``: a trajectory ends once the window's peak intensity
    feature exceeds that multiple of its first value (past the collapse a
    coarse step is no longer a small correction away from the fine steps);
    it also ends at a pair with a non-finite value."""
    import time
    from . import lib as _L
    spacing = spacing or 2 * factor
    cap = int(max_pairs or npairs * len(draws))
    batch = min(batch, cap)
    t0 = time.perf_counter()
    p0 = make_params(draws[0])
    models = [RTModel(p0, precision=_L.FP32, form=form), RTModel(p0, precision=_L.FP64),
              RTModel(p0, precision=_L.FP32, form=form), RTModel(p0, precision=_L.FP32, form=form)]
    m32, m64, mp, tm = models
    stored, sq, ndraws = 0, 0.0, 0
    try:
        c32 = Corrector(m32, geom=geom, capacity=1, batch_max=1)
        probe = Corrector(mp, geom=geom, capacity=1, batch_max=1)
        tc = Corrector(tm, geom=geom, capacity=cap, batch_max=batch, seed=seed)
        for d in draws:
            if stored >= cap:
                break
            ndraws += 1
            m32.reset_field(make_params(d)["initial_condition"])
            z, n, first, dsq = 0.0, 0, None, 0.0
            while n < npairs and stored < cap:
                Xc, T = capture_vs_fp64(m32, c32, m64, probe, z, dz, factor, scheme)
                if not (np.isfinite(Xc).all() and np.isfinite(T).all()):
                    break                      # a step blew up: the trajectory ends here
                peak = float(Xc[2].max())
                first = peak if first is None else first
                if stop_growth and peak > stop_growth * first:
                    break
                tc.store_pair(Xc, T)
                stored += 1
                n += 1
                dsq += float(np.sum(np.square(T))) / T.size
                z += factor * dz
                for _k in range(spacing - factor):
                    m32.zstep(z + dz, dz)
                    z += dz
            sq += dsq
            if log:
                log(f"corrector data: draw {d}: {n} pairs, target rms {math.sqrt(dsq / max(n, 1)):.3e}")
        t1 = time.perf_counter()
        loss = tc.train(iters, batch=min(batch, stored), lr=lr, seed=seed)
        st = tc.state()
    finally:
        for m in models:
            m.close()
    t2 = time.perf_counter()
    info = dict(pairs=stored, draws=ndraws, target_rms=math.sqrt(sq / max(stored, 1)), gather_s=t1 - t0,
                train_s=t2 - t1, last_batch_loss=loss, iters=iters)
    return st, info
