#!/usr/bin/env python3
"""Throughput of the RT z-step on one MI355X: z-steps/s at Nr=1024 x Nt=4096.

One fine "step" is Models.zstep (src/modules/Models.jl:75-121) with the
reference's dz on the whole 1024 x 4096 field: plasma scan, t-FFTs and RK3
nonlinear step (row kernel), then the two Hankel-transform GEMMs with the
propagator multiply.  The field is resident in HBM when the timed region
starts.

Forms, each run over the same trajectory from the same initial field: W
untimed warm-up steps, K timed steps, then untimed steps to the trajectory's
end (the workload's 2000 z-steps = 1 m, or W + K when that is longer), where
the errors against the FP64 build are read:
  fp64    the FP64 build, fine steps: the error reference;
  plain   the plain FP32 port, fine steps;
  chosen  FP32 with the split-BF16 Hankel transform and the FP64 plasma scan,
          advanced by coarse steps of F*dz whose nonlinear and linear parts
          run in the symmetric (Strang) order L(h/2) N(h) L(h/2) (consecutive
          steps share their half linear steps: stepping.CoarseStepper), each
          followed by the learned corrector, with a fine-step verification
          (and fallback) every M coarse steps.  Its rate counts fine-step
          equivalents (a coarse step counts F; verification steps are charged
          to the run).  W is rounded up to whole coarse steps and K to whole
          verification periods (M F fine-step equivalents), so every timed
          window pays for its checks.  The corrector is trained on the device
          during set-up, outside the timed region, on FP64 fine-step targets of
          other pulses.  The same step without the network, and coarse steps in
          the reference's zstep order (Lie) with and without a network of their
          own, are reported beside it.

Prints one JSON line.  `value` is the chosen form's fine-equivalent z-steps/s
when its fluence profile stays within 1e-3 relative L2 of the FP64 build's at
the end of the trajectory, otherwise the plain port's z-steps/s; `steps` and
`warmup` are the counts that form actually ran.  Beside it: `adaptive_loop`,
Filament.run's z loop with the adaptive step (plain FP32, 200 steps), driven
from the host and kept on the device (fil_run_adaptive), with the check that
both end on the same z and field; `cpu_baseline`, the oracle's FP32 instance
on the host's cores and on one core.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PEAK = {"fp32_mfma_tflops": 157.3, "bf16_mfma_tflops": 2500.0, "hbm_gbs": 8000.0, "fp32_valu_tflops": 78.6}
# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (fetch x 2.00, write x 1.00,
# calibrated on a 16 MiB device copy), RT 1024x4096: profiles/rt1024x4096_hbm_traffic_pmc.txt
TRAFFIC_BYTES = {"row_kernel": 39.9e6, "dht_gemm_bf16x3": 83.3e6}
GATE = 1e-3


def load_pkg():
    import importlib.util
    pdir = os.path.join(ROOT, "filament.jl_amd")
    spec = importlib.util.spec_from_file_location("filament_amd", os.path.join(pdir, "__init__.py"),
                                                  submodule_search_locations=[pdir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["filament_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def _plain(o):
    """numpy scalars and arrays as plain JSON values."""
    if hasattr(o, "tolist"):
        return o.tolist()
    raise TypeError(f"{type(o).__name__} is not JSON serialisable")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def fluence_error(F, Fref):
    return float(np.linalg.norm(F - Fref) / np.linalg.norm(Fref))


def diagnostics(m, z):
    """End-of-run quantities compared with the FP64 build: the analyzer's
    fluence profile, peak intensity, peak electron density and pulse energy
    (FieldAnalyzers.jl:166-215; its sums in FP64), and the on-axis spectrum
    |E(r=0, w)|^2 of the downloaded field (positive frequencies)."""
    d = m.analyze(z, arrays=True)
    E0 = m.field()[0]
    sp = np.abs(np.fft.ifft(E0)) ** 2
    return dict(Fr=d["Fr"], Imax=d["Imax"], rhomax=d["rhomax"], W=d["W"], onaxis=sp[: len(sp) // 2 + 1])


def errors_vs(ref, d):
    """Relative errors of a form's end-of-run diagnostics against the FP64 build's."""
    def rl2(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    def rs(a, b):
        return float(abs(a - b) / abs(b)) if b != 0 else float(abs(a))

    return dict(fluence=rl2(d["Fr"], ref["Fr"]), peak_intensity=rs(d["Imax"], ref["Imax"]),
                max_electron_density=rs(d["rhomax"], ref["rhomax"]), onaxis_spectrum=rl2(d["onaxis"], ref["onaxis"]),
                energy=rs(d["W"], ref["W"]))


def kernel_ms(kt):
    n = max(1, kt["count"])
    return dict(row=kt["row_ms"] / n, dht1=kt["dht1_ms"] / n, dht2=kt["dht2_ms"] / n)


def run_fine(pkg, p, precision, form, W, K, dz, every, total):
    """W warm-up and K timed fine steps, then untimed fine steps up to
    ``total`` (the trajectory the errors are read at); per-kernel HIP-event
    times over every ``every``-th timed step (events on the library's stream,
    carried by the kernel dispatches)."""
    m = pkg.RTModel(p, precision=precision, form=form)
    W0 = m.analyze(0.0)["W"]
    z = 0.0
    for _ in range(W):
        z += dz
        m.zstep(z, dz)
    m.synchronize()
    m.timing(every)
    t0 = time.perf_counter()
    for _ in range(K):
        z += dz
        m.zstep(z, dz)
    m.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / K
    kt = m.timing_read()
    m.timing(False)
    for _ in range(total - W - K):
        z += dz
        m.zstep(z, dz)
    dg = diagnostics(m, z)
    dg["energy_ratio"] = dg["W"] / W0
    m.close()
    return ms, kernel_ms(kt), dg


def run_coarse(pkg, p, form, W, K, dz, factor, check_every, tol, corr_state, geom, every, total, scheme):
    """The chosen form: coarse steps (+ corrector) with verification; W, K and
    ``total`` in fine-step equivalents (multiples of the factor)."""
    m = pkg.RTModel(p, precision=pkg.lib.FP32, form=form)
    W0 = m.analyze(0.0)["W"]
    corr = None
    if corr_state is not None:
        corr = pkg.corrector.Corrector(m, geom=geom, capacity=1, batch_max=1)
        corr.load_state(corr_state)
    st = pkg.CoarseStepper(m, factor=factor, check_every=check_every, tol=tol, corrector=corr, scheme=scheme)
    z = st.advance(0.0, dz, W)
    m.synchronize()
    s0 = st.stats()
    m.timing(every)
    t0 = time.perf_counter()
    z = st.advance(z, dz, K)
    m.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / K
    kt = m.timing_read()
    m.timing(False)
    s1 = st.stats()
    z = st.advance(z, dz, total - W - K)
    dg = diagnostics(m, z)
    dg["energy_ratio"] = dg["W"] / W0
    m.close()
    timed = dict(coarse_steps=s1["coarse_steps"] - s0["coarse_steps"], fine_steps=s1["fine_steps"] - s0["fine_steps"],
                 checks=s1["checks"] - s0["checks"], failed_checks=s1["failed_checks"] - s0["failed_checks"],
                 last_check_distance=s1["last_distance"])
    return ms, kernel_ms(kt), dg, timed


def train_corrector(pkg, p, form, dz, factor, draws, npairs, iters, geom, scheme):
    C = pkg.corrector

    def make(d):
        return pkg.workloads.rt_air(p["Nr"], p["Nt"], d["power_pcr"], a0=d["a0"], tau0=d["tau0"], seed=d["seed"])

    # pre-collapse pairs only (a trajectory ends once its window peak has grown threefold)
    st, info = C.fit(pkg.RTModel, make, draws, geom, dz, factor, form, npairs=npairs, iters=iters, log=log,
                     stop_growth=3.0, scheme=scheme)
    log(f"corrector: {info}")
    return st, info


def adaptive_loop(pkg, p, steps):
    """Filament.run's z loop (Filament.jl:129-164: adaptive dz from the last
    analysis, zstep, analysis, Istop test) over ``steps`` plain FP32 steps,
    driven from the host (one analysis read back per step) and kept on the
    device (fil_run_adaptive); phimax makes the Kerr rule set dz = 0.9 dz0 at
    the start.  Both loops take the same steps to the bit."""
    f32 = np.float32
    dz0 = p["dz_initial"]
    m = pkg.RTModel(p)
    phimax = 0.9 * dz0 * m.phik * m.analyze(0.0)["Imax"]
    m.close()
    zmax = 1e9
    res = {}
    for kind in ("host", "device"):
        m = pkg.RTModel(p)
        m.zstep(dz0, dz0)              # code load, then back to the start
        m.close()
        m = pkg.RTModel(p)
        m.synchronize()
        t0 = time.perf_counter()
        if kind == "host":
            z = f32(0.0)
            n = 0
            d = m.analyze(0.0)
            for _ in range(steps):
                dz = f32(m.dz_adaptive(dz0, phimax, d["Imax"], d["rhomax"]))
                z = f32(z + dz)
                m.zstep(float(z), float(dz))
                n += 1
                d = m.analyze(float(z))
                if d["Imax"] > p["Istop"]:
                    break
            z = float(z)
        else:
            z, n, _, _ = m.run_adaptive(0.0, zmax, dz0, phimax, p["Istop"], steps)
        m.synchronize()
        res[kind] = (time.perf_counter() - t0, z, m.field(), n)
        m.close()
    (th, zh, Eh, nh), (td, zd, Ed, nd) = res["host"], res["device"]
    return dict(requested_steps=steps, host_loop_steps=nh, device_loop_steps=nd,
                host_loop_steps_per_s=nh / th, device_loop_steps_per_s=nd / td,
                same_z_and_field=bool(nh == nd and zh == zd and np.array_equal(Eh, Ed)), form="plain FP32")


_VENDOR_GEMM = r"""
import json, sys, torch
Nr, Nh = int(sys.argv[1]), int(sys.argv[2])
dev = torch.device("cuda", 0)
out = {}
for name, dt, K in (("bf16_k3nr", torch.bfloat16, 3 * Nr), ("fp32_knr", torch.float32, Nr)):
    A = torch.randn(Nr, K, device=dev).to(dt)
    B = torch.randn(K, 2 * Nh, device=dev).to(dt)
    for _ in range(5):
        C = A @ B
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 50
    e0.record()
    for _ in range(reps):
        C = A @ B
    e1.record()
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / reps
    out[name] = {"us": us, "tflops": 2.0 * Nr * 2 * Nh * K / (us * 1e-6) / 1e12, "M": Nr, "N": 2 * Nh, "K": K}
print(json.dumps(out))
"""


def library_gemm(Nr, Nh, kt):
    """The vendor GEMM (hipBLASLt, through torch.matmul) on the Hankel products' shape, for comparison with the
    library's own products: the split-BF16 form is three BF16 products of M = Nr, N = 2 Nh real columns, K = Nr,
    i.e. one BF16 GEMM with K = 3 Nr; the plain form one FP32 GEMM with K = Nr.  Runs in a child process: torch
    brings its own HIP runtime, which cannot share this process with the one libfilament_hip.so uses."""
    import subprocess
    r = subprocess.run([sys.executable, "-c", _VENDOR_GEMM, str(Nr), str(Nh)], capture_output=True, text=True,
                       timeout=300)
    if r.returncode != 0:
        return {"error": (r.stderr or "")[-500:]}
    out = json.loads(r.stdout.strip().splitlines()[-1])
    if kt:
        out["this_build_split_bf16_products_us"] = {"dht1": kt["dht1"] * 1e3, "dht2": kt["dht2"] * 1e3}
    return out


def cpu_baseline(pkg, p, sample_steps, sample_steps_1core, dz):
    """The oracle's FP32 instance (the plain port restated in C++, OpenMP over
    rows and over Hankel columns) on this host: ``sample_steps`` z-steps on
    every thread OpenMP is given, then ``sample_steps_1core`` on one thread,
    each after one untimed warm-up step (Filament.jl:166-170)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _helpers as H
    orc = H.load_oracle()
    g = pkg.physics.GridRT(p["rmax"], p["Nr"], p["tmin"], p["tmax"], p["Nt"])
    E0 = p["initial_condition"](g.r, g.t, p["ru"], p["tu"], p["Iu"])
    threads = orc.max_threads()

    def timed(nthreads, nsteps):
        orc.set_threads(nthreads)
        log(f"cpu baseline: oracle setup, {nthreads} threads")
        o = orc.Oracle(H.oracle_config(p), E0, 4)
        o.zstep(dz, dz)
        t0 = time.perf_counter()
        for n in range(nsteps):
            o.zstep((n + 2) * dz, dz)
        dt = time.perf_counter() - t0
        log(f"cpu baseline: {nsteps} steps in {dt:.2f} s on {nthreads} threads")
        return dt

    dt = timed(threads, sample_steps)
    one = None
    if sample_steps_1core > 0:
        dt1 = timed(1, sample_steps_1core)
        one = dict(value=sample_steps_1core / dt1, unit="z-steps/s", cores=1,
                   sample=f"{sample_steps_1core} fine z-steps after 1 warm-up step, {dt1:.1f} s wall")
    orc.set_threads(threads)
    return dict(value=sample_steps / dt, unit="z-steps/s", cores=threads, kind="port",
                sample=f"FP32 C++ restatement of the plain port, RT 1024x4096, {sample_steps} fine z-steps after "
                       f"1 warm-up step, {dt:.1f} s wall", one_core=one)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1600,
                    help="timed fine-step equivalents; coarse forms round it up to whole verification periods "
                         "(the default, two periods, keeps warm-up + timed region inside the workload's 1 m)")
    ap.add_argument("--warmup", type=int, default=16,
                    help="warm-up fine-step equivalents; coarse forms round it up to whole coarse steps")
    ap.add_argument("--factor", type=int, default=25, help="coarse step = factor * dz in the chosen form")
    ap.add_argument("--scheme", default="strang", choices=("strang", "lie"),
                    help="order of the coarse step's nonlinear and linear parts in the chosen form: symmetric "
                         "(Strang, L(h/2) N(h) L(h/2), consecutive steps sharing their half steps) or the "
                         "reference's zstep order (Lie, N(h) L(h))")
    ap.add_argument("--check-every", type=int, default=800,
                    help="fine-step equivalents (z distance in dz) between fine-step verifications; "
                         "a factor F checks every max(1, N/F) coarse steps")
    ap.add_argument("--check-tol", type=float, default=1e-4,
                    help="relative fluence distance of one coarse step from its fine steps that triggers the fallback")
    ap.add_argument("--no-corrector", action="store_true")
    ap.add_argument("--corr-draws", type=int, default=4)
    ap.add_argument("--corr-pairs", type=int, default=24)
    ap.add_argument("--corr-iters", type=int, default=1500)
    ap.add_argument("--cpu-steps", type=int, default=30, help="oracle z-steps timed for the CPU baseline (~10 s)")
    ap.add_argument("--cpu-steps-1core", type=int, default=2,
                    help="oracle z-steps timed on one thread for the CPU baseline's 1-core figure (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-library-gemm", action="store_true",
                    help="skip timing the vendor GEMM (hipBLASLt via torch) on the Hankel products' shape")
    ap.add_argument("--adaptive-steps", type=int, default=200,
                    help="steps of the host-driven and the device-kept adaptive z loop (0: skip)")
    ap.add_argument("--factor-sweep", default="lie:16,lie:4",
                    help="other coarse steps run (own corrector each) and reported beside the headline: "
                         "comma list of [scheme:]factor (scheme defaults to --scheme)")
    ap.add_argument("--timing-every", type=int, default=8,
                    help="per-kernel HIP-event timing on every n-th z-step of the timed region (1: every step)")
    ap.add_argument("--forms", default="fp64,plain,chosen",
                    help="comma list of forms (fp64, plain, chosen); anything but all three is a profiling run "
                         "that prints per-form times only")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the RT z-step runs on a single MI355X; "
              "multi-GPU sharding is a design note, not a measured path.", file=sys.stderr)
        sys.exit(2)

    pkg = load_pkg()
    L = pkg.lib
    p = pkg.workloads.rt_1024x4096()
    dz = p["dz_initial"]
    F = args.factor
    Nr, Nt = p["Nr"], p["Nt"]
    Nh = Nt // 2
    chosen_form = L.FORM_DHT_BF16X3 | L.FORM_SCAN_FP64
    sweep = []
    for x in args.factor_sweep.split(","):
        if x.strip():
            sch, _, f = x.strip().rpartition(":")
            e = (sch or args.scheme, int(f))
            if e != (args.scheme, F) and e not in sweep:
                sweep.append(e)
    sweep_factors = [f for _, f in sweep]

    # Step counts actually run.  Fine forms: W and K as asked.  Coarse forms: the warm-up rounded up to
    # whole coarse steps and the timed region rounded up to whole verification periods, so every timed
    # window carries its share of the fine-step checks.  Every form then continues, untimed, to the same
    # trajectory end (the workload's 2000 z-steps = 1 m, or further when W + K asks for more), where the
    # errors against the FP64 build are read.
    def ceil_to(n, q):
        return -(-max(n, 0) // q) * q

    def coarse_counts(f):
        period = max(1, args.check_every // f) * f
        return ceil_to(args.warmup, f), max(period, ceil_to(args.steps, period)), period

    W, K = args.warmup, max(1, args.steps)
    lcm = 1
    for f in [F] + sweep_factors:
        lcm = lcm * f // math.gcd(lcm, f)
    total = ceil_to(max([p["nsteps"], W + K] + [sum(coarse_counts(f)[:2]) for f in [F] + sweep_factors]), lcm)
    log(f"trajectory: {total} fine-step equivalents ({total * dz:.3f} m); fine forms W={W} K={K}; "
        f"chosen x{F}: W, K, period = {coarse_counts(F)}")

    want = set(args.forms.split(","))
    full = want == {"fp64", "plain", "chosen"}
    forms = {}
    dg64 = Fr64 = None

    def errs(dg):
        return errors_vs(dg64, dg) if dg64 is not None else None

    if "fp64" in want:
        log("fp64: fine steps")
        ms64, kt64, dg64 = run_fine(pkg, p, L.FP64, 0, W, K, dz, args.timing_every, total)
        Fr64 = dg64["Fr"]
        forms["fp64"] = dict(ms_per_step=ms64, z_steps_per_s=1e3 / ms64, kernel_ms=kt64, warmup=W, steps=K,
                             energy_ratio_end_to_start=dg64["energy_ratio"])
        log(f"fp64: {ms64:.4f} ms/step")
    if "plain" in want:
        log("plain FP32: fine steps")
        msp, ktp, dgp = run_fine(pkg, p, L.FP32, 0, W, K, dz, args.timing_every, total)
        forms["plain_fp32"] = dict(ms_per_step=msp, z_steps_per_s=1e3 / msp, kernel_ms=ktp, warmup=W, steps=K,
                                   fluence_err_vs_fp64=fluence_error(dgp["Fr"], Fr64) if Fr64 is not None else None,
                                   errors_vs_fp64=errs(dgp), energy_ratio_end_to_start=dgp["energy_ratio"])
        log(f"plain FP32: {msp:.4f} ms/step")
    if "chosen" in want:
        m0 = pkg.RTModel(p, precision=L.FP32, form=chosen_form)
        geom = pkg.corrector.window_geometry(m0.grid, m0.unit, m0.w0)
        m0.close()
        draws = pkg.corrector.ensemble_draws(args.corr_draws, seed=1)

        def coarse_form(f, scheme, with_corrector):
            """One coarse step (factor and order of its parts), with or without a corrector trained for it."""
            cst, cinfo = None, None
            if with_corrector:
                cst, cinfo = train_corrector(pkg, p, chosen_form, dz, f, draws, args.corr_pairs, args.corr_iters,
                                             geom, scheme)
            Wf, Kf, period = coarse_counts(f)
            tag = f"chosen {scheme} x{f}: corrector {'on' if with_corrector else 'off'}"
            log(f"{tag}, W={Wf} K={Kf}")
            ms, kt, dg, timed = run_coarse(pkg, p, chosen_form, Wf, Kf, dz, f, period // f, args.check_tol, cst, geom,
                                           args.timing_every, total, scheme)
            ef = fluence_error(dg["Fr"], Fr64) if Fr64 is not None else None
            log(f"{tag}: {ms:.4f} ms per fine-step equivalent, fluence err {ef}, {timed}")
            return dict(factor=f, scheme=scheme, corrector=with_corrector, ms_per_fine_step_equivalent=ms,
                        z_steps_per_s=1e3 / ms,
                        kernel_ms_per_zstep=kt, warmup=Wf, steps=Kf, fluence_err_vs_fp64=ef, errors_vs_fp64=errs(dg),
                        energy_ratio_end_to_start=dg["energy_ratio"], timed_region=timed,
                        corrector_training=cinfo)

        use_corr = not args.no_corrector
        forms["chosen"] = coarse_form(F, args.scheme, use_corr)
        forms["chosen"]["description"] = (
            f"FP32 + split-BF16 Hankel MFMA + FP64 plasma scan, coarse steps of {F} dz in "
            f"{'the symmetric (Strang) order' if args.scheme == 'strang' else 'the zstep (Lie) order'}"
            f"{' + learned corrector' if use_corr else ''}, fine-step verification every "
            f"{max(1, args.check_every // F)} coarse steps ({args.check_every} dz, tol {args.check_tol:g})")
        # what the learned part and the order earn: the same step without the network, and the other coarse
        # steps (the zstep order, 4 dz the network's design point) with and without a network trained for each;
        # reported only
        rows = [coarse_form(F, args.scheme, False)] if use_corr else []
        for sch, f in sweep:
            if use_corr:
                rows.append(coarse_form(f, sch, True))
            rows.append(coarse_form(f, sch, False))
        forms["chosen"]["factor_sweep"] = [{k: v for k, v in e.items() if k != "kernel_ms_per_zstep"}
                                           for e in rows]
    if not full:
        print(json.dumps({"profiling_forms": forms}, default=_plain))
        return

    err_c = forms["chosen"]["fluence_err_vs_fp64"]
    if err_c <= GATE:
        fc = forms["chosen"]
        name, ms, kt, Wr, Kr = "chosen", fc["ms_per_fine_step_equivalent"], fc["kernel_ms_per_zstep"], fc["warmup"], \
            fc["steps"]
    else:
        fp = forms["plain_fp32"]
        name, ms, kt, Wr, Kr = "plain_fp32", fp["ms_per_step"], fp["kernel_ms"], fp["warmup"], fp["steps"]

    # roofline of the dominant kernel of the reported form (per launch, HIP-event averages)
    gemm_flop = 2.0 * Nr * (2 * Nh) * Nr          # real GEMM M=Nr, N=2*Nh, K=Nr
    row_bytes = 2.0 * 8 * Nr * Nh                 # one read + one write of the c64 half-spectrum
    if max(kt["dht1"], kt["dht2"]) >= kt["row"]:
        kms = 0.5 * (kt["dht1"] + kt["dht2"])
        peak = PEAK["bf16_mfma_tflops"] if name == "chosen" else PEAK["fp32_mfma_tflops"]
        ach = gemm_flop / (kms * 1e-3) / 1e12
        roof = dict(kernel="dht_gemm", bound="mfma", achieved=ach, peak=peak, unit="TFLOP/s", frac=ach / peak,
                    traffic=TRAFFIC_BYTES["dht_gemm_bf16x3"] if name == "chosen" else None,
                    per_launch=f"{gemm_flop:.4g} flop (2*Nr*Nt*Nr)", avg_launch_ms=kms)
    else:
        ach = row_bytes / (kt["row"] * 1e-3) / 1e9
        roof = dict(kernel="row_kernel", bound="hbm", achieved=ach, peak=PEAK["hbm_gbs"], unit="GB/s",
                    frac=ach / PEAK["hbm_gbs"], traffic=TRAFFIC_BYTES["row_kernel"],
                    per_launch=f"{row_bytes:.4g} B (2*8*Nr*Nt/2)",
                    avg_launch_ms=kt["row"])
    roof["timing"] = (f"HIP events carried by the kernel dispatches on the library's stream, every "
                      f"{args.timing_every}th z-step of the timed region")
    # The row kernel is bound by its FP32 vector work, not by HBM: the same launch against the vector peak,
    # with the algorithmic t-transform work of a step (14 transforms of 5 Nt log2 Nt flop per row, the fewest
    # three responses with RK3 need) and the FP32 vector peak without packed instructions (78.6 TFLOP/s; the
    # packed ones are not used on this machine)
    f_fft_step = 14 * 5.0 * Nt * np.log2(Nt) * Nr
    vach = f_fft_step / (kt["row"] * 1e-3) / 1e12
    vector_roof = dict(kernel="row_kernel", bound="valu_fp32", achieved=vach, peak=PEAK["fp32_valu_tflops"],
                       unit="TFLOP/s", frac=vach / PEAK["fp32_valu_tflops"],
                       per_launch=f"{f_fft_step:.4g} flop (14*5*Nt*log2(Nt)*Nr)", avg_launch_ms=kt["row"])

    # whole-step roofline: T_roof = F_DHT / P_MFMA + max(B_step / BW_HBM, F_FFT / P_FP32) with
    # F_DHT = 2 transforms x 4 Nr^2 Nh flop, B_step = 4 x 8 Nr Nh bytes (the c64 half-spectrum read
    # and written once by the row phase and once by the Hankel phase), F_FFT = 14 x 5 Nt log2(Nt) Nr
    # (the fewest t-transforms three responses with RK3 need); the split-BF16 Hankel products run at
    # one third of the BF16 peak.  A coarse step of the chosen form is one z-step's work.
    f_dht = 2 * 4.0 * Nr * Nr * Nh
    b_step = 4 * 8.0 * Nr * Nh
    f_fft = 14 * 5.0 * Nt * np.log2(Nt) * Nr
    t_rest = max(b_step / (PEAK["hbm_gbs"] * 1e9), f_fft / (PEAK["fp32_mfma_tflops"] * 1e12))
    t_plain = f_dht / (PEAK["fp32_mfma_tflops"] * 1e12) + t_rest
    t_bf16 = f_dht / (PEAK["bf16_mfma_tflops"] * 1e12 / 3) + t_rest
    step_roof = {
        "plain_fp32": {"t_roof_us": t_plain * 1e6, "t_measured_us": forms["plain_fp32"]["ms_per_step"] * 1e3,
                       "frac": t_plain * 1e6 / (forms["plain_fp32"]["ms_per_step"] * 1e3)},
    }
    if "chosen" in forms:
        tc = forms["chosen"]["ms_per_fine_step_equivalent"] * F * 1e3
        step_roof["chosen"] = {"t_roof_us_per_coarse_step": t_bf16 * 1e6, "t_measured_us_per_coarse_step": tc,
                               "frac": t_bf16 * 1e6 / tc,
                               "note": "measured time includes the corrector and the amortised verification"}

    adapt = None
    if args.adaptive_steps > 0:
        log("adaptive z loop: host-driven vs kept on the device")
        adapt = adaptive_loop(pkg, p, args.adaptive_steps)
        log(f"adaptive z loop: {adapt}")

    libgemm = None
    if not args.no_library_gemm:
        log("vendor GEMM on the Hankel products' shape")
        libgemm = library_gemm(Nr, Nh, forms["chosen"]["kernel_ms_per_zstep"] if "chosen" in forms else None)
        log(f"vendor GEMM: {libgemm}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(pkg, p, args.cpu_steps, args.cpu_steps_1core, dz)

    out = {
        "metric": "z-steps/s RT 1024x4096: plain FP32 vs chosen form at <=1e-3 fluence err",
        "value": 1e3 / ms,
        "unit": "z-steps/s",
        "n_gpus": 1,
        "steps": Kr,
        "warmup": Wr,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if name == "plain_fp32" else "fp32+bf16x3-mfma+fp64-scan",
        "data": "synthetic seeded Gaussian pulse (RT air, 5 Pcr), 1e-3 amplitude noise",
        "config": {"workload": "RT Nr=1024 x Nt=4096, air 800 nm, Kerr+plasma, RK3, dz=0.5 mm", "form": name,
                   "Nr": Nr, "Nt": Nt, "steps_unit": "fine-step equivalents of dz", "parallelism": "single GPU",
                   "trajectory_steps": total, "errors_at_z_m": total * dz,
                   "requested": {"steps": args.steps, "warmup": args.warmup}},
        "forms": forms,
        "speedup_chosen_vs_plain": forms["plain_fp32"]["ms_per_step"] / forms["chosen"]["ms_per_fine_step_equivalent"],
        "roofline": roof,
        "vector_roofline": vector_roof,
        "step_roofline": step_roof,
        "adaptive_loop": adapt,
        "vendor_gemm_same_shape": libgemm,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out, default=_plain))


if __name__ == "__main__":
    main()
