#!/usr/bin/env python3
"""Trotter steps/s of the two-site update on one MI355X.

Workload (the Fermi-Hubbard configuration the metric is quoted on): spin-1/2
Hubbard chain L=128, d=4 (basis [0, up, down, up+down]), t=1, U=4, OBC,
2nd-order Trotter with dt=0.01, chi_max=256, half-filled Neel start, 100 steps.
The bond dimensions grow unevenly from 1 and every bulk bond reaches chi=256 by
step ~37.  The first W=40 steps are untimed warm-up; the next K=60 steps (41-100)
are timed, all at the real shape (191 bond updates of up to 1024 x 1024 per step).

`value` is the batched learned-rank path (randomised SVD of chi_c + chi_c/8 + 8
columns, full SVD fallback; the shipped rank model, trained on a different quench
(L=64, chi=128), sends the bonds it predicts to be rejected to the full SVD at once).  The same 100 steps are also run by the exact batched FP64
path: its steps/s are reported, and the learned path's observables at step 100
(energy, <n_i>, <Sz_i>, entanglement entropy in log_d, cumulative truncation error)
are compared with it (`obs_err`).  The plain per-bond FP64 path (`--compare fp64`,
the default) is timed on the first steps of the timed window, started from the
exact path's state; `speedup_vs_fp64` is value over its steps/s.

The mixed-precision form of the same path (theta and gate on the FP32 matrix
cores, `--mixed-mode`) runs the same 100 steps; its steps/s and its observable
errors against the exact batched path are reported under `mixed`.

The timed run splits every half-sweep over three lanes (three HIP streams driven by
three host threads, `--lanes`); the results are bit-identical to one lane.

Prints one JSON line.  `roofline` is the single kernel with the most device time in
the same timed steps run with one lane (`--roofline-lanes`, so that no launch shares
the GPU with the other lane's): its algorithmic work per launch (HBM bytes for the streaming
kernels -- Jacobi rounds and rotations, pivoted-QR panels and trailing updates --,
FP64 flops for the GEMM-shaped ones) / its average launch duration, both measured
with HIP events around each launch on the context's stream inside the library.
`traffic` is the kernel's measured HBM bytes per launch over the same timed steps,
read from the committed rocprofv3 counter summary (profiles/*_traffic.json) when
one exists for that kernel and workload.
`cpu_baseline` times the FP64 CPU oracle (the reference algorithm: padded chi,
full ZGESVD of every (chi d) x (chi d) matrix) on a bounded sample of bond
updates on this host and extrapolates to one step.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "opentebd-extended_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

L, D, CHI, DT, T_HOP, U_INT = 128, 4, 256, 0.01, 1.0, 4.0
FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)
HBM_PEAK_GBS = 8000.0    # MI355X HBM3E


RANK_MODEL = os.path.join(ROOT, "opentebd-extended_amd", "tebd_amd", "rank_model_fh.npy")
WORKLOAD_TAG = "fh_l128_chi256"
# launch classes timed by the library that are one kernel each: (kernel, roofline bound)
SINGLE_KERNEL = {"theta_gate": ("k_theta_gate<4>", "mfma"), "jacobi_round": ("k_jacobi_round", "hbm"),
                 "qr_panel": ("k_qrp_panel_reg<2>", "hbm"), "sketch_gemm": ("k_zgemm", "mfma"),
                 "jacobi_rotate": ("k_jacobi_rotate", "hbm"), "qr_trailing": ("k_qrp_reflect<0>", "hbm")}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` in the timed steps, from the committed counter
    summary (FETCH_SIZE and WRITE_SIZE passes, scaled by a copy kernel of known size)."""
    path = os.path.join(ROOT, "profiles", f"{WORKLOAD_TAG}_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)["kernels"][kernel]
        return rec["hbm_bytes_per_launch"], os.path.relpath(path, ROOT)
    except (OSError, KeyError, ValueError):
        return None, None


def make_context(mode, rank_model=None):
    from tebd_amd import Context, models as M

    H = M.fermi_hubbard(L, T_HOP, U_INT)
    ctx = Context(L, D, CHI, mode=mode)
    if rank_model is not None and "learned" in mode:
        ctx.set_rank_model(rank_model)
    ctx.upload(*M.product_state(M.neel(L, D, 1, 2), CHI))
    ctx.set_gates(M.construct_propagators(H, DT))
    return ctx


class _Roctx:
    """roctx range around the timed region, so a rocprofv3 --marker-trace run can
    restrict its kernel statistics to exactly the launches timed here."""

    def __init__(self):
        self.lib = None
        try:
            import ctypes

            self.lib = ctypes.CDLL("/opt/rocm/lib/librocprofiler-sdk-roctx.so.1")
            self.lib.roctxRangePushA.argtypes = [ctypes.c_char_p]
        except OSError:
            self.lib = None

    def push(self, name):
        if self.lib:
            self.lib.roctxRangePushA(name.encode())

    def pop(self):
        if self.lib:
            self.lib.roctxRangePop()


def device_observables(ctx, H):
    from tebd_amd import models as M

    _, _, ntot, sz, _ = M.fermion_ops(2)
    return dict(E=ctx.energy(H), n=ctx.one_site_expval(ntot), sz=ctx.one_site_expval(sz), S=ctx.chain_entropy())


def run_mode(mode, steps, warmup, log, state=None, keep_state=False, rank_model=None, lanes=3, tag="timed",
             profile=False, graphs=True):
    from tebd_amd import models as M

    ctx = make_context(mode, rank_model)
    ctx.set_lanes(lanes)
    ctx.set_graphs(graphs)
    if state is not None:
        ctx.upload(*state)
    trunc_warm = 0.0
    tw = time.perf_counter()
    for k in range(warmup):
        trunc_warm += ctx.trotter_step()
        if (k + 1) % 10 == 0 or k + 1 == warmup:
            log(f"[bench] {mode} warmup step {k + 1}/{warmup} mean chi={ctx.bond_dims().mean():.1f} "
                f"({time.perf_counter() - tw:.1f}s)")
    ctx.sync()
    warm_s = time.perf_counter() - tw
    snap = ctx.download() if keep_state else None
    # per-launch HIP events only where they are reported (they also turn off the
    # hipGraph replay of the Jacobi sweeps)
    ctx.set_profiling(profile)
    st0 = ctx.stats()
    rx = _Roctx()
    trunc = 0.0
    rx.push(f"bench_{tag}_{mode}")
    t0 = time.perf_counter()
    for k in range(steps):
        trunc += ctx.trotter_step()
        log(f"[bench] {mode} timed step {k + 1}/{steps} ({time.perf_counter() - t0:.1f}s)")
    ctx.sync()
    elapsed = time.perf_counter() - t0
    rx.pop()
    prof = ctx.profile()
    st1 = ctx.stats()
    dims = ctx.bond_dims()
    tuning = ctx.autotune_info()
    obs = device_observables(ctx, M.fermi_hubbard(L, T_HOP, U_INT)) if state is None else None
    ctx.close()
    delta = {k: st1[k] - st0[k] for k in st1}
    return dict(elapsed=elapsed, prof=prof, dims=dims, trunc=trunc + trunc_warm, stats=delta, warm_s=warm_s,
                snap=snap, obs=obs, tuning=tuning)


def copy_calibration(mib):
    """Device-to-device copies of a known size: the HBM byte counters (FETCH_SIZE,
    WRITE_SIZE) of this kernel against the bytes it must move give their scale."""
    import torch

    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "bytes_written": n * 4, "launches": 5}), flush=True)


def cpu_baseline(sample_bonds, log):
    """Reference algorithm on the host: every bond is a padded (chi d)^2 = 1024^2 update,
    on one thread (the reference is a sequential program)."""
    from threadpoolctl import threadpool_limits

    with threadpool_limits(limits=1):
        return _cpu_baseline(sample_bonds, log)


def _cpu_baseline(sample_bonds, log):
    from oracle.oracle import OracleState
    from tebd_amd import models as M

    Ls = 4
    rng = np.random.default_rng(20261202)
    gam = (rng.normal(size=(Ls, CHI, D, CHI)) + 1j * rng.normal(size=(Ls, CHI, D, CHI))) / CHI
    lam = np.sort(rng.uniform(0.0, 1.0, size=(Ls + 1, CHI)), axis=1)[:, ::-1].copy()
    lam /= np.linalg.norm(lam, axis=1, keepdims=True)
    U = M.construct_propagators(M.fermi_hubbard(Ls, T_HOP, U_INT), DT)
    st = OracleState(gam, lam)
    times = []
    for i in range(sample_bonds):
        t0 = time.perf_counter()
        st.two_site_op(2, U[1])
        times.append(time.perf_counter() - t0)
        log(f"[bench] cpu oracle bond update {i + 1}/{sample_bonds}: {times[-1]:.2f}s")
    t_bond = float(np.mean(times))
    return t_bond


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--compare-steps", type=int, default=3)
    ap.add_argument("--mode", default="batched,learned")
    ap.add_argument("--lanes", type=int, default=3, help="streams / host threads per half-sweep of the timed run")
    ap.add_argument("--graphs", type=int, default=1,
                    help="hipGraph replay of the Jacobi sweeps in the timed runs (0 = eager launches)")
    ap.add_argument("--roofline-lanes", type=int, default=1,
                    help="lanes of the separate run whose per-kernel events give the roofline (0 = use the timed run)")
    ap.add_argument("--reference-mode", default="batched", help="exact path for the observable check ('' = skip)")
    ap.add_argument("--mixed-mode", default="batched,learned,mixed",
                    help="mixed-precision path timed and checked against the exact path ('' = skip)")
    ap.add_argument("--rank-model", default=RANK_MODEL if os.path.exists(RANK_MODEL) else "",
                    help="rank model parameters (.npy) for the learned-rank path ('' = none)")
    ap.add_argument("--compare", default="fp64", help="also time this mode on the same workload ('' = skip)")
    ap.add_argument("--cpu-sample-bonds", type=int, default=3)
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only time a device-to-device copy of MIB MiB (HBM counter calibration)")
    args = ap.parse_args()
    log = (lambda s: None) if args.quiet else (lambda s: print(s, file=sys.stderr, flush=True))

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    model = np.load(args.rank_model) if args.rank_model else None
    res = run_mode(args.mode, args.steps, args.warmup, log, rank_model=model, lanes=args.lanes,
                   profile=not args.roofline_lanes, graphs=bool(args.graphs))
    sps = args.steps / res["elapsed"]
    ref = None
    if args.reference_mode:
        ref = run_mode(args.reference_mode, args.steps, args.warmup, log, keep_state=bool(args.compare),
                       graphs=bool(args.graphs))
    mixed = None
    if args.mixed_mode:
        mixed = run_mode(args.mixed_mode, args.steps, args.warmup, log, rank_model=model, graphs=bool(args.graphs))
    # per-kernel event timings from a run whose launches do not share the GPU with a
    # second lane, so each launch's duration is that kernel's own
    roof_run = res
    if args.roofline_lanes:
        roof_run = run_mode(args.mode, args.steps, args.warmup, log, rank_model=model, lanes=args.roofline_lanes,
                            tag="roofline", profile=True)
    prof = roof_run["prof"]
    dom = max(SINGLE_KERNEL, key=lambda n: prof[n][1])
    kernel, bound = SINGLE_KERNEL[dom]
    n_launch, dom_ms, dom_flops, dom_bytes = prof[dom]
    avg_ms = dom_ms / max(n_launch, 1)
    gbs = (dom_bytes / max(n_launch, 1)) / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    tflops = (dom_flops / max(n_launch, 1)) / (avg_ms * 1e-3) / 1e12 if avg_ms > 0 else 0.0
    traffic, traffic_src = measured_traffic(kernel)
    if bound == "hbm":
        roof = {"bound": "hbm", "kernel": kernel, "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": gbs / HBM_PEAK_GBS}
    else:
        roof = {"bound": "mfma", "kernel": kernel, "achieved": tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": tflops / FP64_PEAK_TFLOPS}
    roof.update({"traffic": traffic, "traffic_source": traffic_src, "avg_launch_ms": avg_ms,
                 "launches": int(n_launch), "bytes_per_launch": dom_bytes / max(n_launch, 1),
                 "flops_per_launch": dom_flops / max(n_launch, 1), "fp64_tflops": tflops,
                 "fp64_frac": tflops / FP64_PEAK_TFLOPS, "hbm_gbs": gbs,
                 "lanes": args.roofline_lanes or args.lanes,
                 "device_ms": {n: prof[n][1] for n in prof}})
    st = res["stats"]
    half_steps = 3 * args.steps

    extra = {}
    def obs_errors(run):
        errs = {k: float(np.max(np.abs(np.asarray(run["obs"][k]) - np.asarray(ref["obs"][k])))) for k in run["obs"]}
        errs["trunc"] = abs(run["trunc"] - ref["trunc"])
        return errs

    if ref is not None:
        errs = obs_errors(res)
        extra["obs_err"] = max(errs.values())
        extra["obs_err_by_observable"] = errs
        extra["obs_reference"] = f"{args.reference_mode} (exact FP64 GPU path), same steps 1-{args.warmup + args.steps}"
        extra[f"{args.reference_mode}_steps_per_s"] = args.steps / ref["elapsed"]
    if mixed is not None:
        extra["mixed"] = {"mode": args.mixed_mode, "steps_per_s": args.steps / mixed["elapsed"]}
        if ref is not None:
            merr = obs_errors(mixed)
            extra["mixed"]["obs_err"] = max(merr.values())
            extra["mixed"]["obs_err_by_observable"] = merr
    if args.compare and ref is not None:
        cmp_res = run_mode(args.compare, args.compare_steps, 0, log, state=ref["snap"], graphs=bool(args.graphs))
        extra[f"{args.compare}_steps_per_s"] = args.compare_steps / cmp_res["elapsed"]
        extra[f"{args.compare}_window"] = f"steps {args.warmup + 1}-{args.warmup + args.compare_steps}"
        extra[f"speedup_vs_{args.compare}"] = sps / extra[f"{args.compare}_steps_per_s"]

    cpu = None
    if args.cpu_sample_bonds > 0:
        t_bond = cpu_baseline(args.cpu_sample_bonds, log)
        n_updates = 2 * (L // 2) + (L - 1) // 2  # odd, even, odd half-sweeps: 64 + 63 + 64
        cpu = {"value": 1.0 / (t_bond * n_updates), "unit": "Trotter steps/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_sample_bonds} padded 1024x1024 bond updates of the FP64 oracle "
                         f"(reference algorithm, single-threaded OpenBLAS zgesvd), {t_bond:.2f}s each, "
                         f"extrapolated x191 bond updates per step"}

    out = {
        "metric": "Trotter steps/s, L=128 chi=256: mixed+batched vs FP64 GPU at <=1e-8 obs err",
        "value": sps,
        "unit": "Trotter steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * res["elapsed"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex128 (f64)",
        "data": "synthetic (Neel product state, Fermi-Hubbard gates)",
        "config": {"workload": "Fermi-Hubbard quench L=128 d=4 t=1 U=4 dt=0.01 chi_max=256, Neel start; "
                               f"timed steps {args.warmup + 1}-{args.warmup + args.steps}",
                   "mode": args.mode, "lanes": args.lanes, "graphs": bool(args.graphs), "gemm_autotune": res["tuning"],
                   "chi_max_reached": int(res["dims"].max()),
                   "mean_bond_dim": float(res["dims"].mean()), "warmup_s": res["warm_s"]},
        "roofline": roof,
        "learned_rank": {"rank_model": os.path.basename(args.rank_model) if args.rank_model else None,
                         "sketches_skipped_by_model": st["sketches_skipped"],
                         "sketch_accepted": st["sketch_accepted"], "full_svd_fallbacks": st["fallbacks"],
                         "mean_sketch_width": st["sketch_cols"] / max(st["sketch_accepted"] + st["fallbacks"], 1)},
        "jacobi": {"sweeps_per_half_step_per_lane": st["jacobi_sweeps"] / (half_steps * max(args.lanes, 1)),
                   "pair_visits": st["pair_visits"], "grams_formed": st["pair_grams"],
                   "pairs_rotated": st["pair_rotations"]},
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
