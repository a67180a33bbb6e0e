#!/usr/bin/env python3
"""Trotter steps/s of the two-site update on one MI355X.

Workload (the Fermi-Hubbard configuration the metric is quoted on): spin-1/2
Hubbard chain L=128, d=4 (basis [0, up, down, up+down]), t=1, U=4, OBC,
2nd-order Trotter with dt=0.01, chi_max=256, half-filled Neel start.  The
first W steps of that quench are untimed warm-up: they also grow every bulk
bond to chi=256 (the reference keeps every Schmidt value above 1e-15, so the
bonds saturate within a few steps); the next K steps are timed at that real
shape (191 bond updates of up to 1024 x 1024 per step).

Prints one JSON line.  `roofline` is the Jacobi round kernel (the dominant
kernel): algorithmic FP64 flops per launch / average launch duration, both
measured with HIP events on the context's stream inside the library.
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


def make_context(mode):
    from tebd_amd import Context, models as M

    H = M.fermi_hubbard(L, T_HOP, U_INT)
    ctx = Context(L, D, CHI, mode=mode)
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


def run_mode(mode, steps, warmup, log):
    ctx = make_context(mode)
    tw = time.perf_counter()
    for k in range(warmup):
        ctx.trotter_step()
        log(f"[bench] {mode} warmup step {k + 1}/{warmup} chi_max={ctx.bond_dims().max()} "
            f"({time.perf_counter() - tw:.1f}s)")
    ctx.sync()
    ctx.set_profiling(True)
    rx = _Roctx()
    trunc = 0.0
    rx.push(f"bench_timed_{mode}")
    t0 = time.perf_counter()
    for k in range(steps):
        trunc += ctx.trotter_step()
        log(f"[bench] {mode} timed step {k + 1}/{steps} ({time.perf_counter() - t0:.1f}s)")
    ctx.sync()
    elapsed = time.perf_counter() - t0
    rx.pop()
    prof = ctx.profile()
    dims = ctx.bond_dims()
    ctx.close()
    return dict(elapsed=elapsed, prof=prof, dims=dims, trunc=trunc)


def cpu_baseline(sample_bonds, log):
    """Reference algorithm on the host: every bond is a padded (chi d)^2 = 1024^2 update."""
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
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=6)
    ap.add_argument("--mode", default="batched")
    ap.add_argument("--compare", default="", help="also time this mode (e.g. fp64) on the same workload")
    ap.add_argument("--cpu-sample-bonds", type=int, default=3)
    ap.add_argument("--quiet", action="store_true")
    args = ap.parse_args()
    log = (lambda s: None) if args.quiet else (lambda s: print(s, file=sys.stderr, flush=True))

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    res = run_mode(args.mode, args.steps, args.warmup, log)
    sps = args.steps / res["elapsed"]
    n_upd, jac_ms, jac_flops = res["prof"]["jacobi_round"]
    avg_ms = jac_ms / max(n_upd, 1)
    achieved = (jac_flops / max(n_upd, 1)) / (avg_ms * 1e-3) / 1e12 if avg_ms > 0 else 0.0

    extra = {}
    if args.compare:
        cmp_res = run_mode(args.compare, max(1, args.steps // 3), args.warmup, log)
        extra[f"{args.compare}_steps_per_s"] = max(1, args.steps // 3) / cmp_res["elapsed"]

    cpu = None
    if args.cpu_sample_bonds > 0:
        t_bond = cpu_baseline(args.cpu_sample_bonds, log)
        n_updates = 2 * (L // 2) + (L - 1) // 2  # odd, even, odd half-sweeps: 64 + 63 + 64
        cpu = {"value": 1.0 / (t_bond * n_updates), "unit": "Trotter steps/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_sample_bonds} padded 1024x1024 bond updates of the FP64 oracle "
                         f"(reference algorithm, sequential MKL zgesvd), {t_bond:.2f}s each, "
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
                               "timed steps after the bonds saturate",
                   "mode": args.mode, "chi_max_reached": int(res["dims"].max()),
                   "mean_bond_dim": float(res["dims"].mean())},
        "roofline": {"bound": "mfma", "kernel": "k_jacobi_round", "achieved": achieved,
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": None, "avg_launch_ms": avg_ms, "launches": int(n_upd)},
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
