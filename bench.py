#!/usr/bin/env python3
"""Trotter steps/s of the two-site update on one MI355X, at the saturated shape.

Workload (the Fermi-Hubbard configuration the metric is quoted on): spin-1/2
Hubbard chain L=128, d=4 (basis [0, up, down, up+down]), t=1, U=4, OBC,
2nd-order Trotter with dt=0.01, chi_max=256, half-filled Neel start.

Protocol of every timed run: from the Neel state, Trotter steps are taken until
every bulk bond (bonds 4 .. L-4, whose cap min(4^b, 4^(L-b), 256) is 256) has
reached chi=256 -- whatever --warmup says --, then --warmup more untimed steps,
then exactly --steps timed steps.  The bulk bonds are checked again after the
timed steps; a run whose timed window is not at chi=256 throughout exits with an
error instead of printing a number.  Each timed step is 191 bond updates of up to
1024 x 1024 (64 odd, 63 even, 64 odd bonds).

`value` is the mixed-precision batched learned-rank path: theta's contraction over
the middle bond on the int8 matrix cores from exact power-of-two slices of the FP64
operands (integer accumulation, FP64 combination), the gate and everything after
theta in FP64 (randomised SVD sized per bond, full-SVD fallback, the shipped rank
model), over the timed steps.  Beside it, on the same protocol and window:
  * the exact batched FP64 path (full SVD of every bond) -- its steps/s, and the
    timed path's observables after the last timed step against it
    (`obs_err_vs_exact_gpu`: energy, <n_i>, <Sz_i>, entropy in log_d, truncerr);
  * the same learned-rank path with theta on the FP64 matrix cores;
  * the FP32 form (theta and gate on the FP32 matrix cores, FP32 accumulation);
  * the plain per-bond FP64 path (`fp64`, the drop-in denominator), timed over
    --compare-steps steps from the exact path's state at the start of the window.
`obs_err_vs_oracle` checks each path's bond update against the FP64 CPU oracle:
from the saturated state at the start of the timed window, the Trotter gate is
applied to a few bulk bonds by each path and by the oracle, and the resulting
lambdas and local observables are compared (`meets_1e-8` per path).

`roofline` is the single kernel with the most device time in a separate one-lane,
eagerly launched run of the same protocol and window, with HIP events around each
launch on the stream it runs on; its algorithmic bytes (streaming kernels) or FP64
flops (GEMM-shaped kernels) per launch over its average launch duration.
`traffic` is its measured HBM bytes per launch, read from the committed rocprofv3
counter summary (profiles/*_traffic.json).  `--roofline-only` runs just that run
(the command the committed rocprofv3 kernel trace is taken of); `--rocprof-summary
DIR` turns rocprofv3 CSV output into the tables kept under profiles/.

`number_conserving` is the same chain with the particle number conserved
(TrotterStepNC, numbers [0, 1, 1, 2], total L): exact batched FP64 block SVDs, the
same protocol (its 1e-10 cut of the values saturates the bulk later), steps/s, a
three-bond check against the oracle's TwoSiteOpNC, and that oracle's one-thread time
as its own CPU figure.  `--nc-only` runs just this part, `--nc-steps 0` skips it.

`cpu_baseline` times the FP64 CPU oracle (the reference algorithm: padded chi,
full ZGESVD of every (chi d) x (chi d) bond matrix with the reference's minimum
workspace) on this host: one thread (the reference is a sequential program) on a
few bond updates, and all cores (OpenMP over the bonds of each half-sweep,
sequential LAPACK per thread) on one step of a short chain; both are extrapolated
to the 191 updates of one step.
"""
import argparse
import csv
import glob
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "opentebd-extended_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

L, D, CHI, DT, T_HOP, U_INT = 128, 4, 256, 0.01, 1.0, 4.0
FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 matrix
HBM_PEAK_GBS = 8000.0    # MI355X HBM3E
N_UPDATES = 2 * (L // 2) + (L - 1) // 2  # bond updates per step: 64 + 63 + 64
BULK = slice(4, L - 3)   # bonds 4 .. L-4 of bond_dims(): cap min(4^b, 4^(L-b), CHI) = CHI
MAX_SATURATION_STEPS = 200
ORACLE_BONDS = (32, 64, 97)  # 1-based bulk bonds of the oracle check
OBS_TOL = 1e-8
NC_CONSERV = np.array([0, 1, 1, 2], np.int32)  # particle number of [0, up, down, up+down]

RANK_MODEL = os.path.join(ROOT, "opentebd-extended_amd", "tebd_amd", "rank_model_fh.npy")
WORKLOAD_TAG = "fh_l128_chi256"
# launch classes timed by the library, each counted as one kernel: (kernel, roofline bound)
# (base names as rocprofv3 reports them, without template arguments).  The theta class is
# k_theta_split with its slicing kernels in the split mode and k_theta_gate_mixed in the
# FP32 mode (theta_kernel).
SINGLE_KERNEL = {"theta_gate": ("k_theta_gate_mfma", "mfma"), "jacobi_round": ("k_jacobi_round", "hbm"),
                 "qr_panel": ("k_qrp_panel_reg", "hbm"), "sketch_gemm": ("k_zgemm_mfma", "mfma"),
                 "jacobi_rotate": ("k_jacobi_rotate", "hbm"), "qr_trailing": ("k_qrp_trailing_reg", "hbm")}


class BenchError(RuntimeError):
    pass


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` in the timed steps, from the committed counter
    summary (FETCH_SIZE and WRITE_SIZE passes, scaled by a copy kernel of known size)."""
    path = os.path.join(ROOT, "profiles", f"{WORKLOAD_TAG}_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)["kernels"]
    except (OSError, KeyError, ValueError):
        return None, None
    r = rec.get(kernel)
    if not r or "FETCH_SIZE_bytes" not in r or "WRITE_SIZE_bytes" not in r:
        return None, None
    return r["hbm_bytes_per_launch"], os.path.relpath(path, ROOT)


def gates_and_h():
    from tebd_amd import models as M

    H = M.fermi_hubbard(L, T_HOP, U_INT)
    return H, M.construct_propagators(H, DT)


def make_context(mode, rank_model=None):
    from tebd_amd import Context, models as M

    _, U = gates_and_h()
    ctx = Context(L, D, CHI, mode=mode)
    if rank_model is not None and "learned" in mode:
        ctx.set_rank_model(rank_model)
    ctx.upload(*M.product_state(M.neel(L, D, 1, 2), CHI))
    ctx.set_gates(U)
    return ctx


class _Roctx:
    """roctx range around the timed region, for a rocprofv3 --marker-trace run."""

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


def observables(ctx, H):
    from tebd_amd import models as M

    _, _, ntot, sz, _ = M.fermion_ops(2)
    return dict(E=ctx.energy(H), n=ctx.one_site_expval(ntot), sz=ctx.one_site_expval(sz), S=ctx.chain_entropy())


def bulk_min(ctx):
    return int(ctx.bond_dims()[BULK].min())


def run_mode(mode, steps, warmup, log, pre_steps=None, keep_state=False, rank_model=None, lanes=3, tag="timed",
             profile=False, graphs=True, state=None):
    """One timed run.  Without `state`: from the Neel state, steps until the bulk is
    saturated (or exactly `pre_steps` steps, then checked), `warmup` more, then `steps`
    timed.  With `state` (a saturated snapshot): `steps` timed steps from it.  n_sat in
    the result is the number of steps to saturation, n_pre all untimed steps."""
    ctx = make_context(mode, rank_model)
    ctx.set_lanes(lanes)
    ctx.set_graphs(graphs)
    n_pre = n_sat = 0
    trunc_pre = 0.0
    tw = time.perf_counter()
    if state is not None:
        ctx.upload(*state)
    else:
        while (pre_steps is None and bulk_min(ctx) < CHI) or (pre_steps is not None and n_pre < pre_steps):
            trunc_pre += ctx.trotter_step()
            n_pre += 1
            if n_pre % 10 == 0:
                log(f"[bench] {mode} saturation step {n_pre}: bulk min chi={bulk_min(ctx)} "
                    f"({time.perf_counter() - tw:.1f}s)")
            if n_pre > MAX_SATURATION_STEPS:
                raise BenchError(f"{mode}: bulk bonds not saturated after {n_pre} steps")
        if bulk_min(ctx) < CHI:
            raise BenchError(f"{mode}: bulk bonds not saturated after the {n_pre} pre-steps")
        n_sat = n_pre
        for k in range(warmup):
            trunc_pre += ctx.trotter_step()
            n_pre += 1
        log(f"[bench] {mode}: saturated after {n_sat} steps, {warmup} warm-up steps "
            f"({time.perf_counter() - tw:.1f}s)")
    ctx.sync()
    warm_s = time.perf_counter() - tw
    snap = ctx.download() if keep_state else None
    chi_start = bulk_min(ctx)
    ctx.set_profiling(profile)
    st0 = ctx.stats()
    rx = _Roctx()
    trunc = 0.0
    rx.push(f"bench_{tag}_{mode}")
    t0 = time.perf_counter()
    for k in range(steps):
        trunc += ctx.trotter_step()
        log(f"[bench] {mode} {tag} step {k + 1}/{steps} ({time.perf_counter() - t0:.2f}s)")
    ctx.sync()
    elapsed = time.perf_counter() - t0
    rx.pop()
    prof = ctx.profile()
    st1 = ctx.stats()
    dims = ctx.bond_dims()
    if min(chi_start, int(dims[BULK].min())) < CHI:
        raise BenchError(f"{mode}: timed window left chi={CHI} (bulk min {chi_start} -> {dims[BULK].min()})")
    tuning = ctx.autotune_info()
    H, _ = gates_and_h()
    obs = observables(ctx, H) if state is None else None
    ctx.close()
    delta = {k: st1[k] - st0[k] for k in st1}
    return dict(elapsed=elapsed, prof=prof, dims=dims, trunc=trunc + trunc_pre, stats=delta, warm_s=warm_s,
                snap=snap, obs=obs, tuning=tuning, n_pre=n_pre, n_sat=n_sat)


def oracle_check(snap, modes, log, rank_model=None):
    """Each path's bond update against the FP64 oracle: the Trotter gates of
    ORACLE_BONDS applied to the saturated snapshot by the path and by the oracle; the
    lambdas of those bonds and the local observables of both states (evaluated by the
    FP64 device kernels) are compared."""
    from oracle.oracle import OracleState
    from tebd_amd import Context

    H, U = gates_and_h()
    t0 = time.perf_counter()
    ref = OracleState(snap[0].copy(), snap[1].copy())  # the oracle updates its arrays in place
    for b in ORACLE_BONDS:
        ref.two_site_op(b, U[b - 1])
    log(f"[bench] oracle updates of bonds {ORACLE_BONDS}: {time.perf_counter() - t0:.1f}s")

    def obs_of(gam, lam):
        with Context(L, D, CHI, mode="fp64") as ev:
            ev.upload(gam, lam)
            o = observables(ev, H)
            o["Eb"] = ev.expval_two_site(H).real
        return o

    ref_obs = obs_of(ref.gam, ref.lam)
    out = {}
    for mode in modes:
        ctx = make_context(mode, rank_model)
        ctx.upload(*snap)
        tr = 0.0
        for b in ORACLE_BONDS:
            tr += ctx.two_site_op(b, U[b - 1])
        g, lam = ctx.download()
        st = ctx.stats()
        ctx.close()
        o = obs_of(g, lam)
        errs = {"lambda": float(max(np.max(np.abs(lam[b] - ref.lam[b])) for b in ORACLE_BONDS))}
        errs["E_total"] = abs(o["E"] - ref_obs["E"])
        for k in ("Eb", "n", "sz", "S"):
            errs[k] = float(np.max(np.abs(np.asarray(o[k]) - np.asarray(ref_obs[k]))))
        out[mode] = {"max_err": max(errs.values()), "by_observable": errs,
                     "meets_1e-8": bool(max(errs.values()) <= OBS_TOL),
                     "sketches_accepted": st["sketch_accepted"], "full_svd_fallbacks": st["fallbacks"]}
        log(f"[bench] oracle check {mode}: max err {out[mode]['max_err']:.2e}")
    return out


def sector_spectra_err(LL, lam, LL_ref, lam_ref):
    """Largest difference between the Schmidt values of each number sector of each link
    (which of two equal values of one sector comes first is left to rounding)."""
    worst = 0.0
    for link in range(LL.shape[0]):
        for n in set(LL[link][lam[link] > 0].tolist()) | set(LL_ref[link][lam_ref[link] > 0].tolist()):
            a = np.sort(lam[link][(LL[link] == n) & (lam[link] > 0)])[::-1]
            b = np.sort(lam_ref[link][(LL_ref[link] == n) & (lam_ref[link] > 0)])[::-1]
            k = max(a.size, b.size)
            a, b = np.pad(a, (0, k - a.size)), np.pad(b, (0, k - b.size))
            worst = max(worst, float(np.max(np.abs(a - b))) if k else 0.0)
    return worst


def run_nc(steps, warmup, log, graphs=True, sample_bonds=ORACLE_BONDS):
    """The number-conserving form of the same workload (TrotterStepNC,
    propagation_module.f90:890-1018; TwoSiteOpNC_c, local_operations_module.f90:1845-1964):
    the Hubbard chain with the particle number of the basis [0, up, down, up+down] =
    [0, 1, 1, 2] conserved, total L from the Neel state, every bond's theta factorised
    block by block.  Saturation, warm-up and the timed window as in run_mode (exact FP64,
    batched).  Then from the state at the start of the window the gates of `sample_bonds`
    are applied by the device and by the oracle (one thread, timed: the CPU figure); their
    lambdas, the local observables, every link's number-sector spectra and the labels are
    compared."""
    from threadpoolctl import threadpool_limits

    from oracle.oracle import OracleState
    from tebd_amd import Context, models as M

    H, U = gates_and_h()
    _, _, nop, _, _ = M.fermion_ops(2)
    coeffs = M.neel(L, D, 1, 2)
    tot = L
    mode = "batched"

    def nc_context(state, labels):
        ctx = Context(L, D, CHI, mode=mode)
        ctx.upload(*state)
        ctx.set_gates(U)
        ctx.set_graphs(graphs)
        ctx.set_conserved(NC_CONSERV, tot, labels=labels)
        return ctx

    ctx = nc_context(M.product_state(coeffs, CHI), M.product_labels(coeffs, CHI, NC_CONSERV, tot))
    tw = time.perf_counter()
    n_pre = 0
    while bulk_min(ctx) < CHI:
        ctx.trotter_step()
        n_pre += 1
        if n_pre % 10 == 0:
            log(f"[bench] NC saturation step {n_pre}: bulk min chi={bulk_min(ctx)} ({time.perf_counter() - tw:.1f}s)")
        if n_pre > MAX_SATURATION_STEPS:
            raise BenchError(f"NC: bulk bonds not saturated after {n_pre} steps")
    n_sat = n_pre
    for _ in range(warmup):
        ctx.trotter_step()
        n_pre += 1
    ctx.sync()
    log(f"[bench] NC: saturated after {n_sat} steps, {warmup} warm-up steps ({time.perf_counter() - tw:.1f}s)")
    snap, snap_labels = ctx.download(), ctx.labels(CHI)
    st0 = ctx.stats()
    t0 = time.perf_counter()
    for k in range(steps):
        ctx.trotter_step()
        log(f"[bench] NC timed step {k + 1}/{steps} ({time.perf_counter() - t0:.2f}s)")
    ctx.sync()
    elapsed = time.perf_counter() - t0
    st1 = ctx.stats()
    dims = ctx.bond_dims()
    if int(dims[BULK].min()) < CHI:
        raise BenchError(f"NC: timed window left chi={CHI} (bulk min {dims[BULK].min()})")
    number_err = float(abs(np.sum(ctx.one_site_expval(nop)) - tot))
    ctx.close()

    ref = OracleState(snap[0].copy(), snap[1].copy())
    ref.set_conserved(snap_labels[0].copy(), snap_labels[1].copy(), NC_CONSERV, tot)
    times = []
    with threadpool_limits(limits=1):
        for b in sample_bonds:
            tc = time.perf_counter()
            ref.two_site_op_nc(b, U[b - 1])
            times.append(time.perf_counter() - tc)
    log(f"[bench] NC oracle updates of bonds {tuple(sample_bonds)}: {sum(times):.1f}s")
    dev = nc_context(snap, snap_labels)
    for b in sample_bonds:
        dev.two_site_op(b, U[b - 1])
    g, lam = dev.download()
    LL, LR = dev.labels(CHI)
    dev.close()
    with Context(L, D, CHI, mode="fp64") as ev:
        ev.upload(g, lam)
        o = observables(ev, H)
        ev.upload(ref.gam, ref.lam)
        r = observables(ev, H)
    errs = {"lambda": float(max(np.max(np.abs(lam[b] - ref.lam[b])) for b in sample_bonds)),
            "E_total": abs(o["E"] - r["E"]),
            "sectors": sector_spectra_err(LL, lam, ref.LL, ref.lam)}
    for k in ("n", "sz", "S"):
        errs[k] = float(np.max(np.abs(np.asarray(o[k]) - np.asarray(r[k]))))
    t_bond = float(np.mean(times))
    model, _ = host_cpu()
    return {"workload": f"Fermi-Hubbard quench L={L} d={D} t=1 U=4 dt={DT} chi_max={CHI}, Neel start, particle "
                        f"number conserved (numbers [0,1,1,2], total {tot}); timed steps {n_pre + 1}-"
                        f"{n_pre + steps} with every bulk bond at chi={CHI}",
            "mode": mode, "graphs": graphs, "steps": steps, "steps_per_s": steps / elapsed, "ms_per_step": 1e3 * elapsed / steps,
            "saturation_steps": n_sat, "mean_bond_dim": float(dims.mean()),
            "block_svd_batches_per_step": float(np.sum(st1["solve_sweeps"] - st0["solve_sweeps"])) / steps,
            "jacobi_sweeps_per_step": (st1["jacobi_sweeps"] - st0["jacobi_sweeps"]) / steps,
            "particle_number_err": number_err,
            "obs_err_vs_oracle": {"max_err": max(errs.values()), "by_observable": errs,
                                  "labels_equal": bool(np.array_equal(LL, ref.LL) and np.array_equal(LR, ref.LR)),
                                  "meets_1e-8": bool(max(errs.values()) <= OBS_TOL), "bonds": list(sample_bonds)},
            "cpu_baseline": {"value": 1.0 / (t_bond * N_UPDATES), "unit": "Trotter steps/s", "cores": 1,
                             "kind": "port", "cpu_model": model,
                             "sample": f"the NC updates of bonds {tuple(sample_bonds)} above by the FP64 oracle "
                                       f"(TwoSiteOpNC on padded chi={CHI}) on one thread, {t_bond:.2f}s each, "
                                       f"x{N_UPDATES} bond updates per step"}}


def copy_calibration(mib):
    """Device-to-device copies of a known size through the HIP runtime the library uses
    (its blit kernel reads and writes mib MiB each): the HBM byte counters (FETCH_SIZE,
    WRITE_SIZE) of these dispatches against the bytes they must move give their scale."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    n = mib * (1 << 20)
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(src), n) or hip.hipMalloc(ctypes.byref(dst), n):
        raise BenchError("copy calibration: allocation failed")
    hip.hipMemset(src, 1, n)
    for _ in range(5):
        if hip.hipMemcpy(dst, src, n, 3):  # device to device
            raise BenchError("copy calibration: copy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_calibration_bytes_read": n, "bytes_written": n, "launches": 5}), flush=True)


def host_cpu():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, avail


def random_padded_state(Ls, seed=20261202):
    rng = np.random.default_rng(seed)
    gam = (rng.normal(size=(Ls, CHI, D, CHI)) + 1j * rng.normal(size=(Ls, CHI, D, CHI))) / CHI
    lam = np.sort(rng.uniform(0.0, 1.0, size=(Ls + 1, CHI)), axis=1)[:, ::-1].copy()
    lam /= np.linalg.norm(lam, axis=1, keepdims=True)
    return gam, lam


def cpu_baseline(sample_bonds, threads, log):
    """Reference algorithm on the host, every bond a padded (chi d)^2 = 1024^2 update.
    One thread: `sample_bonds` updates.  All cores: one Trotter step of a chain of
    2 * threads sites (3 threads - 1 updates in three half-sweeps, the bonds of each
    half-sweep in parallel, sequential LAPACK per thread)."""
    from threadpoolctl import threadpool_limits

    from oracle.oracle import OracleState
    from tebd_amd import models as M

    model, avail = host_cpu()
    res = {}
    with threadpool_limits(limits=1):
        Ls = 4
        U = M.construct_propagators(M.fermi_hubbard(Ls, T_HOP, U_INT), DT)
        st = OracleState(*random_padded_state(Ls))
        times = []
        for i in range(sample_bonds):
            t0 = time.perf_counter()
            st.two_site_op(2, U[1])
            times.append(time.perf_counter() - t0)
            log(f"[bench] cpu oracle bond update {i + 1}/{sample_bonds}: {times[-1]:.2f}s")
        t_bond = float(np.mean(times))
        res["one"] = {"value": 1.0 / (t_bond * N_UPDATES), "unit": "Trotter steps/s", "cores": 1, "kind": "port",
                      "cpu_model": model,
                      "sample": f"{sample_bonds} padded 1024x1024 bond updates of the FP64 oracle on one thread "
                                f"(MKL sequential zgesvd with the reference's minimum workspace), {t_bond:.2f}s each, "
                                f"x{N_UPDATES} bond updates per step"}
        if threads > 1:
            Ls = 2 * threads
            U = M.construct_propagators(M.fermi_hubbard(Ls, T_HOP, U_INT), DT)
            st = OracleState(*random_padded_state(Ls, seed=7))
            t0 = time.perf_counter()
            st.trotter_step(U, nthreads=threads)
            t_step = time.perf_counter() - t0
            per_update = t_step / (3 * threads - 1)
            log(f"[bench] cpu oracle step of {Ls} sites on {threads} threads: {t_step:.2f}s")
            res["all"] = {"value": 1.0 / (per_update * N_UPDATES), "unit": "Trotter steps/s", "cores": threads,
                          "kind": "port", "cpu_model": model, "cpus_available": avail,
                          "sample": f"one Trotter step of a {Ls}-site chain at padded chi={CHI} ({3 * threads - 1} "
                                    f"bond updates of 1024x1024, OpenMP over the bonds of each half-sweep, "
                                    f"sequential MKL zgesvd per thread) in {t_step:.2f}s, "
                                    f"x{N_UPDATES}/{3 * threads - 1}"}
    return res


def theta_kernel(mode):
    return "k_theta_split" if "split" in mode else ("k_theta_gate_mixed" if "mixed" in mode else "k_theta_gate_mfma")


def roofline_of(prof, lanes, mode=""):
    names = dict(SINGLE_KERNEL)
    names["theta_gate"] = (theta_kernel(mode), "mfma")
    dom = max(names, key=lambda n: prof[n][1])
    kernel, bound = names[dom]
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
                 "fp64_frac": tflops / FP64_PEAK_TFLOPS, "hbm_gbs": gbs, "lanes": lanes,
                 "device_ms": {n: prof[n][1] for n in prof},
                 "per_kernel": {names[n][0]: {"launches": int(prof[n][0]),
                                                      "avg_ms": prof[n][1] / max(prof[n][0], 1),
                                                      "bytes_per_launch": prof[n][3] / max(prof[n][0], 1),
                                                      "flops_per_launch": prof[n][2] / max(prof[n][0], 1)}
                                for n in names if prof[n][0] > 0}})
    return roof


def kernel_base(name):
    """Name of a kernel as rocprofv3 prints it, reduced to the library's launch classes:
    no return type, namespace or parameter list, and no template arguments except those
    of k_qrp_reflect, whose instances belong to different classes
    ("void (anonymous namespace)::k_x<2>(int, ...)" -> "k_x")."""
    n = name.strip()
    if n.startswith("void "):
        n = n[5:]
    n = n.replace("(anonymous namespace)::", "")
    n = n.split("(")[0].strip()
    base = n.split("<")[0].split("::")[-1]
    if base == "k_qrp_reflect" and "<" in n:
        return base + "<" + n.split("<", 1)[1]
    return base


def _read_csv(pattern, path):
    rows = []
    for f in sorted(glob.glob(os.path.join(path, "**", pattern), recursive=True)):
        with open(f) as fh:
            rows.extend(csv.DictReader(fh))
    return rows


def rocprof_summary(path, window_file=None, traffic_out=None, calib_dir=None, calib_mib=0):
    """Per-kernel tables from rocprofv3 CSV output under `path`.

    * *_kernel_trace.csv (--kernel-trace): per-dispatch durations.  With `window_file`
      (the JSON line of `bench.py --roofline-only`), each kernel's statistics are taken
      over its last N dispatches, N = its launches in the timed window as counted by the
      library; otherwise over all dispatches.
    * *_counter_collection.csv (--pmc FETCH_SIZE, or WRITE_SIZE; KiB per dispatch): the
      same last-N rule.  `calib_dir` holds the counter runs of `--copy-calibration
      calib_mib`: its largest FETCH_SIZE dispatch is a copy of calib_mib MiB and its
      WRITE_SIZE dispatches each write calib_mib MiB, which scales both counters to bytes.
    """
    window = {}
    if window_file:
        with open(window_file) as fh:
            line = [ln for ln in fh if ln.strip().startswith("{")][-1]
        for k, v in json.loads(line)["roofline"]["per_kernel"].items():
            window[kernel_base(k)] = int(v["launches"])
    out = []
    trace = _read_csv("*kernel_trace.csv", path)
    if trace:
        by = {}
        for r in trace:
            by.setdefault(kernel_base(r["Kernel_Name"]), []).append(
                (int(r["Start_Timestamp"]), int(r["End_Timestamp"]) - int(r["Start_Timestamp"])))
        stats = []
        for name, v in by.items():
            v.sort()
            n = window.get(name, 0) if window else 0
            sel = v[-n:] if n else ([] if window else v)
            if not sel:
                continue
            tot = sum(x[1] for x in sel)
            stats.append((name, len(sel), tot, tot / len(sel)))
        stats.sort(key=lambda s: -s[2])
        tot_all = sum(s[2] for s in stats) or 1
        scope = "the timed window (last N dispatches per kernel)" if window else "all dispatches"
        out.append(f"Kernel durations over {scope}:")
        out.append("")
        out.append("| kernel | dispatches | total ms | average us | % of listed |")
        out.append("|---|---:|---:|---:|---:|")
        for name, n, tot, avg in stats[:30]:
            out.append(f"| `{name}` | {n} | {tot / 1e6:.1f} | {avg / 1e3:.2f} | {100 * tot / tot_all:.1f} |")
        out.append("")
        if window:
            # every kernel that started inside the window, which opens at the first timed dispatch of
            # the kernel with the most timed launches (kernels the library does not count are included)
            ref = max(window, key=lambda k: window[k])
            if ref in by and window[ref] <= len(by[ref]):
                t0 = by[ref][-window[ref]][0]
                allk = []
                for name, v in by.items():
                    sel = [x for x in v if x[0] >= t0]
                    if sel:
                        tot = sum(x[1] for x in sel)
                        allk.append((name, len(sel), tot, tot / len(sel)))
                allk.sort(key=lambda s: -s[2])
                tot_w = sum(s[2] for s in allk) or 1
                out.append(f"Every kernel dispatched from the first timed `{ref}` on (the timed steps and the "
                           "run's closing measurements):")
                out.append("")
                out.append("| kernel | dispatches | total ms | average us | % |")
                out.append("|---|---:|---:|---:|---:|")
                for name, n, tot, avg in allk[:30]:
                    out.append(f"| `{name}` | {n} | {tot / 1e6:.1f} | {avg / 1e3:.2f} | {100 * tot / tot_w:.1f} |")
                out.append("")
    cnt = _read_csv("*counter_collection.csv", path)
    kern = {}
    if cnt:
        scale = {"FETCH_SIZE": 1.0, "WRITE_SIZE": 1.0}
        if calib_dir and calib_mib:
            nbytes = calib_mib * (1 << 20)
            cal = _read_csv("*counter_collection.csv", calib_dir)
            f = [float(r["Counter_Value"]) for r in cal if r["Counter_Name"] == "FETCH_SIZE"]
            w = sorted(float(r["Counter_Value"]) for r in cal if r["Counter_Name"] == "WRITE_SIZE")
            if f:
                scale["FETCH_SIZE"] = nbytes / (max(f) * 1024.0)
            if w:
                scale["WRITE_SIZE"] = nbytes / (w[len(w) // 2] * 1024.0)
        by = {}
        for r in cnt:
            key = (kernel_base(r["Kernel_Name"]), r["Counter_Name"])
            by.setdefault(key, []).append((int(r["Dispatch_Id"]), float(r["Counter_Value"])))
        out.append("HBM counters per dispatch (KiB as read, bytes after scaling by the copy calibration"
                   f" {scale}):")
        out.append("")
        out.append("| kernel | counter | dispatches | mean KiB | bytes per dispatch |")
        out.append("|---|---|---:|---:|---:|")
        for (name, cn), v in sorted(by.items()):
            v.sort()
            n = window.get(name, 0) if window else 0
            sel = v[-n:] if n else ([] if window else v)
            if not sel or cn not in scale:
                continue
            mean = sum(x[1] for x in sel) / len(sel)
            b = mean * 1024.0 * scale[cn]
            k = kern.setdefault(name, {"dispatches": len(sel)})
            k[cn + "_bytes"] = b
            out.append(f"| `{name}` | {cn} | {len(sel)} | {mean:.1f} | {b:.4g} |")
        for k in kern.values():
            k["hbm_bytes_per_launch"] = k.get("FETCH_SIZE_bytes", 0.0) + k.get("WRITE_SIZE_bytes", 0.0)
        out.append("")
        if traffic_out:
            old, old_scale = {}, {}
            if os.path.exists(traffic_out):
                with open(traffic_out) as fh:
                    prev = json.load(fh)
                old, old_scale = prev.get("kernels", {}), prev.get("scale", {})
            seen = {cn for (_, cn) in by}
            old_scale.update({cn: v for cn, v in scale.items() if cn in seen})
            for name, k in kern.items():  # one counter per pass: merge with the other pass
                o = old.get(name, {})
                o.update(k)
                o["hbm_bytes_per_launch"] = o.get("FETCH_SIZE_bytes", 0.0) + o.get("WRITE_SIZE_bytes", 0.0)
                old[name] = o
            with open(traffic_out, "w") as fh:
                json.dump({"workload": WORKLOAD_TAG, "scale": old_scale, "kernels": old}, fh, indent=1)
    print("\n".join(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed steps, all at chi=256 in the bulk")
    ap.add_argument("--warmup", type=int, default=5, help="untimed steps after the bulk has saturated")
    ap.add_argument("--compare-steps", type=int, default=3)
    ap.add_argument("--mode", default="batched,learned,split")
    ap.add_argument("--lanes", type=int, default=3, help="streams / host threads per half-sweep of the timed run")
    ap.add_argument("--graphs", type=int, default=1,
                    help="hipGraph replay of the Jacobi sweeps in every run (0 = eager launches)")
    ap.add_argument("--roofline-lanes", type=int, default=1, help="lanes of the per-kernel event run")
    ap.add_argument("--roofline-only", action="store_true", help="only the per-kernel event run")
    ap.add_argument("--reference-mode", default="batched", help="exact GPU path of the comparison ('' = skip)")
    ap.add_argument("--mixed-mode", default="batched,learned,mixed",
                    help="FP32-theta path timed and checked ('' = skip)")
    ap.add_argument("--fp64-theta-mode", default="batched,learned",
                    help="learned-rank path with FP64 theta, timed and checked ('' = skip)")
    ap.add_argument("--rank-model", default=RANK_MODEL if os.path.exists(RANK_MODEL) else "",
                    help="rank model parameters (.npy) for the learned-rank path ('' = none)")
    ap.add_argument("--compare", default="fp64", help="per-bond FP64 drop-in timed on the window ('' = skip)")
    ap.add_argument("--oracle-check", type=int, default=1, help="bond updates of every path vs the CPU oracle")
    ap.add_argument("--cpu-sample-bonds", type=int, default=3)
    ap.add_argument("--nc-steps", type=int, default=10,
                    help="timed steps of the number-conserving form of the workload (0 = skip)")
    ap.add_argument("--nc-only", action="store_true", help="only the number-conserving run")
    ap.add_argument("--cpu-threads", type=int, default=0, help="threads of the all-cores CPU figure (0 = auto)")
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only time a device-to-device copy of MIB MiB (HBM counter calibration)")
    ap.add_argument("--rocprof-summary", default="", metavar="DIR", help="summarise rocprofv3 CSV output and exit")
    ap.add_argument("--window", default="", help="with --rocprof-summary: output of --roofline-only (timed window)")
    ap.add_argument("--traffic-out", default="", help="with --rocprof-summary: merge per-kernel HBM bytes into this")
    ap.add_argument("--calib-dir", default="", help="with --rocprof-summary: counter runs of --copy-calibration")
    ap.add_argument("--calib-mib", type=int, default=0, help="with --rocprof-summary: size of the calibration copy")
    args = ap.parse_args()
    log = (lambda s: None) if args.quiet else (lambda s: print(s, file=sys.stderr, flush=True))

    if args.rocprof_summary:
        rocprof_summary(args.rocprof_summary, args.window or None, args.traffic_out or None, args.calib_dir or None,
                        args.calib_mib)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    if args.nc_only:
        print(json.dumps({"number_conserving": run_nc(args.nc_steps, args.warmup, log, bool(args.graphs))}), flush=True)
        return
    model = np.load(args.rank_model) if args.rank_model else None
    graphs = bool(args.graphs)
    if args.roofline_only:
        roof_run = run_mode(args.mode, args.steps, args.warmup, log, rank_model=model, lanes=args.roofline_lanes,
                            tag="roofline", profile=True, graphs=graphs)
        print(json.dumps({"roofline": roofline_of(roof_run["prof"], args.roofline_lanes, args.mode),
                          "ms_per_step": 1e3 * roof_run["elapsed"] / args.steps, "steps": args.steps,
                          "pre_steps": roof_run["n_pre"]}), flush=True)
        return

    res = run_mode(args.mode, args.steps, args.warmup, log, rank_model=model, lanes=args.lanes, graphs=graphs,
                   keep_state=bool(args.oracle_check))
    n_pre, n_sat = res["n_pre"], res["n_sat"]
    sps = args.steps / res["elapsed"]
    extra = {}
    ref = None
    if args.reference_mode:
        ref = run_mode(args.reference_mode, args.steps, args.warmup, log, pre_steps=n_sat,
                       keep_state=bool(args.compare), lanes=args.lanes, graphs=graphs)
        extra[f"{args.reference_mode}_steps_per_s"] = args.steps / ref["elapsed"]
    f64t = None
    if args.fp64_theta_mode:
        f64t = run_mode(args.fp64_theta_mode, args.steps, args.warmup, log, pre_steps=n_sat, rank_model=model,
                        lanes=args.lanes, graphs=graphs)
    mixed = None
    if args.mixed_mode:
        mixed = run_mode(args.mixed_mode, args.steps, args.warmup, log, pre_steps=n_sat, rank_model=model,
                         lanes=args.lanes, graphs=graphs)
    roof_run = run_mode(args.mode, args.steps, args.warmup, log, pre_steps=n_sat, rank_model=model,
                        lanes=args.roofline_lanes, tag="roofline", profile=True, graphs=graphs)
    roof = roofline_of(roof_run["prof"], args.roofline_lanes, args.mode)

    def obs_errors(run):
        errs = {k: float(np.max(np.abs(np.asarray(run["obs"][k]) - np.asarray(ref["obs"][k])))) for k in run["obs"]}
        errs["trunc"] = abs(run["trunc"] - ref["trunc"])
        return errs

    if ref is not None:
        errs = obs_errors(res)
        extra["obs_err_vs_exact_gpu"] = {"max": max(errs.values()), "by_observable": errs,
                                         "against": f"{args.reference_mode} (exact FP64 GPU path), "
                                                    f"after the same {n_pre + args.steps} steps"}
    if f64t is not None:
        extra["fp64_theta"] = {"mode": args.fp64_theta_mode, "steps_per_s": args.steps / f64t["elapsed"]}
        if ref is not None:
            ferr = obs_errors(f64t)
            extra["fp64_theta"]["obs_err_vs_exact_gpu"] = {"max": max(ferr.values()), "by_observable": ferr}
    if mixed is not None:
        extra["fp32_theta"] = {"mode": args.mixed_mode, "steps_per_s": args.steps / mixed["elapsed"]}
        if ref is not None:
            merr = obs_errors(mixed)
            extra["fp32_theta"]["obs_err_vs_exact_gpu"] = {"max": max(merr.values()), "by_observable": merr}
    if args.compare and ref is not None:
        cmp_res = run_mode(args.compare, args.compare_steps, 0, log, state=ref["snap"], graphs=graphs, lanes=1)
        extra[f"{args.compare}_steps_per_s"] = args.compare_steps / cmp_res["elapsed"]
        extra[f"{args.compare}_window"] = f"steps {n_pre + 1}-{n_pre + args.compare_steps}"
        extra[f"speedup_vs_{args.compare}"] = sps / extra[f"{args.compare}_steps_per_s"]
    if args.oracle_check:
        modes = [args.mode] + [m for m in (args.fp64_theta_mode, args.reference_mode, args.mixed_mode, args.compare)
                               if m]
        extra["obs_err_vs_oracle"] = oracle_check(res["snap"], modes, log, rank_model=model)
        extra["obs_err_vs_oracle"]["bonds"] = list(ORACLE_BONDS)

    if args.nc_steps > 0:
        extra["number_conserving"] = run_nc(args.nc_steps, args.warmup, log, graphs)

    cpu = None
    if args.cpu_sample_bonds > 0:
        model_name, avail = host_cpu()
        threads = args.cpu_threads or max(1, min(16, avail))
        cb = cpu_baseline(args.cpu_sample_bonds, threads, log)
        cpu = cb["one"]
        if "all" in cb:
            extra["cpu_baseline_all_cores"] = cb["all"]

    st = res["stats"]
    half_steps = 3 * args.steps
    out = {
        "metric": "Trotter steps/s, L=128 \u03c7=256: mixed+batched vs FP64 GPU at \u22641e-8 obs err",
        "value": sps,
        "unit": "Trotter steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * res["elapsed"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "i8 slices -> i32 (theta contraction), complex128 f64 (gate, SVD, lambda)" if "split" in args.mode
                 else "complex128 (f64)",
        "data": "synthetic (Neel product state, Fermi-Hubbard gates)",
        "config": {"workload": f"Fermi-Hubbard quench L={L} d={D} t=1 U=4 dt={DT} chi_max={CHI}, Neel start; "
                               f"timed steps {n_pre + 1}-{n_pre + args.steps} with every bulk bond at chi={CHI}",
                   "mode": args.mode, "lanes": args.lanes, "graphs": graphs, "gemm_autotune": res["tuning"],
                   "saturation_steps": n_sat, "chi_max_reached": int(res["dims"].max()),
                   "bulk_chi_min": int(res["dims"][BULK].min()),
                   "mean_bond_dim": float(res["dims"].mean()), "warmup_s": res["warm_s"]},
        "roofline": roof,
        "learned_rank": {"rank_model": os.path.basename(args.rank_model) if args.rank_model else None,
                         "bonds_sent_to_full_svd_by_model": st["model_full_svd"],
                         "sketch_accepted": st["sketch_accepted"], "full_svd_fallbacks": st["fallbacks"],
                         "mean_sketch_width": st["sketch_cols"] / max(st["sketch_accepted"] + st["fallbacks"], 1)},
        "jacobi": {"sweeps_per_half_step_per_lane": st["jacobi_sweeps"] / (half_steps * max(args.lanes, 1)),
                   "pair_visits": st["pair_visits"], "grams_formed": st["pair_grams"],
                   "pairs_rotated": st["pair_rotations"],
                   "rotations_by_sweep": [int(x) for x in np.trim_zeros(st["sweep_rotations"], "b")],
                   "solves_by_sweep_count": [int(x) for x in np.trim_zeros(st["solve_sweeps"], "b")]},
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    try:
        main()
    except BenchError as e:
        print(f"bench.py: {e}", file=sys.stderr)
        sys.exit(3)
