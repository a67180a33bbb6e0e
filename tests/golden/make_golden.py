"""Write the golden fixtures in tests/golden/ from the reference itself.

Runs oracle/_ref/golden_driver (the reference's Fortran modules compiled from
/root/reference by ``make -C oracle ref``, plus a project-written driver) on
seeded jobs and stores its outputs as .npz data: bond Hamiltonians, Trotter
gates, per-step observables and final states.  Inputs that need randomness
(fields, hoppings, product states, random Vidal tensors) come from numpy with
fixed seeds and are written to the job file, so both sides see the same bytes.

Usage:  make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "golden_driver")


def fmt(x):
    return " ".join(repr(float(v)) for v in np.ravel(x))


def write_job(path, model, L, param, chi, consts, field, dt, nsteps, canon, coeffs, mode=0, extra=None):
    with open(path, "w") as f:
        f.write(f"{model} {L} {param} {chi}\n")
        f.write(fmt(consts) + "\n")
        if field is not None and len(field):
            f.write(fmt(field) + "\n")
        f.write(f"{float(dt.real)!r} {float(dt.imag)!r} {nsteps} {canon}\n")
        for s in range(L):
            f.write(" ".join(f"{float(c.real)!r} {float(c.imag)!r}" for c in coeffs[s]) + "\n")
        f.write(f"{mode}\n")
        if extra:
            f.write(extra)


# progress lines the reference's Swapping prints on every call (local_operations_module.f90:680-693)
SWAP_CHATTER = ("chis", "theta", "SVD", "L1", "G1", "G2")


def parse(text):
    blocks = []
    for line in text.splitlines():
        tok = line.split()
        if not tok or tok[0] in SWAP_CHATTER:
            continue
        if tok[0][0].isalpha():
            head = tok[0]
            nums = []
            rest = tok[1:]
            blocks.append((head, nums))
            # integer tag arguments (indices) for H/U/LAMBDA/GAMMA/AOP/DIMS
            if head in ("H", "U", "LAMBDA", "GAMMA", "AOP", "DIMS", "MTB", "MCORR", "MFCORR", "LABELL", "LABELR"):
                blocks[-1] = (head, nums, [int(t) for t in rest])
                continue
            nums.extend(float(t) for t in rest)
            blocks[-1] = (head, nums, [])
        else:
            blocks[-1][1].extend(float(t) for t in tok)
    return blocks


def run(job):
    out = subprocess.run([DRIVER, job], check=True, capture_output=True, text=True)
    return parse(out.stdout)


def run_itp(job):
    """A mode-2 job (ImagTimeProp): the reference's own convergence printout -- one
    "ITP step j ..." and one "Percent difference t convergence Criterion c" line per
    judgement -- split into stages where j restarts, plus the usual blocks."""
    import re

    out = subprocess.run([DRIVER, job], check=True, capture_output=True, text=True).stdout
    num = re.compile(r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eEdD][-+]?\d+)?")
    judged, rest = [], []
    in_judgement = False  # list-directed output wraps the "ITP step j" line at varying places
    for line in out.splitlines():
        t = line.strip()
        if t.startswith("ITP step j"):
            judged.append([int(num.findall(t)[0])])
            in_judgement = True
        elif t.startswith("Percent difference"):
            v = [float(x.replace("D", "E").replace("d", "e")) for x in num.findall(t)]
            judged[-1] += v[:2]
            in_judgement = False
        elif in_judgement or t.startswith("number") or t.startswith("Sx"):
            continue  # the wrapped tail of a judgement line, and the printed totals
        elif re.fullmatch(r"\d+", t):
            continue  # ImagTimePropNC prints each stage's chi (propagation_module.f90:2031)
        else:
            rest.append(line)
    stages, cur = [], []
    for j in judged:
        if cur and j[0] <= cur[-1][0]:
            stages.append(cur)
            cur = []
        cur.append(j)
    if cur:
        stages.append(cur)
    return stages, parse("\n".join(rest))


def collect(blocks, L, d, chi):
    d2 = d * d
    res = {"H": [], "U": [], "energy": [], "trunc": [], "sz": [], "n": [], "S": [],
           "lam": [], "gam": [], "aop": [], "labl": [], "labr": []}
    for head, nums, args in blocks:
        a = np.array(nums)
        if head in ("H", "U"):
            m = a.reshape(d2, 2 * d2)
            res[head].append(m[:, 0::2] + 1j * m[:, 1::2])
        elif head == "SZOP":
            res["szop"] = a.reshape(d, d)
        elif head == "NOP":
            res["nop"] = a.reshape(d, d)
        elif head == "AOP":
            res["aop"].append(a.reshape(d, d))
        elif head == "STEP":  # step index, energy, truncerr of that step
            res["energy"].append(a[1])
            res["trunc"].append(a[2])
        elif head == "SZ":
            res["sz"].append(a)
        elif head == "N":
            res["n"].append(a)
        elif head == "S":
            res["S"].append(a)
        elif head == "TRUNC":
            res["trunc"].append(a[0])
        elif head == "LAMBDA":
            res["lam"].append(a)
        elif head == "LABELL":
            res["labl"].append(a.astype(np.int64))
        elif head == "LABELR":
            res["labr"].append(a.astype(np.int64))
        elif head == "CONSERV":
            res["conserv"] = a.astype(np.int64)
        elif head == "GAMMA":
            z = a[0::2] + 1j * a[1::2]
            # Fortran (a, i, b) column-major -> [b, i, a]
            res["gam"].append(z.reshape(chi, d, chi))
    out = {}
    for k, v in res.items():
        if isinstance(v, list):
            if len(v):
                out[k] = np.array(v)
        else:
            out[k] = v
    return out


def neel(L, d, up, dn):
    c = np.zeros((L, d), np.complex128)
    for s in range(L):
        c[s, up if s % 2 == 0 else dn] = 1.0
    return c


def main():
    if not os.path.exists(DRIVER):
        sys.exit("build the reference driver first: make -C oracle ref")
    tmp = tempfile.mkdtemp()
    jobs = {}

    # XXZ spin-1/2 (the CPU-runnable workload): L=32, Jx=Jy=Jz=1, h_i~U[-1/2,1/2]
    # seed 1, Neel, chi=32, dt=0.01, 100 steps.
    L = 32
    h = np.random.default_rng(1).uniform(-0.5, 0.5, L)
    jobs["xxz_l32_chi32"] = dict(model=1, L=L, param=1, chi=32, consts=[1.0, 1.0, 1.0], field=h,
                                 dt=0.01 + 0j, nsteps=100, canon=0, coeffs=neel(L, 2, 0, 1), d=2)
    # Same chain at chi=8: truncation active from the first steps.
    jobs["xxz_l16_chi8"] = dict(model=1, L=16, param=1, chi=8, consts=[1.0, 1.0, 0.5], field=h[:16],
                                dt=0.01 + 0j, nsteps=60, canon=0, coeffs=neel(16, 2, 0, 1), d=2)
    # Bose-Hubbard JVector: d=5 (maxFilling 4), U=1, J_i~U[0.05,0.15] seed 2, Mott n=1.
    L = 12
    J = np.random.default_rng(2).uniform(0.05, 0.15, L - 1)
    c = np.zeros((L, 5), np.complex128)
    c[:, 1] = 1.0
    jobs["bh_l12_chi24"] = dict(model=2, L=L, param=4, chi=24, consts=[1.0, 0.0, 0.0], field=J,
                                dt=0.005 + 0j, nsteps=40, canon=0, coeffs=c, d=5)
    # Fermi-Hubbard spin-1/2, maxFilling 2 (d=4), t=1, U=4, Neel (up=index 2, down=index 3).
    jobs["fh_l10_chi32"] = dict(model=3, L=10, param=2, chi=32, consts=[1.0, 4.0, 0.0], field=None,
                                dt=0.01 + 0j, nsteps=30, canon=0, coeffs=neel(10, 4, 1, 2), d=4)
    # Truncated FH run: a random complex product start (seed 3) keeps the Schmidt
    # spectrum non-degenerate at the chi=8 cut (a Neel start truncates through
    # exactly degenerate multiplets, where the kept vectors are not unique).
    r = np.random.default_rng(3)
    c = r.normal(size=(8, 4)) + 1j * r.normal(size=(8, 4))
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    jobs["fh_l8_chi8"] = dict(model=3, L=8, param=2, chi=8, consts=[1.0, 4.0, 0.0], field=None,
                              dt=0.02 + 0j, nsteps=25, canon=0, coeffs=c, d=4)
    # Spin-1 Heisenberg, random normalised complex product state seed 4.
    L = 8
    r = np.random.default_rng(4)
    c = r.normal(size=(L, 3)) + 1j * r.normal(size=(L, 3))
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    jobs["s1_l8_chi12"] = dict(model=1, L=L, param=2, chi=12, consts=[1.0, 1.0, 1.0], field=np.zeros(L),
                               dt=0.01 + 0j, nsteps=30, canon=0, coeffs=c, d=3)
    # Imaginary time (dt = -i dtau), CanonicalFormAll after every step;
    # spin-1/2 Heisenberg with h_i~U[-1,1] seed 5, equal-weight product start (AllStates).
    L = 10
    hf = np.random.default_rng(5).uniform(-1.0, 1.0, L)
    c = np.full((L, 2), 1.0 / np.sqrt(2.0), np.complex128)
    jobs["itp_l10_chi8"] = dict(model=1, L=L, param=1, chi=8, consts=[1.0, 1.0, 1.0], field=hf,
                                dt=-0.05j, nsteps=40, canon=1, coeffs=c, d=2)

    for name, j in jobs.items():
        path = os.path.join(tmp, name + ".txt")
        write_job(path, j["model"], j["L"], j["param"], j["chi"], j["consts"], j["field"], j["dt"],
                  j["nsteps"], j["canon"], j["coeffs"])
        res = collect(run(path), j["L"], j["d"], j["chi"])
        meta = dict(model=j["model"], L=j["L"], param=j["param"], chi=j["chi"], d=j["d"],
                    consts=np.array(j["consts"], float), dt=np.complex128(j["dt"]), nsteps=j["nsteps"],
                    canon=j["canon"], coeffs=j["coeffs"],
                    field=np.array([] if j["field"] is None else j["field"], float))
        if j["L"] * j["chi"] * j["d"] * j["chi"] > 4096:  # keep fixtures small
            res.pop("gam", None)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **meta, **res)
        print(name, "energy[-1]=", res["energy"][-1], "sum trunc=", np.sum(res["trunc"]))

    # Single TwoSiteOp on a random, non-canonical state (mode 1): L=4, d=2, chi=4, link 2.
    r = np.random.default_rng(7)
    L, d, chi, link = 4, 2, 4, 2
    gam = r.normal(size=(L, chi, d, chi)) + 1j * r.normal(size=(L, chi, d, chi))
    lam = np.sort(r.uniform(0.1, 1.0, size=(L + 1, chi)), axis=1)[:, ::-1].copy()
    lam /= np.linalg.norm(lam, axis=1, keepdims=True)
    extra = f"{link}\n"
    for s in range(L):
        z = gam[s].reshape(-1)  # [b, i, a] C-order == Fortran (a,i,b)
        extra += " ".join(f"{float(v.real)!r} {float(v.imag)!r}" for v in z) + "\n"
    for l in range(L + 1):
        extra += fmt(lam[l]) + "\n"
    hz = r.uniform(-0.5, 0.5, L)
    path = os.path.join(tmp, "tso.txt")
    write_job(path, 1, L, 1, chi, [1.0, 0.7, 1.3], hz, 0.05 + 0j, 0, 0, neel(L, 2, 0, 1), mode=1, extra=extra)
    res = collect(run(path), L, d, chi)
    np.savez_compressed(os.path.join(HERE, "two_site_op_l4.npz"), gam_in=gam, lam_in=lam, link=link,
                        field=hz, consts=np.array([1.0, 0.7, 1.3]), dt=np.complex128(0.05), L=L, d=d, chi=chi,
                        **res)
    print("two_site_op_l4 trunc=", res["trunc"])


def itp_stages():
    """ImagTimeProp with chi stages chiIn=4 -> chiMax=8 (propagation_module.f90:636-760):
    spin-1/2 Heisenberg L=10, h_i~U[-1,1] seed 5, equal-weight product start, dtITP=0.2,
    stepsForJudge=10, convCriterion1=1e-4, convCriterion2=1e-5, maxITPsteps=600.  Stored: for every stage
    the judgements (j, largest relative change of lambda_1, criterion), and the final
    state's observables, lambdas and Gammas."""
    L, chi, chi_in = 10, 8, 4
    hf = np.random.default_rng(5).uniform(-1.0, 1.0, L)
    c = np.full((L, 2), 1.0 / np.sqrt(2.0), np.complex128)
    params = dict(chi_in=chi_in, dtau=0.2, max_steps=600, judge=10, conv1=1e-4, conv2=1e-5)
    tmp = tempfile.mkdtemp()
    path = os.path.join(tmp, "itp_stages.txt")
    extra = f"{chi_in} {params['dtau']!r} {params['max_steps']} {params['judge']} {params['conv1']!r} {params['conv2']!r}\n"
    write_job(path, 1, L, 1, chi, [1.0, 1.0, 1.0], hf, 0j, 0, 0, c, mode=2, extra=extra)
    stages, blocks = run_itp(path)
    res = collect(blocks, L, 2, chi)
    steps = [s[-1][0] for s in stages]
    conv = [bool(s[-1][1] < s[-1][2]) for s in stages]
    # the last judgement of each stage and the one before must not sit at the criterion (the
    # device's lambdas agree with the reference's to ~1e-13, far inside this margin)
    for s in stages:
        for j, t, cr in s[-2:]:
            assert abs(t / cr - 1.0) > 1e-4, (j, t, cr)
    meta = dict(model=1, L=L, param=1, chi=chi, d=2, consts=np.ones(3), field=hf, coeffs=c,
                stage_steps=np.array(steps), stage_converged=np.array(conv),
                judgements=np.array([[k, j, t, cr] for k, s in enumerate(stages) for j, t, cr in s]),
                **{k: np.array(v) for k, v in params.items()})
    np.savez_compressed(os.path.join(HERE, "itp_stages_l10.npz"), **meta, **res)
    print("itp_stages_l10 stages", steps, conv, "energy", res["energy"][-1])


def schedules():
    """Trajectories under the reference's other Trotter schedules (mode 3): 5th-order
    Forest-Ruth with open boundaries, and 2nd / 5th order with periodic boundaries (edge
    gate after swapping site L next to site 1), for even and odd L.  Spin-1/2 XXZ with
    h_i~U[-1/2,1/2] (seed 21) from a random product state (seed 22); chi=16 is full rank
    at L=8, chi=8 truncates at L=9."""
    tmp = tempfile.mkdtemp()
    cases = {"sched_o5_obc_l10": (10, 16, 5, "O", 20), "sched_o2_pbc_l8": (8, 16, 2, "P", 15),
             "sched_o2_pbc_l9": (9, 8, 2, "P", 15), "sched_o5_pbc_l8": (8, 16, 5, "P", 10),
             "sched_o5_pbc_l9": (9, 8, 5, "P", 10)}
    for name, (L, chi, order, bc, nsteps) in cases.items():
        h = np.random.default_rng(21).uniform(-0.5, 0.5, L)
        r = np.random.default_rng(22)
        c = r.normal(size=(L, 2)) + 1j * r.normal(size=(L, 2))
        c /= np.linalg.norm(c, axis=1, keepdims=True)
        path = os.path.join(tmp, name + ".txt")
        write_job(path, 1, L, 1, chi, [1.0, 1.0, 0.7], h, 0.05 + 0j, nsteps, 0, c, mode=3, extra=f"{order} {bc}\n")
        res = collect(run(path), L, 2, chi)
        meta = dict(model=1, L=L, param=1, chi=chi, d=2, consts=np.array([1.0, 1.0, 0.7]), dt=np.complex128(0.05),
                    nsteps=nsteps, canon=0, coeffs=c, field=h, order=order, bc=bc)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **meta, **res)
        print(name, "energy[-1]=", res["energy"][-1], "sum trunc=", np.sum(res["trunc"]), "gates", len(res["U"]))


def measures():
    """EvaluateMeasures (observables_module.f90:2286-2526) after a trajectory (mode 5):
    Fermi-Hubbard L=6 chi=16 (with the fermi-phase correlator c+_up x c_up), spin-1/2 XXZ
    L=8 chi=8 (truncated), and a periodic XXZ chain L=7 (energy with the edge bond), each
    from a random product state."""
    tmp = tempfile.mkdtemp()
    cases = {"meas_fh_l6": (3, 6, 2, 16, [1.0, 4.0, 0.0], None, 4, 0.05, 8, 2, "O"),
             "meas_xxz_l8": (1, 8, 1, 8, [1.0, 1.0, 0.6], 2, 2, 0.05, 12, 2, "O"),
             "meas_xxz_pbc_l7": (1, 7, 1, 8, [1.0, 1.0, 0.6], 2, 2, 0.05, 8, 2, "P")}
    for name, (model, L, param, chi, consts, nf, d, dt, nsteps, order, bc) in cases.items():
        r = np.random.default_rng(51)
        c = r.normal(size=(L, d)) + 1j * r.normal(size=(L, d))
        c /= np.linalg.norm(c, axis=1, keepdims=True)
        field = np.random.default_rng(52).uniform(-0.5, 0.5, L) if model == 1 else None
        path = os.path.join(tmp, name + ".txt")
        write_job(path, model, L, param, chi, consts, field, dt + 0j, nsteps, 0, c, mode=5, extra=f"{order} {bc}\n")
        blocks = run(path)
        res = collect(blocks, L, d, chi)
        meas = {}
        for head, nums, args in blocks:
            a = np.array(nums)
            if head == "MEN":
                meas["m_en"] = a[0]
            elif head == "MQME":
                meas["m_qme"] = a[0]
            elif head == "MLOCAL":
                meas["m_local"] = a[0::2] + 1j * a[1::2]
            elif head == "MAVG":
                meas["m_avg"] = a[0] + 1j * a[1]
            elif head == "MVN":
                meas["m_vn"] = a
            elif head == "MCHAIN":
                meas["m_chain"] = a
            elif head in ("MTB", "MCORR", "MFCORR"):
                key = {"MTB": "m_tbvn", "MCORR": "m_corr", "MFCORR": "m_fcorr"}[head]
                row = a if head == "MTB" else a[0::2] + 1j * a[1::2]
                meas.setdefault(key, []).append(row)
        meas = {k: np.array(v) for k, v in meas.items()}
        meta = dict(model=model, L=L, param=param, chi=chi, d=d, consts=np.array(consts), dt=np.complex128(dt),
                    nsteps=nsteps, canon=0, coeffs=c, field=np.array([] if field is None else field), order=order,
                    bc=bc)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **meta, **res, **meas)
        print(name, {k: v.shape for k, v in meas.items()})


def checkpoint():
    """The reference's own checkpoint files (io_module RecordLambdas / RecordGammas, kinds
    'S' and 'B', mode 4) of a spin-1/2 XXZ state, L=4, chi=4 (bond 2 truncated from 4 to
    4 -- full rank -- and padded zeros at the edges), after 6 steps from a random product
    state (seed 31); kept as the files themselves in tests/golden/ckpt_l4/ beside the state."""
    import shutil

    L, chi = 4, 4
    r = np.random.default_rng(31)
    c = r.normal(size=(L, 2)) + 1j * r.normal(size=(L, 2))
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    h = np.random.default_rng(32).uniform(-0.5, 0.5, L)
    tmp = tempfile.mkdtemp()
    path = os.path.join(tmp, "ckpt.txt")
    write_job(path, 1, L, 1, chi, [1.0, 1.0, 0.5], h, 0.05 + 0j, 6, 0, c, mode=4)
    out = subprocess.run([DRIVER, path], check=True, capture_output=True, text=True, cwd=tmp).stdout
    res = collect(parse(out), L, 2, chi)
    dest = os.path.join(HERE, "ckpt_l4")
    os.makedirs(dest, exist_ok=True)
    for f in ("ckpt_lambdas_S.dat", "ckpt_gammas_S.dat", "ckpt_lambdas_B.dat", "ckpt_gammas_B.dat"):
        shutil.copy(os.path.join(tmp, f), os.path.join(dest, f))
    np.savez_compressed(os.path.join(dest, "state.npz"), L=L, d=2, chi=chi, gam=res["gam"][0] if res["gam"].ndim == 5
                        else res["gam"], lam=res["lam"], H=res["H"], U=res["U"], coeffs=c)
    print("checkpoint files written", os.listdir(dest))


def nc():
    """Number-conserving trajectories (mode 6: ProductStateLabels + TrotterStepNC):
    Bose-Hubbard d=3 (maxFilling 2) from the unit-filling Mott state, J_i~U[0.8,1.2]
    (seed 31), U=2, at L=8 chi=16 (truncated) and L=6 chi=32 (full rank); Fermi-Hubbard
    d=4 (t=1, U=4, number = the model's Conserv, intDegFree) from the Neel state at L=6
    chi=64 (full rank: its spin-symmetric spectra have degenerate pairs, and a cut through
    one leaves the kept state to rounding)."""
    tmp = tempfile.mkdtemp()
    cases = {
        "nc_bh_l8": dict(model=2, L=8, param=2, chi=16, consts=[2.0, 0.0, 0.0], up=1, dn=1, nsteps=20, intdeg=0),
        "nc_bh_l6_full": dict(model=2, L=6, param=2, chi=32, consts=[2.0, 0.0, 0.0], up=1, dn=1, nsteps=15, intdeg=0),
        "nc_fh_l6": dict(model=3, L=6, param=2, chi=64, consts=[1.0, 4.0, 0.0], up=1, dn=2, nsteps=12, intdeg=1),
    }
    for name, cs in cases.items():
        L, chi, model = cs["L"], cs["chi"], cs["model"]
        d = 3 if model == 2 else 4
        field = np.random.default_rng(31).uniform(0.8, 1.2, L - 1) if model == 2 else None
        c = neel(L, d, cs["up"], cs["dn"])
        tot = int(sum(np.argmax(np.abs(c[s])) for s in range(L))) if model == 2 else L
        path = os.path.join(tmp, name + ".txt")
        write_job(path, model, L, cs["param"], chi, cs["consts"], field, 0.05 + 0j, cs["nsteps"], 0, c, mode=6,
                  extra=f"{tot} {cs['intdeg']}\n")
        res = collect(run(path), L, d, chi)
        meta = dict(model=model, L=L, param=cs["param"], chi=chi, d=d, consts=np.array(cs["consts"]),
                    dt=np.complex128(0.05), nsteps=cs["nsteps"], canon=0, coeffs=c, tot_num=tot, intdeg=cs["intdeg"])
        if field is not None:
            meta["field"] = field
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **meta, **res)
        print(name, "energy[-1]=", res["energy"][-1], "sum trunc=", np.sum(res["trunc"]),
              "kept", [(l > 0).sum() for l in res["lam"]])


def nc_sched():
    """Number-conserving trajectories under the other Trotter schedules (mode 8):
    TrotterStep5thOrderNC on an open chain, TrotterStep2ndOrderPBCNC and
    TrotterStep5thOrderPBCNC (SwappingNC moves site L next to site 1 and back) on periodic
    chains of even and odd L.  Bose-Hubbard d=3 (maxFilling 2), J_i~U[0.8,1.2] (seed 31;
    the edge J_L the next draw), U=2, from the unit-filling Mott state; chi=12 truncates."""
    tmp = tempfile.mkdtemp()
    cases = {"nc_sched_o5_obc_l8": (8, 12, 5, "O", 10), "nc_sched_o2_pbc_l8": (8, 12, 2, "P", 12),
             "nc_sched_o2_pbc_l7": (7, 12, 2, "P", 12), "nc_sched_o5_pbc_l8": (8, 12, 5, "P", 6),
             "nc_sched_o5_pbc_l7": (7, 12, 5, "P", 6)}
    for name, (L, chi, order, bc, nsteps) in cases.items():
        d = 3
        field = np.random.default_rng(31).uniform(0.8, 1.2, L)
        c = neel(L, d, 1, 1)
        tot = L
        extra = f"{order} {bc}\n" + (f"{float(field[L - 1])!r}\n" if bc == "P" else "") + f"{tot} 0\n"
        path = os.path.join(tmp, name + ".txt")
        write_job(path, 2, L, 2, chi, [2.0, 0.0, 0.0], field[:L - 1], 0.05 + 0j, nsteps, 0, c, mode=8, extra=extra)
        res = collect(run(path), L, d, chi)
        meta = dict(model=2, L=L, param=2, chi=chi, d=d, consts=np.array([2.0, 0.0, 0.0]), dt=np.complex128(0.05),
                    nsteps=nsteps, canon=0, coeffs=c, tot_num=tot, intdeg=0, order=order, bc=bc,
                    field=field if bc == "P" else field[:L - 1])
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **meta, **res)
        print(name, "energy[-1]=", res["energy"][-1], "sum trunc=", np.sum(res["trunc"]), "gates", len(res["U"]),
              "kept", [(l > 0).sum() for l in res["lam"]])


def itp_nc():
    """Number-conserving imaginary time (mode 7: ProductStateLabels + ImagTimePropNC,
    propagation_module.f90:1986-2121, CanonicalFormAllNC :1943-1983 every step):
    Bose-Hubbard d=3 (maxFilling 2), J_i~U[0.8,1.2] (seed 31), U=2, from the unit-filling
    Mott state at L=8 with chi stages 4 -> 12; Fermi-Hubbard d=4 (t=1, U=4, the model's
    Conserv, intDegFree) from the Neel state at L=6 in one full-rank stage (chi=64: its
    spin-degenerate spectra leave a truncated cut ambiguous).  Stored as itp_stages: the
    judgements per stage, the final state, and every link's labels."""
    tmp = tempfile.mkdtemp()
    cases = {
        "itp_nc_bh_l8": dict(model=2, L=8, param=2, chi=12, chi_in=4, consts=[2.0, 0.0, 0.0], up=1, dn=1, intdeg=0,
                             dtau=0.1, max_steps=500, judge=10, conv1=1e-4, conv2=1e-6),
        "itp_nc_fh_l6": dict(model=3, L=6, param=2, chi=64, chi_in=64, consts=[1.0, 4.0, 0.0], up=1, dn=2, intdeg=1,
                             dtau=0.1, max_steps=100, judge=10, conv1=1e-4, conv2=1e-6),
    }
    for name, cs in cases.items():
        L, chi, model = cs["L"], cs["chi"], cs["model"]
        d = 3 if model == 2 else 4
        field = np.random.default_rng(31).uniform(0.8, 1.2, L - 1) if model == 2 else None
        c = neel(L, d, cs["up"], cs["dn"])
        tot = int(sum(np.argmax(np.abs(c[s])) for s in range(L))) if model == 2 else L
        params = {k: cs[k] for k in ("chi_in", "dtau", "max_steps", "judge", "conv1", "conv2")}
        extra = (f"{tot} {cs['intdeg']}\n{cs['chi_in']} {cs['dtau']!r} {cs['max_steps']} {cs['judge']} "
                 f"{cs['conv1']!r} {cs['conv2']!r}\n")
        path = os.path.join(tmp, name + ".txt")
        write_job(path, model, L, cs["param"], chi, cs["consts"], field, 0j, 0, 0, c, mode=7, extra=extra)
        stages, blocks = run_itp(path)
        res = collect(blocks, L, d, chi)
        steps = [st[-1][0] for st in stages]
        conv = [bool(st[-1][1] < st[-1][2]) for st in stages]
        for st in stages:  # no judgement of a stage's end sits at the criterion
            for j, t, cr in st[-2:]:
                assert abs(t / cr - 1.0) > 1e-4, (name, j, t, cr)
        meta = dict(model=model, L=L, param=cs["param"], chi=chi, d=d, consts=np.array(cs["consts"]), coeffs=c,
                    tot_num=tot, intdeg=cs["intdeg"], stage_steps=np.array(steps), stage_converged=np.array(conv),
                    judgements=np.array([[k, j, t, cr] for k, st in enumerate(stages) for j, t, cr in st]),
                    **{k: np.array(v) for k, v in params.items()})
        if field is not None:
            meta["field"] = field
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **meta, **res)
        print(name, "stages", steps, conv, "energy", res["energy"][-1], "kept", [(l > 0).sum() for l in res["lam"]])


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "checkpoint":
        checkpoint()
    elif len(sys.argv) > 1 and sys.argv[1] == "measures":
        measures()
    elif len(sys.argv) > 1 and sys.argv[1] == "itp_stages":
        itp_stages()
    elif len(sys.argv) > 1 and sys.argv[1] == "schedules":
        schedules()
    elif len(sys.argv) > 1 and sys.argv[1] == "nc":
        nc()
    elif len(sys.argv) > 1 and sys.argv[1] == "itp_nc":
        itp_nc()
    elif len(sys.argv) > 1 and sys.argv[1] == "nc_sched":
        nc_sched()
    else:
        main()
