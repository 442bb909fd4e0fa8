"""Benchmark: two-site DMRG sweeps at bond dimension D=1024 on one MI355X,
reduced-precision contractions against the FP64 path from the same state.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp32|bf16x3]
                    [--model hubbard|heisenberg]

Default workload (the configuration that actually reaches D=1024 at the
reference's default cutoff 1e-12): spinful Hubbard chain on native d=4 sites,
N=64, t=1, U=4 (particle-hole symmetric form), open boundaries, w=6, D ramped
128 -> 1024.  ``--model heisenberg`` runs the open S=1/2 Heisenberg chain
(N=100, d=2, w=5), whose ground state saturates near D=200 at that cutoff.

A step is one full DMRG sweep (left-to-right and back, 2(N-1) bond updates:
Lanczos ground state of H_eff, FP64 SVD truncation, environment update).
Setup (untimed): seeded random MPS, the D ramp and W warm-up sweeps at D.
Timed: K sweeps in the chosen precision; then, from the same starting
state, K sweeps with rocSOLVER dsyevd in place of the library's own
tridiagonalisation as the Gram eigensolver (the A/B of that kernel), and K
sweeps in FP64 for the comparison (|dE|/N between the final FP64 Rayleigh
quotients).  H_eff throughput comes from HIP events recorded
around every H_eff apply on the engine's stream.  The CPU baseline times the
oracle (NumPy complex128, the reference algorithm) on a bounded sample of
the same workload: one H_eff matvec, one truncated SVD and one environment
update at the D=1024 bond shape, combined with the measured Lanczos
iteration count per bond.
"""

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# "mixed" runs BF16x3 and FP32 applies; it is priced against the FP32 peak
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0, "mixed": 157.3}
METRIC = "DMRG sweep time + H_eff TFLOP/s at D=1024: mixed vs FP64, |ΔE/N|<1e-9"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def calibrate_copy(nbytes, reps):
    """Device-to-device copies of a known size, the reference point that
    scales the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    n = nbytes // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibration_copy_bytes": n * 8, "reps": reps}))


CPU_SAMPLE = r"""
import json, os, sys, time
import numpy as np
sys.path.insert(0, sys.argv[1])
from oracle import dmrg as od, linalg as ol, mps as om
spec = json.loads(sys.argv[2])
d, w, reps = spec["d"], spec["w"], spec["reps"]
rng = np.random.default_rng(0)
c = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)
out = []
for Dl, Dr, iters in spec["bonds"]:
    L, R = c(Dl, w, Dl), c(Dr, w, Dr)
    W1, W2 = c(w, d, d, w), c(w, d, d, w)
    th = c(Dl, d, d, Dr)
    A = c(Dl, d, Dr)
    mv = od.heff2_matvec(L, W1, W2, R, th.shape)
    x = th.reshape(-1)
    mv(x)                                   # warm-up (BLAS threads, page faults)
    best = None
    for _ in range(reps):
        t0 = time.perf_counter()
        for _ in range(iters):
            mv(x)
        t1 = time.perf_counter()
        ol.svd_truncated(th.reshape(Dl * d, d * Dr), 1e-12, max(Dl, Dr), 1)
        t2 = time.perf_counter()
        om.extend_left(L, A, W1, A)
        t3 = time.perf_counter()
        r = (t1 - t0, t2 - t1, t3 - t2)
        best = r if best is None or sum(r) < sum(best) else best
    out.append(best)
print(json.dumps(out))
"""


CPU_BOND = r"""
import json, sys, time
import numpy as np
sys.path.insert(0, sys.argv[1])
from oracle import dmrg as od, linalg as ol, mps as om
spec = json.loads(sys.argv[2])
Dl, Dr, d, w, maxdim = spec["Dl"], spec["Dr"], spec["d"], spec["w"], spec["maxdim"]
rng = np.random.default_rng(0)
c = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)
L, R = c(Dl, w, Dl), c(Dr, w, Dr)
W1, W2 = c(w, d, d, w), c(w, d, d, w)
th = c(Dl, d, d, Dr)
A = c(Dl, d, Dr)
mv2 = od.heff2_matvec(L, W1, W2, R, th.shape)
mv1 = od.heff1_matvec(L, W1, R, A.shape)
x2, x1 = th.reshape(-1), A.reshape(-1)
mv2(x2)
mv1(x1)                                     # warm-up (BLAS threads, page faults)
t = {}
t0 = time.perf_counter(); mv2(x2); t["heff2"] = time.perf_counter() - t0
t0 = time.perf_counter(); mv1(x1); t["heff1"] = time.perf_counter() - t0
t0 = time.perf_counter(); ol.svd_truncated(th.reshape(Dl * d, d * Dr), 1e-12, maxdim, 1)
t["svd"] = time.perf_counter() - t0
t0 = time.perf_counter(); om.extend_left(L, A, W1, A); t["env"] = time.perf_counter() - t0
print(json.dumps(t))
"""


def cpu_bond_sample(Dl, Dr, d, w, maxdim, threads):
    """One bond of the oracle (NumPy complex128, the reference algorithm):
    a two-site and a one-site H_eff matvec, the truncated SVD of theta and
    an environment update, timed once each after a warm-up, on `threads`
    BLAS threads (set before numpy is imported)."""
    env = dict(os.environ)
    for k in ("OPENBLAS_NUM_THREADS", "OMP_NUM_THREADS", "MKL_NUM_THREADS"):
        env[k] = str(threads)
    spec = {"Dl": Dl, "Dr": Dr, "d": d, "w": w, "maxdim": maxdim}
    p = subprocess.run([sys.executable, "-c", CPU_BOND, ROOT, json.dumps(spec)], env=env,
                       capture_output=True, text=True, timeout=1800)
    if p.returncode != 0:
        log("cpu bond sample failed:", p.stderr[-2000:])
        return None
    return json.loads(p.stdout.strip().splitlines()[-1])


def stratified_bonds(shapes, iters, strata=4):
    """Bond shapes of the GPU's timed sweep grouped into `strata` groups of
    equal bond count by cost (Gram order min(Dl d1, d2 Dr), then Dl Dr); the
    median bond of each group represents it.  Returns [(Dl, Dr, iters,
    group size)]."""
    rows = sorted(zip(shapes, iters), key=lambda r: (min(r[0][1] * r[0][2], r[0][3] * r[0][4]),
                                                      r[0][1] * r[0][4]))
    k = max(1, min(strata, len(rows)))
    out = []
    for g in range(k):
        grp = rows[g * len(rows) // k:(g + 1) * len(rows) // k]
        if not grp:
            continue
        (b, Dl, d1, d2, Dr), it = grp[len(grp) // 2]
        out.append((Dl, Dr, max(1, int(round(sum(r[1] for r in grp) / len(grp)))), len(grp)))
    return out


def cpu_baseline(shapes, iters, d, w, threads_all, strata=4):
    """The oracle (NumPy complex128 restatement of the reference algorithm)
    on the real bond shapes of the GPU's timed sweep: a stratified sample
    (median bond of each cost quartile), per bond its measured Lanczos
    matvec count + truncated SVD + environment update, each group's time
    scaled by its bond count; all cores (best of 3) and 1 thread (one run),
    BLAS threads set before numpy is imported."""
    sample = stratified_bonds(shapes, iters, strata)
    res = {}
    t_wall = time.time()
    for threads, reps in ((threads_all, 3), (1, 1)):
        env = dict(os.environ)
        for k in ("OPENBLAS_NUM_THREADS", "OMP_NUM_THREADS", "MKL_NUM_THREADS"):
            env[k] = str(threads)
        spec = {"d": d, "w": w, "reps": reps, "bonds": [[Dl, Dr, it] for Dl, Dr, it, _ in sample]}
        p = subprocess.run([sys.executable, "-c", CPU_SAMPLE, ROOT, json.dumps(spec)], env=env,
                           capture_output=True, text=True, timeout=1800)
        if p.returncode != 0:
            log("cpu baseline failed:", p.stderr[-2000:])
            return None
        times = json.loads(p.stdout.strip().splitlines()[-1])
        sweep_s = sum(sum(t) * cnt for t, (_, _, _, cnt) in zip(times, sample))
        res[threads] = (sweep_s, times)
        log("cpu baseline {} thread(s): sweep {:.1f}s from {}".format(threads, sweep_s, times))
    sweep_all, times_all = res[threads_all]
    desc = ", ".join("Dl={} Dr={} x{} bonds ({} matvecs {:.2f}s, SVD {:.2f}s, env {:.2f}s)".format(
        Dl, Dr, cnt, it, t[0], t[1], t[2]) for (Dl, Dr, it, cnt), t in zip(sample, times_all))
    return {"value": 1.0 / sweep_all, "unit": "sweeps/s", "cores": threads_all, "kind": "port",
            "sample": ("oracle (NumPy complex128 restatement of the reference algorithm) on the bond shapes of "
                       "the GPU's timed sweep, stratified into {} cost quartiles (median bond each, scaled by its "
                       "bond count, measured Lanczos matvecs per bond): {}; sweep {:.0f}s on {} BLAS threads "
                       "(best of 3; host has os.cpu_count()={}), {:.0f}s on 1 thread; wall {:.0f}s").format(
                len(sample), desc, sweep_all, threads_all, os.cpu_count(), res[1][0], time.time() - t_wall),
            "one_thread": {"value": 1.0 / res[1][0], "unit": "sweeps/s", "cores": 1}}


def run_sweeps(ch, psi, K, prec_name, maxdim):
    import torch
    from dmrgpy_amd import algorithms as alg
    from dmrgpy_amd.chain import RES_TOL
    from dmrgpy_amd.engine import PREC_NAMES
    eng = ch.eng
    prec = PREC_NAMES[prec_name]
    stats = alg.Stats()
    eng.prof_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(K):
        alg.dmrg_sweep(eng, psi, ch.H, maxdim, ch.cutoff, 2, prec, stats, res_tol=RES_TOL.get(prec, 0.0))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    tms, flops, calls = eng.prof_read()
    nbytes = eng.prof_read_bytes()
    eng.prof_enable(False)
    nrm = alg.overlap(eng, psi, psi).real
    e = (alg.expectation(eng, psi, ch.H, psi) / nrm).real
    docc = None
    if getattr(ch, "docc_mpo", None) is not None:
        docc = (alg.expectation(eng, psi, ch.docc_mpo, psi) / nrm).real
    return {"sweep_s": dt / K, "energy": e, "double_occupancy": docc,
            "heff_ms": tms[0], "heff_flops": flops[0], "heff_calls": calls[0],
            "heff_bytes": nbytes[0],
            "env_ms": tms[1], "svd_ms": tms[2], "svd_calls": calls[2],
            "trd_ms": tms[4], "trd_calls": calls[4], "trd_bytes": nbytes[4],
            "vec_ms": tms[3], "vec_calls": calls[3], "vec_bytes": nbytes[3],
            "iters_per_bond": sum(stats.lanczos_iters) / max(1, len(stats.lanczos_iters)),
            "svd_guided_fraction": sum(stats.svd_guided) / max(1, len(stats.svd_guided)),
            "max_bond": max(psi.bond_dims()),
            "shapes": list(stats.bond_shapes[-2 * (psi.length() - 1):]),
            "iters": list(stats.lanczos_iters[-2 * (psi.length() - 1):]),
            "truncation_weight": sum(t[2] for t in stats.truncation[-2 * (psi.length() - 1):])}


def heff_bulk(eng, mpo, D, prec_name, reps=20, seed=0):
    """H_eff applies at the bulk shape Dl = Dr = D of a chain (the MPO
    tensors of its two middle sites, random environments and theta in the
    engine layouts), timed with the library's HIP events on the engine
    stream: (ms per apply, FLOP per apply, algorithmic bytes per apply)."""
    import numpy as np
    import torch
    from dmrgpy_amd.engine import PREC_NAMES
    prec = PREC_NAMES[prec_name]
    n = len(mpo.W)
    W1, W2 = mpo.W[n // 2 - 1], mpo.W[n // 2]
    d1, d2 = W1.shape[3], W2.shape[3]
    wl, wr = W1.shape[1], W2.shape[2]
    g = torch.Generator(device="cpu").manual_seed(seed)
    et = eng.env_dtype(prec)
    p = eng.env_pitch(D)
    L = (torch.randn((1, wl, D, p), generator=g, dtype=torch.float64) / D).to(et).to(eng.device)
    R = (torch.randn((1, wr, D, p), generator=g, dtype=torch.float64) / D).to(et).to(eng.device)
    x = torch.randn((1, D * d1 * d2 * D), generator=g, dtype=torch.float64).to(eng.device)
    y = torch.empty_like(x)
    desc = eng.heff_desc(L, R, W1, W2, D, D, d1, d2, 0, prec)
    for _ in range(3):
        eng.heff_apply(desc, x, y)
    eng.sync()
    eng.prof_enable(True)
    eng.prof_read()            # clears nothing: read totals before and after
    t0, f0, c0 = (v[0] for v in eng.prof_read())
    b0 = eng.prof_read_bytes()[0]
    for _ in range(reps):
        eng.heff_apply(desc, x, y)
    eng.sync()
    t1, f1, c1 = (v[0] for v in eng.prof_read())
    b1 = eng.prof_read_bytes()[0]
    eng.prof_enable(False)
    calls = max(1, c1 - c0)
    return (t1 - t0) / calls, (f1 - f0) / calls, (b1 - b0) / calls


def workload_heff(args):
    """H_eff applies alone at the bulk D of the default model (the apply a
    D=1024 sweep spends its contraction time in), every precision; one JSON
    line per precision.  The shape rocprofv3 profiles are taken on."""
    from dmrgpy_amd.mpo import build_mpo
    from dmrgpy_amd.engine import get_engine
    from dmrgpy_amd.multioperator import heisenberg_terms, hubbard_terms
    spec = MODELS[args.model]
    n = args.n or spec["n"]
    terms = hubbard_terms(n, 1.0, 4.0) if args.model == "hubbard" else heisenberg_terms(n)
    mpo = build_mpo([spec["code"]] * n, terms)
    eng = get_engine(0)
    precs = [args.precision] if args.only_precision else ["fp32", "fp64", "bf16x3"]
    dims = [int(x) for x in args.heff_dims.split(",")] if args.heff_dims else [args.D]
    for D in dims:
        for pn in precs:
            ms, flops, nbytes = heff_bulk(eng, mpo, D, pn, reps=max(1, args.steps))
            mult = 3.0 if pn == "bf16x3" else 1.0
            tf = flops / (ms * 1e-3) / 1e12
            print(json.dumps({"workload": "H_eff apply at the bulk shape Dl=Dr={} of the {} (d={}, w={})".format(
                D, spec["desc"].format(n=n), spec["d"], spec["w"]), "D": D, "precision": pn, "ms_per_apply": ms,
                "gflop_per_apply": flops / 1e9, "tflops": tf, "peak": PEAK_TFLOPS[pn],
                "frac": mult * tf / PEAK_TFLOPS[pn], "algorithmic_bytes": nbytes, "applies": max(1, args.steps)}))


def measured_traffic(model, D, precision):
    """The rocprofv3 FETCH_SIZE / WRITE_SIZE measurement of this workload's
    timed-sweep H_eff applies committed under profiles/ (traffic and
    algorithmic bytes per apply of the profiled applies, and the command),
    or None when no measurement of this workload exists."""
    path = os.path.join(ROOT, "profiles", "{}_D{}_{}_heff_traffic.json".format(model, D, precision))
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def workload_tdvp(args):
    """Two-site TDVP after a local quench (configuration 5, sizes from the
    flags): S=1 Heisenberg chain, DMRG ground state (FP32 ramp, final FP64
    sweep), psi1 = psi2 = S^z_{N/2} |psi0>, quench_tdvp's steps of dt = 0.05 on
    H - E0 with Krylov dimension 20 and bond dimension cap D, FP64
    contractions.  Times `steps` TDVP steps after `warmup` untimed ones."""
    import numpy as np
    import torch
    from dmrgpy_amd import algorithms as alg
    from dmrgpy_amd.chain import Chain, RES_TOL
    from dmrgpy_amd.mpo import build_mpo
    from dmrgpy_amd.multioperator import heisenberg_terms
    n = args.n or 128
    D = args.D
    ch = Chain([3] * n, field="real")
    terms = heisenberg_terms(n)
    ch.set_hamiltonian(terms)
    eng = ch.eng
    np.random.seed(0)
    psi = ch._default_mps(maxm=64)
    guided = args.svd == "guided"
    t0 = time.perf_counter()
    for m in [m for m in (64, 128, 256, 512) if m < D] + [D, D]:
        e = alg.dmrg_sweep(eng, psi, ch.H, m, 1e-12, 2, 1, None, res_tol=RES_TOL[1])
        log("ground state ramp D={} E={:.10f} maxbond={} t={:.1f}s".format(m, e, max(psi.bond_dims()),
                                                                       time.perf_counter() - t0))
    e = alg.dmrg_sweep(eng, psi, ch.H, D, 1e-12, 2, 0, None)
    e0 = (alg.expectation(eng, psi, ch.H, psi) / alg.overlap(eng, psi, psi)).real
    Hs = build_mpo(ch.site_types, list(terms) + [(-e0, [("Id", 1)])])
    A = build_mpo(ch.site_types, [(1.0, [("Sz", n // 2)])])
    psi1 = alg.apply_mpo(eng, A, psi, 1e-12, D)
    psi2 = psi1.copy()
    alg.position(eng, psi1, 1)
    norm0 = float(abs(alg.overlap(eng, psi1, psi1))) ** 0.5
    dt = 0.05
    stats = alg.Stats()
    eng.svd_guided = guided   # the TDVP splits only (the ground state above used the full split)

    def step():
        alg.tdvp_step(eng, psi1, Hs, dt, 1e-12, D, 20, 0, stats)
        alg.normalize(eng, psi1)
        alg.scale_site(psi1, psi1.center or 1, norm0)

    corr = []
    for _ in range(args.warmup):
        corr.append(alg.overlap(eng, psi2, psi1))
        step()
    eng.prof_enable(True)
    stats.reset()
    torch.cuda.synchronize()
    ts = time.perf_counter()
    for k in range(args.steps):
        corr.append(alg.overlap(eng, psi2, psi1))
        step()
        if (k + 1) % 10 == 0:
            log("tdvp step {} t={:.2f} max bond {} elapsed {:.1f}s C={}{}".format(
                k + 1, (k + 1) * dt, max(psi1.bond_dims()), time.perf_counter() - ts, corr[-1],
                " guided {:.2f}".format(sum(stats.svd_guided) / max(1, len(stats.svd_guided))) if guided else ""))
    torch.cuda.synchronize()
    dt_s = (time.perf_counter() - ts) / args.steps
    tms, flops, calls = eng.prof_read()
    eng.prof_enable(False)
    kry = sum(stats.krylov_iters) / max(1, len(stats.krylov_iters))
    cpu = None
    if not args.no_cpu:
        # the oracle on one bond at the bond-dimension cap (every bulk bond of
        # a saturated step), scaled to a step: 2(N-1) two-site evolutions,
        # splits and environment updates, 2(N-2) one-site evolutions, each
        # evolution the measured mean Krylov matvec count
        t = cpu_bond_sample(D, D, 3, 5, D, args.cpu_threads)
        if t is not None:
            step = 2 * (n - 1) * (kry * t["heff2"] + t["svd"] + t["env"]) + 2 * (n - 2) * kry * t["heff1"]
            cpu = {"value": 1.0 / step, "unit": "steps/s", "cores": args.cpu_threads, "kind": "port",
                   "sample": ("oracle (NumPy complex128, the reference algorithm) on one bulk bond at Dl = Dr = {} "
                              "(d=3, w=5): two-site matvec {:.2f}s, one-site matvec {:.2f}s, truncated SVD {:.2f}s, "
                              "environment update {:.2f}s, scaled to a step with every bond at the cap "
                              "(2(N-1) splits, {:.1f} Krylov matvecs per evolution as measured on the GPU): {:.0f}s "
                              "per step on {} BLAS threads (one run; host os.cpu_count()={})").format(
                                  D, t["heff2"], t["heff1"], t["svd"], t["env"], kry, step, args.cpu_threads,
                                  os.cpu_count())}
            log("cpu baseline (TDVP step at the cap): {}".format(cpu))
    out = {"metric": "two-site TDVP step time after a local quench", "value": 1.0 / dt_s,
           "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": dt_s * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "fp64", "data": "synthetic: DMRG ground state of the model, seeded random initial MPS",
           "config": {"workload": "TDVP quench, open S=1 Heisenberg chain N={}, d=3, w=5, D cap {}, dt=0.05, "
                                  "Krylov dim 20, Sz at site {}{}".format(
                                      n, D, n // 2, ", subspace-guided splits" if guided else ""),
                      "N": n, "D": D, "E0": e0, "svd": args.svd},
           "svd_guided_fraction": sum(stats.svd_guided) / max(1, len(stats.svd_guided)),
           "max_bond": max(psi1.bond_dims()),
           "krylov_iters_per_call": kry,
           "time_split_s_per_step": {"heff": tms[0] / 1e3 / args.steps, "env": tms[1] / 1e3 / args.steps,
                                     "svd": tms[2] / 1e3 / args.steps},
           "heff_tflops": flops[0] / (tms[0] * 1e-3) / 1e12 if tms[0] > 0 else 0.0,
           "correlator": [[c.real, c.imag] for c in corr],
           "total_s": dt_s * args.steps, "cpu_baseline": cpu}
    print(json.dumps(out))


SMALL_CPU = r"""
import json, sys, time
import numpy as np
sys.path.insert(0, sys.argv[1])
from oracle import dmrg as od, mpo as omp, mps as om, ed
n, D, ns = int(sys.argv[2]), int(sys.argv[3]), int(sys.argv[4])
np.random.seed(0)
psi = om.random_mps([2] * n, D)
Ws = omp.to_mpo([2] * n, ed.heisenberg_terms(n))
t0 = time.perf_counter()
es = []
od.dmrg(psi, Ws, [D] * ns, 1e-12, 2, callback=lambda k, e, p: es.append(e))
print(json.dumps({"t": time.perf_counter() - t0, "energies": es}))
"""


def workload_small(args):
    """The reference's CPU-runnable configuration (configuration 1): open
    S=1/2 Heisenberg chain N=32, D=64, 6 sweeps, FP64, random initial MPS
    from np.random.seed(0) drawn as the reference draws it.  The GPU session
    and the oracle (the reference algorithm in NumPy, one full run on the
    host, 1 BLAS thread and all cores) run the same sweeps from the same
    state; their per-sweep energies must agree to 1e-10 per site."""
    import numpy as np
    import torch
    from dmrgpy_amd.chain import Chain
    from dmrgpy_amd.multioperator import heisenberg_terms
    n = args.n or 32
    D = 64 if args.D == 1024 else args.D
    ns = 6
    ch = Chain([2] * n)
    ch.set_sweep_params(D, ns, 1e-12, 0.0)
    ch.set_hamiltonian(heisenberg_terms(n))
    np.random.seed(0)
    ch.wf0 = ch._default_mps()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ch.gs_energy()
    torch.cuda.synchronize()
    t_gpu = time.perf_counter() - t0
    es = list(ch.sweep_energies)
    cpu = {}
    for threads in (1, args.cpu_threads):
        env = dict(os.environ)
        for k in ("OPENBLAS_NUM_THREADS", "OMP_NUM_THREADS", "MKL_NUM_THREADS"):
            env[k] = str(threads)
        p = subprocess.run([sys.executable, "-c", SMALL_CPU, ROOT, str(n), str(D), str(ns)], env=env,
                           capture_output=True, text=True, timeout=1800)
        if p.returncode != 0:
            log("oracle run failed:", p.stderr[-2000:])
            continue
        cpu[threads] = json.loads(p.stdout.strip().splitlines()[-1])
    ref = cpu.get(args.cpu_threads) or cpu.get(1)
    dE = max(abs(a - b) for a, b in zip(es, ref["energies"])) / n if ref else None
    out = {"metric": "DMRG time of the reference's CPU configuration (N=32, D=64, 6 sweeps, FP64)",
           "value": ns / t_gpu, "unit": "sweeps/s", "n_gpus": 1, "steps": ns, "warmup": 0,
           "ms_per_step": t_gpu / ns * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "fp64", "data": "synthetic: seeded random initial MPS (np.random.seed(0))",
           "config": {"workload": "two-site DMRG, open S=1/2 Heisenberg N={}, D={}, {} sweeps".format(n, D, ns),
                      "N": n, "D": D},
           "energies": es, "max_dE_per_site_vs_oracle": dE,
           "cpu_baseline": [{"value": ns / r["t"], "unit": "sweeps/s", "cores": t, "kind": "port",
                             "sample": "the full 6-sweep run of the oracle"} for t, r in sorted(cpu.items())]}
    print(json.dumps(out))


def workload_gradients(args):
    """Coupling gradients (configuration 4, sizes from the flags): random-bond
    XXZ chain, J_i, Delta_i ~ U[0.5, 1.5] from default_rng(seed 4), D cap,
    FP32 contractions with a final FP64 sweep; all dE0/dJ_i, dE0/dDelta_i
    from one backward pass of the MPO-gradient kernels, two couplings checked
    by central finite differences of warm-started DMRG energies."""
    import numpy as np
    import torch
    from dmrgpy_amd.chain import Chain
    from dmrgpy_amd.gradients import xxz_gradients, xxz_terms
    n = args.n or 64
    D = args.D
    rng = np.random.default_rng(4)
    J = rng.uniform(0.5, 1.5, n - 1)
    Dz = rng.uniform(0.5, 1.5, n - 1)

    def solve(Jx, Dx, wf=None, nsw=8):
        ch = Chain([2] * n, precision="fp32", field="real")
        ch.set_sweep_params(D, nsw, 1e-12, 0.0)
        ch.precision_schedule = ["fp32"] * (nsw - 1) + ["fp64"]
        ch.set_hamiltonian(xxz_terms(Jx, Dx))
        if wf is not None:
            ch.set_wavefunction(wf.copy())
        else:
            np.random.seed(1)
        return ch, ch.gs_energy()

    t0 = time.perf_counter()
    ch, e = solve(J, Dz)
    t_gs = time.perf_counter() - t0
    psi = ch.gs_wavefunction()
    xxz_gradients(ch.eng, psi, J, Dz)       # warm-up (workspace growth)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        gJ, gD = xxz_gradients(ch.eng, psi, J, Dz)
    torch.cuda.synchronize()
    t_grad = (time.perf_counter() - t0) / args.steps
    h = 1e-4
    checks = []
    for i in (0, n // 2):
        Jp, Jm = J.copy(), J.copy()
        Jp[i] += h
        Jm[i] -= h
        fd = (solve(Jp, Dz, psi, 4)[1] - solve(Jm, Dz, psi, 4)[1]) / (2 * h)
        checks.append({"coupling": "J_{}".format(i + 1), "grad": gJ[i], "fd": fd,
                       "rel_err": abs(fd - gJ[i]) / abs(gJ[i])})
    out = {"metric": "dE0/dJ and dE0/dDelta for every coupling (backward MPO-gradient kernels)",
           "value": 1.0 / t_grad, "unit": "gradient passes/s", "n_gpus": 1, "steps": args.steps,
           "warmup": 1, "ms_per_step": t_grad * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "fp64", "data": "synthetic: seeded random-bond XXZ couplings",
           "config": {"workload": "random-bond XXZ chain N={}, D={}, {} couplings".format(n, D, 2 * (n - 1)),
                      "N": n, "D": D},
           "ground_state_s": t_gs, "energy": e, "fd_checks": checks}
    if not args.no_cpu:
        # the oracle on one bulk bond at the cap: the gradient pass is one
        # left and one right environment sweep plus one local contraction
        # per site (~3N environment updates); a DMRG sweep is 2(N-1) bonds of
        # 4 Lanczos matvecs, a split and an environment update
        t = cpu_bond_sample(D, D, 2, 5, D, args.cpu_threads)
        if t is not None:
            grad_s = 3 * n * t["env"]
            sweep_s = 2 * (n - 1) * (4 * t["heff2"] + t["svd"] + t["env"])
            out["cpu_baseline"] = {
                "value": 1.0 / grad_s, "unit": "gradient passes/s", "cores": args.cpu_threads, "kind": "port",
                "sample": ("oracle (NumPy complex128, the reference algorithm) on one bulk bond at Dl = Dr = {} (d=2, "
                           "w=5): environment update {:.3f}s, two-site matvec {:.3f}s, truncated SVD {:.3f}s; gradient "
                           "pass = 3N environment updates = {:.2f}s, DMRG sweep = 2(N-1) x (4 matvecs + SVD + "
                           "environment) = {:.1f}s, on {} BLAS threads (one run)").format(
                               D, t["env"], t["heff2"], t["svd"], grad_s, sweep_s, args.cpu_threads),
                "dmrg_sweep_s": sweep_s}
            log("cpu baseline (gradients): {}".format(out["cpu_baseline"]))
    print(json.dumps(out))


MODELS = {
    "hubbard": {"n": 64, "code": 1, "d": 4, "w": 6, "ramp": [128, 256, 512],
                "desc": "spinful Hubbard chain (native d=4 sites), N={n}, t=1, U=4, open, w=6"},
    "heisenberg": {"n": 100, "code": 2, "d": 2, "w": 5, "ramp": [64, 128, 256, 512],
                   "desc": "open S=1/2 Heisenberg chain, N={n}, J=1, d=2, w=5"},
}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--model", default="hubbard", choices=sorted(MODELS))
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--D", type=int, default=1024)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "bf16x3", "mixed"])
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--svd", default="full", choices=["guided", "full"],
                    help="two-site splits: the full split (the reference's rule), or the subspace split guided "
                         "by the neighbouring isometry (falls back to the full split when it misses weight)")
    ap.add_argument("--syevd-check", action="store_true",
                    help="also time the sweeps with rocSOLVER dsyevd as the Gram eigensolver from the same state")
    ap.add_argument("--no-syevd-check", action="store_true", help="(default; kept for older command lines)")
    ap.add_argument("--full-reduction-check", action="store_true",
                    help="also time the sweeps from the same state with the complete tridiagonal reduction "
                         "(eigensolver mode 3) instead of the default early-stopping one")
    ap.add_argument("--no-full-svd-check", action="store_true",
                    help="skip the timed sweep with the full split from the same state (guided runs only)")
    ap.add_argument("--cpu-threads", type=int,
                    default=int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1),
                    help="BLAS threads of the all-cores CPU baseline (default: OMP_NUM_THREADS, the host "
                         "share, else os.cpu_count())")
    ap.add_argument("--heff-dims", default="", help="--workload heff: comma-separated bond dimensions")
    ap.add_argument("--only-precision", action="store_true",
                    help="--workload heff: only the --precision given (default: fp32, fp64 and bf16x3)")
    ap.add_argument("--workload", default="dmrg", choices=["dmrg", "small", "tdvp", "gradients", "heff"],
                    help="dmrg: the headline sweep benchmark; tdvp / gradients: the TDVP-quench and "
                         "coupling-gradient configurations (sizes from --n/--D/--steps)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run device copies of this size (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: runs on one GPU only (DMRG sweeps are sequential over bonds; H_eff would "
              "shard over the row index of the left environment with one all-gather of the Krylov "
              "vector per matvec); --gpus {} is not supported".format(args.gpus))
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy, max(1, args.steps))
        return
    if args.workload == "tdvp":
        workload_tdvp(args)
        return
    if args.workload == "gradients":
        workload_gradients(args)
        return
    if args.workload == "small":
        workload_small(args)
        return
    if args.workload == "heff":
        workload_heff(args)
        return

    import numpy as np
    from dmrgpy_amd import algorithms as alg
    from dmrgpy_amd.chain import Chain, RES_TOL
    from dmrgpy_amd.engine import PREC_NAMES
    from dmrgpy_amd.multioperator import heisenberg_terms, hubbard_terms

    spec = MODELS[args.model]
    n = args.n or spec["n"]
    D = args.D
    terms = hubbard_terms(n, 1.0, 4.0) if args.model == "hubbard" else heisenberg_terms(n)
    ch = Chain([spec["code"]] * n, precision=args.precision, field="real")
    ch.set_sweep_params(D, 1, 1e-12, 0.0)
    ch.set_hamiltonian(terms)
    ch.docc_mpo = None
    if args.model == "hubbard":   # double occupancy (1/N) sum_i <n_up n_dn>_i
        from dmrgpy_amd.mpo import build_mpo
        ch.docc_mpo = build_mpo(ch.site_types, [(1.0 / n, [("Nup", i), ("Ndn", i)]) for i in range(1, n + 1)])
    np.random.seed(0)
    ramp = [m for m in spec["ramp"] if m < D]
    psi = ch._default_mps(maxm=ramp[0] if ramp else D)
    eng = ch.eng
    eng.svd_guided = args.svd == "guided"
    t_setup = time.perf_counter()
    prec = PREC_NAMES[args.precision]
    for m in ramp:
        e = alg.dmrg_sweep(eng, psi, ch.H, m, ch.cutoff, 2, prec, None, res_tol=RES_TOL.get(prec, 0.0))
        log("ramp D={} E={:.12f} maxbond={} t={:.1f}s".format(m, e, max(psi.bond_dims()),
                                                             time.perf_counter() - t_setup))
    for k in range(args.warmup):
        e = alg.dmrg_sweep(eng, psi, ch.H, D, ch.cutoff, 2, prec, None, res_tol=RES_TOL.get(prec, 0.0))
        log("warmup sweep {} D={} E={:.12f} maxbond={} t={:.1f}s".format(k, D, e, max(psi.bond_dims()),
                                                                        time.perf_counter() - t_setup))
    start = psi.copy()
    res = run_sweeps(ch, psi, args.steps, args.precision, D)
    log("timed {}: {}".format(args.precision, {k: v for k, v in res.items() if k not in ("shapes", "iters")}))
    ressy = None
    if args.syevd_check:
        eng.set_eigensolver(0)
        ressy = run_sweeps(ch, start.copy(), args.steps, args.precision, D)
        eng.set_eigensolver(1)   # back to the default
        log("timed {} with rocSOLVER dsyevd: {}".format(
            args.precision, {k: v for k, v in ressy.items() if k not in ("shapes", "iters")}))
    resred = None
    if args.full_reduction_check:
        eng.set_eigensolver(3)
        resred = run_sweeps(ch, start.copy(), args.steps, args.precision, D)
        eng.set_eigensolver(1)
        log("timed {} with the complete reduction: {}".format(
            args.precision, {k: v for k, v in resred.items() if k not in ("shapes", "iters")}))
    resfull = None
    if args.svd == "guided" and not args.no_full_svd_check:
        eng.svd_guided = False
        resfull = run_sweeps(ch, start.copy(), args.steps, args.precision, D)
        eng.svd_guided = True
        log("timed {} with the full split: {}".format(args.precision, resfull))
    res64 = None
    if not args.no_fp64:
        psi64 = start.copy()
        res64 = run_sweeps(ch, psi64, args.steps, "fp64", D)
        log("timed fp64: {}".format({k: v for k, v in res64.items() if k not in ("shapes", "iters")}))

    # the H_eff apply at the bulk shape Dl = Dr = D itself (the sweep's bonds
    # stay below D at the reference cutoff): the same kernels on the model's
    # middle MPO tensors, HIP events on the engine stream
    bulk_prec = "fp32" if args.precision == "mixed" else args.precision
    bulk_ms, bulk_flops, bulk_bytes = heff_bulk(eng, ch.H, D, bulk_prec, reps=10)
    bulk_mult = 3.0 if bulk_prec == "bf16x3" else 1.0
    bulk_tf = bulk_flops / (bulk_ms * 1e-3) / 1e12
    log("bulk H_eff at D={}: {:.3f} ms, {:.1f} TFLOP/s".format(D, bulk_ms, bulk_tf))
    heff_tflops = res["heff_flops"] / (res["heff_ms"] * 1e-3) / 1e12 if res["heff_ms"] > 0 else 0.0
    avg_ms = res["heff_ms"] / max(1, res["heff_calls"])
    flops_per = res["heff_flops"] / max(1, res["heff_calls"])
    peak = PEAK_TFLOPS[args.precision]
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(res["shapes"], res["iters"], spec["d"], spec["w"], args.cpu_threads)
        log("cpu baseline: {}".format(cpu))
    # split-BF16x3 runs three BF16 MFMA products per FP32-equivalent flop:
    # priced against the BF16 peak with the products it actually issues
    mfma_mult = 3.0 if args.precision == "bf16x3" else 1.0
    achieved = mfma_mult * flops_per / (avg_ms * 1e-3) / 1e12 if avg_ms > 0 else 0.0
    prof = measured_traffic(args.model, D, args.precision)
    alg_bytes = res["heff_bytes"] / max(1, res["heff_calls"])
    traffic, traffic_note = None, None
    if prof is not None:
        # measured bytes per algorithmic byte of the profiled applies, applied
        # to this run's algorithmic bytes (the profiled sweeps follow this
        # command's trajectory when its arguments match)
        ratio = prof["traffic_bytes_per_apply"] / prof["algorithmic_bytes_per_apply"]
        traffic = ratio * alg_bytes
        traffic_note = "rocprofv3 FETCH_SIZE+WRITE_SIZE of {} applies of `{}`: {:.3f} GB measured per apply, " \
                       "{:.2f}x its algorithmic bytes; scaled to this run's algorithmic bytes".format(
                           prof.get("applies", "?"), prof.get("profiled_command", "?"),
                           prof["traffic_bytes_per_apply"] / 1e9, ratio)
    out = {
        "metric": METRIC,
        "value": 1.0 / res["sweep_s"],
        "unit": "sweeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["sweep_s"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "{}-contract/fp64-krylov-svd".format(args.precision),
        "data": "synthetic: seeded random initial MPS (np.random.seed(0)), model Hamiltonian",
        "config": {"workload": "two-site DMRG sweep at D={}: {}, D ramp {} then {} warm-up sweep(s), "
                               "{} contractions".format(D, spec["desc"].format(n=n), ramp, args.warmup,
                                                        args.precision),
                   "model": args.model, "N": n, "D": D, "d": spec["d"], "w": spec["w"], "cutoff": 1e-12,
                   "precision": args.precision, "field": "real", "svd": args.svd},
        "max_bond": res["max_bond"],
        "heff_tflops": heff_tflops,
        "heff_calls": res["heff_calls"],
        "lanczos_iters_per_bond": res["iters_per_bond"],
        "svd_guided_fraction": res["svd_guided_fraction"],
        "sweep_s": res["sweep_s"],
        "time_split_s_per_sweep": {"heff": res["heff_ms"] / 1e3 / args.steps,
                                   "env": res["env_ms"] / 1e3 / args.steps,
                                   "svd": res["svd_ms"] / 1e3 / args.steps},
        "energy": res["energy"],
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": traffic, "traffic_source": traffic_note,
                     "algorithmic_bytes": alg_bytes,
                     "kernel": "H_eff two-site apply (L-GEMM, W small-op, R-GEMM), {:.3g} GFLOP per apply "
                               "averaged over the sweep's bonds{}, {:.3f} ms average".format(
                                   flops_per / 1e9, " (x3 BF16 products)" if mfma_mult > 1 else "", avg_ms)},
    }
    out["heff_bulk"] = {"D": D, "precision": bulk_prec, "ms_per_apply": bulk_ms, "gflop_per_apply": bulk_flops / 1e9,
                        "tflops": bulk_tf, "peak": PEAK_TFLOPS[bulk_prec],
                        "frac": bulk_mult * bulk_tf / PEAK_TFLOPS[bulk_prec],
                        "kernel": "H_eff two-site apply at Dl = Dr = {} with the model's bulk MPO tensors and random "
                                  "environments, 10 applies".format(D)}
    if res["trd_calls"] > 0 and res["trd_ms"] > 0:
        # the Gram eigensolver's tridiagonalisation, the largest cost of a sweep:
        # algorithmic bytes = the lower triangle of the trailing matrix once per column
        gbs = res["trd_bytes"] / (res["trd_ms"] * 1e-3) / 1e9
        out["roofline_tridiagonalisation"] = {
            "bound": "hbm", "achieved": gbs, "peak": 8000.0, "unit": "GB/s", "frac": gbs / 8000.0, "traffic": None,
            "algorithmic_bytes": res["trd_bytes"] / res["trd_calls"],
            "kernel": "own Householder tridiagonalisation of the Gram matrices (trd_col_k + trd_symv_k per column, "
                      "panel GEMMs), {} reductions per sweep, {:.2f} ms average, {:.2f} s per sweep".format(
                          int(res["trd_calls"] / args.steps), res["trd_ms"] / res["trd_calls"],
                          res["trd_ms"] / 1e3 / args.steps)}
    if res["vec_calls"] > 0 and res["vec_ms"] > 0:
        # Lanczos vector passes (alpha/beta dots, re-orthogonalisation dot and
        # subtract, norms): bytes = the vectors each pass reads and writes once
        vgbs = res["vec_bytes"] / (res["vec_ms"] * 1e-3) / 1e9
        out["krylov_vectors"] = {"bound": "hbm", "achieved": vgbs, "peak": 8000.0, "unit": "GB/s",
                                 "frac": vgbs / 8000.0, "passes_per_sweep": int(res["vec_calls"] / args.steps),
                                 "s_per_sweep": res["vec_ms"] / 1e3 / args.steps}
    if ressy is not None:
        out["rocsolver_syevd"] = {"sweep_s": ressy["sweep_s"], "energy": ressy["energy"],
                                  "svd_s_per_sweep": ressy["svd_ms"] / 1e3 / args.steps,
                                  "max_bond": ressy["max_bond"],
                                  "dE_per_site_vs_own_eigensolver": abs(res["energy"] - ressy["energy"]) / n}
        out["speedup_own_eigensolver_vs_syevd"] = ressy["sweep_s"] / res["sweep_s"]
    if resred is not None:
        out["complete_reduction"] = {"sweep_s": resred["sweep_s"], "energy": resred["energy"],
                                     "svd_s_per_sweep": resred["svd_ms"] / 1e3 / args.steps,
                                     "max_bond": resred["max_bond"],
                                     "dE_per_site_vs_early_stop": abs(res["energy"] - resred["energy"]) / n}
        out["speedup_early_stop_vs_complete_reduction"] = resred["sweep_s"] / res["sweep_s"]
    if resfull is not None:
        out["full_svd"] = {"sweep_s": resfull["sweep_s"], "energy": resfull["energy"],
                           "time_split_s_per_sweep": {"heff": resfull["heff_ms"] / 1e3 / args.steps,
                                                      "env": resfull["env_ms"] / 1e3 / args.steps,
                                                      "svd": resfull["svd_ms"] / 1e3 / args.steps},
                           "max_bond": resfull["max_bond"],
                           "dE_per_site_vs_guided": abs(res["energy"] - resfull["energy"]) / n}
        out["speedup_guided_vs_full_svd"] = resfull["sweep_s"] / res["sweep_s"]
    if res64 is not None:
        out["fp64"] = {"sweep_s": res64["sweep_s"], "energy": res64["energy"],
                       "heff_tflops": res64["heff_flops"] / (res64["heff_ms"] * 1e-3) / 1e12,
                       "lanczos_iters_per_bond": res64["iters_per_bond"],
                       "time_split_s_per_sweep": {"heff": res64["heff_ms"] / 1e3 / args.steps,
                                                  "env": res64["env_ms"] / 1e3 / args.steps,
                                                  "svd": res64["svd_ms"] / 1e3 / args.steps}}
        out["speedup_vs_fp64"] = res64["sweep_s"] / res["sweep_s"]
        out["dE_per_site_vs_fp64"] = abs(res["energy"] - res64["energy"]) / n
        # observables of the final states against FP64 (truncation weight:
        # summed discarded weight of the last timed sweep's 2(N-1) splits)
        out["observables_vs_fp64"] = {
            "energy": [res["energy"], res64["energy"]],
            "double_occupancy": [res["double_occupancy"], res64["double_occupancy"]],
            "truncation_weight": [res["truncation_weight"], res64["truncation_weight"]]}
        if res["double_occupancy"] is not None:
            out["observables_vs_fp64"]["d_double_occupancy"] = abs(res["double_occupancy"] -
                                                                   res64["double_occupancy"])
    if cpu is not None:
        out["cpu_baseline"] = {k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample")}
        out["cpu_baseline_one_thread"] = cpu["one_thread"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
