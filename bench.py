"""Benchmark: RHF Fock build (fused Rys ERI + J/K) on one MI355X.

Workload (default): a seeded cluster of 32 water molecules in a 10 A cube
with the even-tempered s,p,d basis (512 shells, 992 cartesian functions, up
to 5 Rys roots), Schwarz threshold 1e-12.  One step = one full two-electron
Fock build G = J - K/2 from a density already resident in HBM
(core-Hamiltonian guess density): every surviving canonical shell quartet's
ERIs are recomputed and contracted.  The headline value is the FP64 build;
`mixed_vs_fp64` times the same build with FP32 recurrences for the tasks
whose Schwarz bound times the largest density block they read is below
tau_mix (1e-8 .. 1e-5), with the energy error against the FP64 build.
n-hexadecane C16H34/6-31G is timed the same way under `extra_workloads`.

Prints one JSON line (metric, value = canonical surviving ERI quartets per
second, ms_per_step = Fock-build time per iteration, roofline of the
dominant kernel from HIP events, CPU baselines from the oracle on bounded
samples).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 (vector and matrix) dense peak
HBM_PEAK = 8000.0         # GB/s

# Counter-measured figures of the dominant kernel per launch (one build: the
# sum over the launch's dispatches), from the committed rocprofv3 summary
# profiles/valu_counters_fock.md, keyed by (workload, precision, kernel):
#   TRAFFIC: HBM bytes, (FETCH_SIZE x 2 + WRITE_SIZE) x 1024 (KiB counters,
#            each scaled by what it reads for a 256 MiB device copy: FETCH
#            half the bytes read, WRITE all the bytes written);
#   COUNTER_FLOPS: executed FP64 VALU flops, 64 x (2 FMA + ADD + MUL + TRANS)
#            wavefront instructions.
TRAFFIC = {
    ("w32_et", "fp64", "jk_group_kernel<0, 1, 0, 0, false, true>"): (2 * 1.463e6 + 2.614e6) * 1024,
    ("c16h34_631g", "fp64", "jk_group_kernel<0, 1, 0, 0, false, false>"): (2 * 2.229e4 + 1.484e5) * 1024,
}
COUNTER_FLOPS = {
    ("w32_et", "fp64", "jk_group_kernel<0, 1, 0, 0, false, true>"): 6.6876e10,
    ("c16h34_631g", "fp64", "jk_group_kernel<0, 1, 0, 0, false, false>"): 1.4862e10,
}


def core_guess_density(m):
    from librint_amd import scf
    S = scf.int1e(m, "ovlp")
    H = scf.int1e(m, "kin") + scf.int1e(m, "nuc")
    s, U = np.linalg.eigh(S)
    X = U / np.sqrt(s)
    e, C = np.linalg.eigh(X.T @ H @ X)
    C = X @ C
    nocc = m.nelectron // 2
    return 2.0 * C[:, :nocc] @ C[:, :nocc].T


def workload(name):
    from librint_amd import mol
    if name == "c16h34_631g":
        return mol.c16h34_631g(), 1e-12
    if name == "h2o_sto3g":
        return mol.h2o_sto3g(), 0.0
    if name.startswith("alkane"):
        return mol.alkane_631g(int(name[6:])), 1e-12
    if name.startswith("w") and name.endswith("_et"):
        return mol.water_cluster(int(name[1:-3])), 1e-12
    raise SystemExit("unknown workload " + name)


def cpu_baseline(m, P, seconds, seed=0):
    """Oracle (plain-C restatement of integral2e_fock: full nbas^4 loop, no
    symmetry, no screening, FP64, one core) on a random sample of ordered
    shell quartets; returns the rate and the extrapolated Fock-build time."""
    from oracle import refcpu
    nb = m.nbas
    rng = np.random.default_rng(seed)
    total = 0
    t_used = 0.0
    batch = 2000
    while t_used < seconds:
        q = rng.integers(0, nb, size=(batch, 4)).astype(np.int32)
        t0 = time.perf_counter()
        refcpu.fock2e_sample(m, P, q)
        t_used += time.perf_counter() - t0
        total += batch
    rate = total / t_used
    return rate, float(nb) ** 4 / rate, total, t_used


def cpu_threaded(m, P, tau, seconds, stride=16):
    """Second CPU comparator: the oracle's Schwarz-screened, 8-fold-symmetric,
    threaded Fock build (the quartet set the GPU evaluates), timed over whole
    interleaved slices of the bra pairs (ij % stride == phase) until `seconds`
    have passed; the Schwarz diagonal is computed first and not timed (the
    GPU computes it once per context, outside the timed builds as well)."""
    from oracle import refcpu
    nth = refcpu.host_threads()
    Q = refcpu.schwarz(m, nth)
    done, used, phases = 0, 0.0, 0
    while phases < stride and used < seconds:
        t0 = time.perf_counter()
        _, c = refcpu.fock2e_screened(m, P, tau, Q=Q, stride=stride, phase=phases, nthreads=nth)
        used += time.perf_counter() - t0
        done += c
        phases += 1
    return done / used, nth, done, used, phases


def bench_shares(args, m, P, ctx):
    """One device's share of the Fock build, for every share of a
    G-device split (the deal of set_shard): each share timed on its own at
    its real shape, the whole build beside them, and the sum of the shares'
    G's checked against the whole G."""
    import torch
    dev = torch.device("cuda:0")
    Pd = torch.from_numpy(np.ascontiguousarray(P)).to(dev)
    Gd = torch.empty_like(Pd)

    def timed():
        for _ in range(args.warmup):
            ctx.fock_device(Pd, Gd)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.fock_device(Pd, Gd)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / args.steps * 1e3

    whole_ms = timed()
    G = Gd.cpu().numpy()
    whole = ctx.stats()
    shares, Gsum = [], np.zeros_like(G)
    for k in range(args.shares):
        ctx.set_shard(args.shares, k)
        ms = timed()
        Gsum += Gd.cpu().numpy()
        st = ctx.stats()
        shares.append({"share": k, "ms": ms, "quartets": int(st["quartets"]),
                       "prim_quartets": int(st["prim_quartets"]), "flops": st["flops"]})
    ctx.set_shard(1, 0)
    ms = np.array([x["ms"] for x in shares])
    fl = np.array([x["flops"] for x in shares])
    line = {
        "metric": "RHF Fock-build s/iter, one device's share of a %d-device split" % args.shares,
        "value": float(ms.max()) * 1e-3, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": float(ms.max()), "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic geometry, core-Hamiltonian density",
        "config": {"workload": args.workload, "nao": m.nao(), "schwarz": ctx.stats()["schwarz_tol"],
                   "shares": args.shares, "quartets": int(whole["quartets"])},
        "whole_ms": whole_ms,
        "share_ms_max_over_mean": float(ms.max() / ms.mean()),
        "share_flops_max_over_mean": float(fl.max() / fl.mean()),
        "speedup_if_shares_ran_in_parallel": whole_ms / float(ms.max()),
        "max_abs_sum_of_shares_minus_whole": float(np.abs(Gsum - G).max()),
        "shares": shares,
    }
    print(json.dumps(line), flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def bench_grad(args):
    """RHF energy gradient w.r.t. every nuclear coordinate and Gaussian
    exponent: converged SCF (untimed), then K timed gradient evaluations at
    that density (adjoint Rys kernel on the GPU + host one-electron parts).
    Optionally checks against FD of re-converged energies (5-point central
    stencil, steps 1e-4 bohr and 1e-4*max(zeta, 1) in the exponents; the
    3-point quotient at the same step is reported beside it)."""
    from librint_amd import autograd
    # the gradient workload is the 8-water cluster unless another one is named
    name = "w8_et" if args.workload == "w32_et" else args.workload
    m, _ = workload(name)
    tau = 1e-14
    E0, P0, it, ctx = autograd.rhf(m, conv=1e-10, schwarz_tol=tau)
    for _ in range(args.warmup):
        autograd.gradient(ctx, P0)
    t0 = time.perf_counter()
    kms = []
    for _ in range(args.steps):
        gn, de = autograd.gradient(ctx, P0)
        kms.append(ctx.stats()["grad_kernel_ms"])
    wall = (time.perf_counter() - t0) / args.steps
    st = ctx.stats()
    slots = autograd.exponent_slots(m)
    line = {
        "metric": "RHF analytic gradient (nuclear + exponents) s/eval",
        "value": wall, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "seeded synthetic water cluster, even-tempered s,p,d basis",
        "config": {"workload": name, "nao": m.nao(), "natm": len(m._atm), "exponents": len(slots),
                   "schwarz": tau, "quartets": int(st["quartets"]), "scf_energy": E0, "scf_iterations": it},
        "adjoint_kernel_ms": float(np.mean(kms)),
        "grad_norm_nuclear": float(np.linalg.norm(gn)),
    }
    ctx.close()
    if args.fd != "0":
        env0 = m._env.copy()
        params = [("R", a, d, int(m._atm[a][1]) + d) for a in range(len(m._atm)) for d in range(3)]
        params += [("exp", -1, -1, int(k)) for k in slots]
        if args.fd != "all":
            rng = np.random.default_rng(5)
            params = [params[i] for i in rng.choice(len(params), int(args.fd), replace=False)]
        errs, errs3 = [], []
        for kind, a, d, k in params:
            h = 1e-4 if kind == "R" else 1e-4 * max(env0[k], 1.0)
            es = {}
            for sg in (2, 1, -1, -2):
                m._env[k] = env0[k] + sg * h
                E, _, _, c = autograd.rhf(m, conv=1e-10, schwarz_tol=tau, P0=P0)
                c.close()
                es[sg] = E
            m._env[:] = env0
            fd3 = (es[1] - es[-1]) / (2 * h)
            fd = (8.0 * (es[1] - es[-1]) - (es[2] - es[-2])) / (12.0 * h)
            an = gn[a, d] if kind == "R" else de[k]
            errs.append(abs(fd - an))
            errs3.append(abs(fd3 - an))
            print("fd %s %d %d env[%d] fd5 %.10f fd3 %.10f analytic %.10f diff5 %.2e diff3 %.2e"
                  % (kind, a, d, k, fd, fd3, an, abs(fd - an), abs(fd3 - an)), file=sys.stderr, flush=True)
        line["fd_check"] = {"params": len(params), "stencil": "5-point, h = 1e-4 bohr / 1e-4*max(zeta,1)",
                            "max_abs_diff": float(max(errs)), "mean_abs_diff": float(np.mean(errs)),
                            "max_abs_diff_3point": float(max(errs3))}
    print(json.dumps(line), flush=True)


def bench_scf(args):
    """Full RHF on the GPU (core-Hamiltonian guess, DIIS, density and energy
    criteria), FP64 and with mixed precision, on one workload: iterations,
    wall time, time per iteration and the energy difference."""
    from librint_amd.context import FockContext
    m, tau = workload(args.workload)
    out = []
    for label, tau_mix in [("fp64", 0.0)] + [("mixed", t) for t in args.scf_tau_mix]:
        ctx = FockContext(m, schwarz_tol=tau)
        if tau_mix > 0.0:
            ctx.set_precision(1, tau_mix, 0.0)
        ctx.scf(imax=2, conv=1e-6, guess=1, diis=True)  # warm-up: lists, streams, code objects
        t0 = time.perf_counter()
        E, _, it = ctx.scf(imax=100, conv=1e-6, guess=1, diis=True, conv_e=1e-8)
        wall = time.perf_counter() - t0
        purified = int(ctx.stats()["scf_purified"])
        ctx.close()
        out.append({"precision": label, "tau_mix": tau_mix, "energy_Eh": E, "iterations": it, "wall_s": wall,
                    "ms_per_iteration": wall / max(it, 1) * 1e3, "purified_densities": purified})
    for r in out[1:]:
        r["energy_error_vs_fp64_gpu_Eh"] = r["energy_Eh"] - out[0]["energy_Eh"]
    line = {
        "metric": "RHF SCF on the GPU (core-H guess, DIIS, ||dP||/2 <= 1e-6 and |dE| <= 1e-8) s",
        "value": out[0]["wall_s"], "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": out[0]["wall_s"] * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic seeded geometry",
        "config": {"workload": args.workload, "nao": m.nao(), "schwarz": tau},
        "runs": out,
    }
    print(json.dumps(line), flush=True)


def timed_builds(ctx, Pd, Gd, steps, warmup):
    """Wall ms per Fock build over `steps` builds after `warmup`, bracketed by
    device synchronisation."""
    import torch
    for _ in range(warmup):
        ctx.fock_device(Pd, Gd)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.fock_device(Pd, Gd)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps * 1e3


def fock_line(name, steps, warmup, tau_mix_list, traffic_precision="fp64"):
    """FP64 Fock-build timing of one workload (core-Hamiltonian density),
    per-class device times from the library's HIP events, the roofline of the
    dominant launch and, for each tau_mix, the mixed-precision build timed the
    same way with its energy error against the FP64 GPU build.  Returns the
    result dict and (molecule, density, context, FP64 G, wall ms)."""
    import torch
    from librint_amd.context import FockContext
    m, tau = workload(name)
    P = core_guess_density(m)
    ctx = FockContext(m, schwarz_tol=tau)
    dev = torch.device("cuda:0")
    Pd = torch.from_numpy(np.ascontiguousarray(P)).to(dev)
    Gd = torch.empty_like(Pd)
    ms_per_step = timed_builds(ctx, Pd, Gd, steps, warmup)
    # device time of the J/K kernels (HIP events recorded by the library on
    # its streams around the class launches), more builds timed one by one
    ev_ms, cs_runs = [], []
    for _ in range(min(steps, 10)):
        ctx.fock(P)
        ev_ms.append(ctx.stats()["kernel_ms"])
        cs_runs.append(ctx.class_stats())
    st = ctx.stats()
    cstats = cs_runs[-1]
    for k, c in enumerate(cstats):
        c["ms"] = float(np.mean([r[k]["ms"] for r in cs_runs]))
        c["reduce_ms"] = float(np.mean([r[k]["reduce_ms"] for r in cs_runs]))
    classes = [{"class": "(%d%d|%d%d)" % (int(c["la"]), int(c["lb"]), int(c["lc"]), int(c["ld"])),
                "kernel": c["kernel"], "quartets": int(c["quartets"]), "prim_quartets": int(c["prim_quartets"]),
                "lane_util": c["quartets"] / max(c["lanes"], 1), "ms": c["ms"], "slab_reduction_ms": c["reduce_ms"],
                "tflops": c["flops"] / max(c["ms"], 1e-9) / 1e9} for c in cstats]
    jk_ms = float(np.mean(ev_ms))
    # dominant kernel: the class launch with the most algorithmic flops.  The
    # library runs it first and alone, so its event time is its own; the
    # other launches then overlap on several streams and their times include
    # the sharing.  A launch whose Coulomb slabs run in chunks reduces each
    # chunk right after it inside its events: the kernel's own time is the
    # launch time less those reductions (the library's events around them).
    dom = max(cstats, key=lambda c: c["flops"])
    dom_kernel_ms = dom["ms"] - dom["reduce_ms"]
    dom_achieved = dom["flops"] / (dom_kernel_ms * 1e-3) / 1e12
    achieved_all = st["flops"] / (jk_ms * 1e-3) / 1e12
    quartets = st["quartets"]
    line = {
        "value": quartets / (ms_per_step * 1e-3),
        "ms_per_step": ms_per_step,
        "fock_s_per_iter": ms_per_step * 1e-3,
        "integrals_per_s": st["integrals"] / (ms_per_step * 1e-3),
        "prim_quartets_per_s": st["prim_quartets"] / (ms_per_step * 1e-3),
        "config": {"workload": name, "nbas": m.nbas, "nao": m.nao(), "schwarz": tau,
                   "precision": "fp64", "quartets": int(quartets), "prim_quartets": int(st["prim_quartets"]),
                   "integrals": int(st["integrals"]), "candidates": int(st["candidates"]),
                   "launches": int(st["launches"]),
                   "streams": int(ctx.stats()["streams"])},
        "roofline": {"bound": "valu_fp64", "achieved": dom_achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": dom_achieved / FP64_PEAK_TFLOPS,
                     "traffic": TRAFFIC.get((name, traffic_precision, dom["kernel"])),
                     "kernel": dom["kernel"], "kernel_ms": dom_kernel_ms, "launch_ms": dom["ms"],
                     "slab_reduction_ms": dom["reduce_ms"], "flops_per_launch": dom["flops"],
                     "counter_flops_per_launch": COUNTER_FLOPS.get((name, traffic_precision, dom["kernel"])),
                     "frac_counter": (COUNTER_FLOPS[(name, traffic_precision, dom["kernel"])] / (dom_kernel_ms * 1e-3)
                                      / 1e12 / FP64_PEAK_TFLOPS
                                      if (name, traffic_precision, dom["kernel"]) in COUNTER_FLOPS else None),
                     "all_classes": {"achieved": achieved_all, "frac": achieved_all / FP64_PEAK_TFLOPS,
                                     "ms": jk_ms, "flops": st["flops"]}},
        "classes": classes,
    }
    G64 = ctx.fock(P)
    mixed = []
    for tau_mix in tau_mix_list:
        # FP32 recurrences for the 64-ket tasks whose largest
        # Q_ab Q_cd max|D_block| is below tau_mix, decided per build
        ctx.set_precision(1, tau_mix, 0.0)
        ms_mixed = timed_builds(ctx, Pd, Gd, steps, warmup)
        Gm = ctx.fock(P)
        fc = ctx.flag_counts()
        q64, q32 = fc["quartets"][:2]
        mixed.append({
            "tau_mix": tau_mix, "ms_per_step": ms_mixed, "fp64_ms_per_step": ms_per_step,
            "speedup": ms_per_step / ms_mixed,
            "energy_error_vs_fp64_gpu_Eh": float(0.5 * np.sum(P * (Gm - G64))),
            "max_abs_dG": float(np.abs(Gm - G64).max()),
            "quartets_fp64_fp32": [q64, q32], "fp32_quartet_share": q32 / max(q64 + q32, 1)})
    if tau_mix_list:
        ctx.set_precision(0, 0.0, 0.0)
        line["mixed_vs_fp64"] = mixed
    return line, (m, P, ctx, G64, ms_per_step)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="w32_et")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tau-mix-compare", type=float, nargs="*", default=[1e-8, 1e-7, 1e-6, 1e-5],
                    help="also time the mixed-precision build at these thresholds")
    ap.add_argument("--no-mixed-compare", action="store_true")
    ap.add_argument("--extra", nargs="*", default=["c16h34_631g"],
                    help="further workloads timed the same way (FP64 and mixed), reported under 'extra_workloads'")
    ap.add_argument("--shares", type=int, default=1,
                    help="time each device's share of a G-device split of the Fock build on this one GPU")
    ap.add_argument("--grad", action="store_true",
                    help="time the analytic RHF gradient (adjoint ERI kernel) instead of the Fock build")
    ap.add_argument("--copy-mb", type=float, default=0.0,
                    help="only copy this many MiB device to device --steps times (PMC calibration)")
    ap.add_argument("--scf", action="store_true",
                    help="time a full RHF (core-H guess, DIIS) in FP64 and mixed precision instead")
    ap.add_argument("--scf-tau-mix", type=float, nargs="*", default=[1e-5],
                    help="with --scf: tau_mix values of the mixed-precision runs")
    ap.add_argument("--fd", default="0",
                    help="with --grad: check 'all' or N seeded parameters against re-converged central FD")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the Fock build is sharded by quartet bins "
              "across GPUs in the design, not in this build); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from librint_amd.context import FockContext

    if args.grad:
        return bench_grad(args)
    if args.scf:
        return bench_scf(args)
    if args.copy_mb > 0:
        # device-to-device copy of a known size: the reference point for the
        # FETCH_SIZE / WRITE_SIZE counters of the roofline traffic figure
        n = int(args.copy_mb * 2 ** 20) // 8
        a = torch.arange(n, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(max(args.steps, 1)):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 8 * n, "copy_bytes_written": 8 * n, "copies": max(args.steps, 1)}))
        return
    if args.shares > 1:
        m, tau = workload(args.workload)
        P = core_guess_density(m)
        return bench_shares(args, m, P, FockContext(m, schwarz_tol=tau))
    taus = [] if args.no_mixed_compare else args.tau_mix_compare
    body, (m, P, ctx, G64, ms_per_step) = fock_line(args.workload, args.steps, args.warmup, taus)
    quartets = body["config"]["quartets"]
    line = {
        "metric": "RHF Fock-build s/iter + ERI quartets/s, mixed vs FP64 at <=1 uEh error",
        "value": body.pop("value"),
        "unit": "quartets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": body.pop("ms_per_step"),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic seeded geometry, core-Hamiltonian density; value = FP64 fused J/K build, "
                "mixed_vs_fp64 = the same build with FP32 recurrences below tau_mix",
    }
    line.update(body)
    ctx.close()
    if not args.no_cpu:
        rate, t_full, nq, used = cpu_baseline(m, P, args.cpu_seconds)
        line["cpu_baseline"] = {
            "value": rate, "unit": "quartets/s", "cores": 1, "kind": "port",
            "sample": "%d random ordered shell quartets of the nbas^4 reference loop (%.1f s, oracle "
                      "integral2e_fock restatement: no symmetry, no screening, FP64, 1 thread)" % (nq, used),
            "fock_s_per_iter_extrapolated": t_full,
        }
        stride = max(16, int(quartets / 2e6))
        trate, nth, tq, tused, phases = cpu_threaded(m, P, line["config"]["schwarz"], args.cpu_seconds, stride)
        line["cpu_threaded"] = {
            "value": trate, "unit": "quartets/s", "cores": nth, "kind": "port",
            "cpu": cpu_model(),
            "sample": "%d of the %d Schwarz-surviving canonical quartets (bra pairs ij %% %d in 0..%d, %.1f s; "
                      "oracle screened 8-fold-symmetric Fock build, OpenMP)" % (tq, quartets, stride, phases - 1,
                                                                                 tused),
            "fock_s_per_iter_extrapolated": quartets / trate,
            "gpu_speedup_fock": quartets / trate / (line["ms_per_step"] * 1e-3),
        }
    extra = {}
    for name in args.extra:
        if name == args.workload:
            continue
        b, (_, _, c2, _, _) = fock_line(name, args.steps, args.warmup, [] if args.no_mixed_compare else
                                        [1e-8, 1e-6, 1e-5])
        c2.close()
        b.pop("classes")
        extra[name] = b
    if extra:
        line["extra_workloads"] = extra
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
