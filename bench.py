"""Benchmark: psi(s) throughput of the 8x8 FermiU1 PEPS (D=8 as {Q-1:3, Q:3,
Q+1:2}, boundary-MPS cut Dc=64) at half filling (N=32), batched engine on one
MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]

A step evaluates one batch of B configurations (drawn uniformly at half
filling, fresh per step) with inputs already resident in HBM.  The line
printed reports configurations/sec over the K timed steps, the roofline of
the dominant kernel measured with HIP events on the launch stream during the
timed steps, and the reference TAT (header-only C++ built from the reference
sources into oracle/_ref) timed on this host's cores on a bounded sample of
the same workload.

Kernel durations come from a separate pass outside the timed region: one
chunk of the last timed batch on one lane with HIP events on the launch
stream (events under concurrent lanes would also time the queueing behind
the other lanes' kernels), so they agree with a rocprofv3 kernel trace of
a one-lane run.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"hbm": 8.0e12, "fp64": 78.6e12, "fp32": 157.3e12}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--chunk", type=int, default=128, help="configurations per lockstep pass of one lane")
    ap.add_argument("--timing-chunk", type=int, default=256,
                    help="configurations per launch of the one-lane kernel timing pass after the timed steps")
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--lanes", type=int, default=16, help="concurrent lockstep lanes (host thread + HIP stream each)")
    ap.add_argument("--hw-queues", type=int, default=16,
                    help="GPU_MAX_HW_QUEUES for this process (0: leave the environment's); with fewer hardware "
                         "queues than lanes the lanes' streams share queues and a long kernel of one lane holds "
                         "back the launches queued behind it")
    ap.add_argument("--sampler-lanes", type=int, default=4,
                    help="lanes of the sampling legs (sweep, ESS, direct, gradient descent): their batched ops "
                         "split the chains, so fewer, larger lanes do better there")
    ap.add_argument("--workload", default="C4", choices=["C2", "C3", "C4"])
    ap.add_argument("--cpu-configs", type=int, default=4, help="configs per host process for the CPU baseline")
    ap.add_argument("--cpu-procs", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain", type=int, default=4,
                    help="also time N configs of the last timed batch on the plain per-config path and compare "
                         "their psi with the timed engine's")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32", "bf16x3", "bf16x6"],
                    help="contraction GEMM precision of the timed engine")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default GEMM tile per size class")
    ap.add_argument("--traffic-calibration", action="store_true",
                    help="first copy 1 GiB device to device (the reference kernel of known traffic that scales "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    ap.add_argument("--ess", default="", help="also measure ESS/sec of plain vs learned sampling on this workload")
    ap.add_argument("--ess-chains", type=int, default=256)
    ap.add_argument("--ess-diagonal", default="keep", choices=["keep", "hole"],
                    help="doubly occupied bonds: keep psi, or refresh it with hole(()) as tetragono does")
    ap.add_argument("--ess-engine-sweep", action="store_true",
                    help="also measure the sweep that recomputes psi(s') with the amplitude engine")
    ap.add_argument("--ess-samples", type=int, default=16)
    ap.add_argument("--ess-default", default="C4",
                    help="workload of the bounded ESS leg of the default run (learned proposal vs sweep)")
    ap.add_argument("--no-ess", action="store_true", help="skip the ESS leg of the default run")
    ap.add_argument("--ess-train-sweeps", type=int, default=6)
    ap.add_argument("--ess-epochs", type=int, default=6)
    ap.add_argument("--ess-learned-k", default="",
                    help="proposals per learned sample, comma-separated (default: bonds / 2)")
    ap.add_argument("--sweep", default="C4", help="also time tetragono's sweep (cached-environment replace) and "
                                                "its observer energy on this workload")
    ap.add_argument("--sweep-only", action="store_true", help="time only the sweep and observer (one JSON line)")
    ap.add_argument("--sweep-chains", type=int, default=256)
    ap.add_argument("--sweep-diagonal", default="hole", choices=["keep", "hole"],
                    help="doubly occupied bonds in the sweep timing: tetragono's hole(()) refresh (the reference "
                         "chain) or keep psi")
    ap.add_argument("--sweep-count", type=int, default=1, help="timed sweeps after one untimed sweep")
    ap.add_argument("--no-sweep", action="store_true", help="skip the sweep timing of the default run")
    ap.add_argument("--vmc", default="", help="run tetragono-style gradient descent on this workload (one JSON line)")
    ap.add_argument("--vmc-steps", type=int, default=10)
    ap.add_argument("--vmc-chains", type=int, default=256)
    ap.add_argument("--vmc-sweeps", type=int, default=1, help="sweeps (samples per chain) per gradient step")
    ap.add_argument("--vmc-step-size", type=float, default=0.01)
    ap.add_argument("--vmc-natural", action="store_true", help="natural gradient by conjugate gradient")
    ap.add_argument("--vmc-direct", type=int, default=0, help="direct samples per gradient step (0: sweep sampling)")
    ap.add_argument("--direct", default="", help="time tetragono's direct sampling on this workload (one JSON line)")
    ap.add_argument("--direct-samples", type=int, default=256)
    ap.add_argument("--direct-default", default="C2", help="workload of the direct-sampling timing of the default run")
    ap.add_argument("--no-direct", action="store_true", help="skip the direct-sampling timing of the default run")
    ap.add_argument("--direct-cut2", type=int, default=4, help="double-layer bond dimension of direct sampling")
    return ap.parse_args()


def vmc_run(args, log):
    """Gradient descent with sweep sampling (tnsp_amd.vmc), inputs resident:
    per step the sweeps, the observer's energy and holes, the gradient and
    the update are timed together."""
    import numpy as np

    from tnsp_amd import peps
    from tnsp_amd.engine import AmplitudeEngine
    from tnsp_amd.environment import EnvironmentSampler
    from tnsp_amd.vmc import gradient_descent

    spec = getattr(peps, args.vmc)
    lat = peps.Lattice(spec)
    eng = AmplitudeEngine(lat, chunk=256, lanes=args.sampler_lanes)
    cand = peps.sample_configurations(spec, 4 * args.vmc_chains, seed=77)
    start = cand[eng.evaluate(cand) != 0][:args.vmc_chains]
    smp = EnvironmentSampler(eng, start, seed=1, diagonal=args.ess_diagonal, lanes=args.sampler_lanes)
    direct = None
    if args.vmc_direct:
        from tnsp_amd.direct import DirectSampler

        direct = DirectSampler(lat, cut2=args.direct_cut2, seed=2)
    recs = gradient_descent(smp, args.vmc_steps, args.vmc_step_size, sweeps=args.vmc_sweeps,
                            natural=args.vmc_natural, log=log, direct=direct, direct_samples=args.vmc_direct)
    return {"workload": args.vmc, "chains": len(start), "lanes": args.sampler_lanes, "natural": args.vmc_natural,
            "step_size": args.vmc_step_size, "sweeps_per_step": args.vmc_sweeps,
            "direct_samples_per_step": args.vmc_direct, "direct_cut2": args.direct_cut2 if direct else None,
            "instability": [r.get("instability") for r in recs],
            "energy_per_site": [r["energy_per_site"] for r in recs],
            "seconds_per_step": float(np.mean([r["seconds"] for r in recs[1:]] or [recs[0]["seconds"]]))}


def direct_timing(args, log):
    """Time DirectSampler draws (tetragono's DirectSampling, the samples in
    lockstep) on ``args.direct``, inputs resident: the double layer once
    (reported apart), one untimed draw of a few samples, then one timed draw
    of ``args.direct_samples``; the weights |psi|^2 / p use the engine's psi."""
    import numpy as np
    import torch

    from tnsp_amd import peps
    from tnsp_amd.direct import DirectSampler, instability, reweight
    from tnsp_amd.engine import AmplitudeEngine

    spec = getattr(peps, args.direct)
    lat = peps.Lattice(spec)
    torch.cuda.synchronize()
    t0 = time.time()
    dsm = DirectSampler(lat, cut2=args.direct_cut2, seed=1)
    torch.cuda.synchronize()
    t_double = time.time() - t0
    log(f"direct: double layer built in {t_double:.2f}s")
    dsm.sample(min(8, args.direct_samples), log=lambda m: log(f"direct (warm-up): {m}"))
    torch.cuda.synchronize()
    t0 = time.time()
    poss, cfg = dsm.sample(args.direct_samples, log=lambda m: log(f"direct: {m}"))
    torch.cuda.synchronize()
    dt = time.time() - t0
    eng = AmplitudeEngine(lat, chunk=256, lanes=args.sampler_lanes)
    w = reweight(eng.evaluate(cfg), poss)
    return {"workload": args.direct, "samples": args.direct_samples, "cut": spec.cut, "cut2": args.direct_cut2,
            "double_layer_s": t_double, "draw_s": dt, "samples_per_s": args.direct_samples / dt,
            "redrawn": dsm.redrawn, "zero_psi": int(np.sum(w == 0)), "instability": float(instability(w))}


def sweep_timing(args, log, eng=None):
    """Time EnvironmentSampler sweeps (tetragono's SweepSampling with
    replace over cached environments) and one observer pass over all
    chains, inputs resident, on ``args.sweep`` (reusing ``eng`` when it
    holds that workload's PEPS)."""
    import numpy as np
    import torch

    from tnsp_amd import peps
    from tnsp_amd.engine import AmplitudeEngine
    from tnsp_amd.environment import EnvironmentSampler

    spec = getattr(peps, args.sweep)
    if eng is None:
        eng = AmplitudeEngine(peps.Lattice(spec), chunk=256, lanes=args.sampler_lanes)
    cand = peps.sample_configurations(spec, 4 * args.sweep_chains, seed=77)
    start = cand[eng.evaluate(cand) != 0][:args.sweep_chains]
    eng.trim()
    smp = EnvironmentSampler(eng, start, seed=1, diagonal=args.sweep_diagonal, lanes=args.sampler_lanes)
    t0 = time.time()
    smp.sweep(log=lambda m: log(f"sweep (warm-up): {m} ({time.time() - t0:.1f} s)"))
    log("sweep: warm-up sweep done")
    torch.cuda.synchronize()
    t0 = time.time()
    for k in range(args.sweep_count):
        smp.sweep(log=lambda m: log(f"sweep {k + 1}: {m} ({time.time() - t0:.1f} s)"))
        log(f"sweep: {k + 1}/{args.sweep_count} timed sweeps")
    torch.cuda.synchronize()
    dt = (time.time() - t0) / args.sweep_count
    t0 = time.time()
    e = smp.local_energy()
    torch.cuda.synchronize()
    de = time.time() - t0
    return {"workload": args.sweep, "chains": len(start), "lanes": args.sampler_lanes, "diagonal": args.sweep_diagonal,
            # "keep" skips tetragono's hole(()) refresh of doubly occupied
            # bonds: a cheaper chain than the reference's at a truncating cut
            "reference_chain": args.sweep_diagonal == "hole",
            "sweep_s": dt, "chain_sweeps_per_s": len(start) / dt, "bonds": len(smp.bonds),
            "proposals_per_s": smp.proposals / (args.sweep_count + 1) / dt,
            "acceptance": smp.accepted / max(1, smp.proposals), "hole_refreshes": smp.refreshed,
            "respun_chains": smp.respun, "observer_s": de,
            "energy_per_site": float(np.nanmean(e) / (spec.L1 * spec.L2))}


def measured_traffic(kernel, args):
    """HBM bytes per launch of ``kernel`` from the committed FETCH_SIZE /
    WRITE_SIZE profile of the same workload and launch size (rocprofv3 --pmc
    runs scaled by the counters' reading for a device copy of known size)."""
    path = os.path.join(ROOT, "profiles", "c4_traffic.json")
    if args.workload != "C4" or args.timing_chunk != 256 or args.precision != "fp64" or not os.path.exists(path):
        return None
    with open(path) as f:
        k = json.load(f)["kernels"].get(kernel)
    return k["traffic_bytes"] if k else None


def host_cores():
    """CPU threads this process may use: the cgroup CPU quota when one is set
    (a GPU box shares its host), else the affinity mask; and nproc."""
    nproc = os.cpu_count() or 1
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = nproc
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            quota, period = f.read().split()[:2]
        if quota != "max":
            usable = min(usable, max(1, int(int(quota) / int(period))))
    except (OSError, ValueError):
        pass
    return usable, nproc


def kernel_rates(kstats):
    """Achieved rate of each kernel kind over the launches of the kernel
    timing pass (HIP events on the launch stream): flops for GEMM / QR / SVD
    against the fp64 matrix peak, bytes for permute / fill / diag against
    HBM."""
    out = {}
    for kind, s in kstats.items():
        if s["launches"] == 0 or s["ms"] <= 0:
            continue
        rate = s["units"] / (s["ms"] / 1e3)
        if kind in ("permute", "fill", "diag"):
            out[kind] = {"GB/s": rate / 1e9, "frac_hbm": rate / PEAK["hbm"]}
        else:
            out[kind] = {"GFLOP/s": rate / 1e9, "frac_fp64": rate / PEAK["fp64"]}
    return out


def psi_gate(psi_engine, psi_plain, bound=1e-6):
    """|d log psi| and sign agreement of the timed engine against the plain
    fp64 path on the same configurations (zeros must agree exactly)."""
    import numpy as np

    a = np.asarray(psi_engine, dtype=np.float64)
    b = np.asarray(psi_plain, dtype=np.float64)
    both = (a != 0) & (b != 0)
    zeros_agree = bool(np.all((a == 0) == (b == 0)))
    dl = np.abs(np.log(np.abs(a[both])) - np.log(np.abs(b[both]))) if both.any() else np.zeros(0)
    sign = float(np.mean(np.sign(a[both]) == np.sign(b[both]))) if both.any() else 1.0
    worst = float(dl.max()) if dl.size else 0.0
    return {"configs": int(a.size), "nonzero": int(both.sum()), "max_abs_dlogpsi": worst,
            "sign_agreement": sign, "zeros_agree": zeros_agree, "bound": bound,
            "gate_met": bool(worst <= bound and sign == 1.0 and zeros_agree)}


def dominant_kernel(kstats):
    best = max(kstats.items(), key=lambda kv: kv[1]["ms"])
    kind, s = best
    if s["launches"] == 0:
        return None
    avg_s = s["ms"] / s["launches"] / 1e3
    per_launch = s["units"] / s["launches"]
    if kind in ("permute", "fill", "diag"):
        bound, unit, peak, scale = "hbm", "GB/s", PEAK["hbm"], 1e9
    else:
        bound, unit, peak, scale = "mfma", "TFLOP/s", PEAK["fp64"], 1e12
    achieved = per_launch / avg_s
    return kind, {"bound": bound, "achieved": achieved / scale, "peak": peak / scale, "unit": unit,
                  "frac": achieved / peak, "traffic": None, "kernel": kind, "launches": int(s["launches"]),
                  "avg_launch_ms": avg_s * 1e3, "per_launch": per_launch}


def cpu_baseline(spec, storages, configs, per_proc, procs):
    """Reference TAT (oracle/_ref/tat_interp) on this host: ``procs``
    single-threaded processes, each evaluating ``per_proc`` configurations
    (projection + amplitude in tetragono's order); aggregate configs/sec."""
    from oracle import peps as opeps
    from oracle import ref_runner as rr

    if not rr.available():
        return None
    sd = {"kind": spec.kind, "L1": spec.L1, "L2": spec.L2, "N": spec.N, "segments": list(spec.segments)}
    sites = opeps.build_sites(sd, storages)
    lines = [f"sym {spec.kind}"]
    for l1 in range(spec.L1):
        for l2 in range(spec.L2):
            lines += rr.tensor_lines(f"s_{l1}_{l2}", sites[l1][l2])
    head = "\n".join(lines)
    scripts = []
    for p in range(procs):
        occ = configs[p * per_proc:(p + 1) * per_proc].reshape(-1)
        scripts.append(head + f"\ntime_raw {spec.L1} {spec.L2} {spec.cut} s {per_proc} " + " ".join(map(str, occ)) + "\n")
    env = dict(os.environ, OMP_NUM_THREADS="1", MKL_NUM_THREADS="1")
    t0 = time.time()
    ps = [subprocess.Popen([rr.BINARY], stdin=subprocess.PIPE, stdout=subprocess.PIPE, stderr=subprocess.PIPE, env=env,
                           text=True) for _ in range(procs)]
    for p, s in zip(ps, scripts):
        p.stdin.write(s)
        p.stdin.close()
    inner = []
    for p in ps:
        out = p.stdout.read()
        p.wait()
        for line in out.splitlines():
            if line.startswith("time"):
                inner.append(float(line.split()[1]))
    wall = time.time() - t0
    if len(inner) != procs:
        return None
    total = per_proc * procs
    # each process's own timed loop excludes script parsing
    rate = total / max(inner)
    return {"value": rate, "unit": "configs/s", "cores": procs, "kind": "reference",
            "sample": f"{total} configurations ({per_proc} per process x {procs} single-threaded processes) of the "
                      f"same workload; reference TAT built from the reference sources, projection + amplitude timed "
                      f"in-process, max over processes; wall incl. parsing {wall:.1f}s"}


def ess_comparison(args, log):
    """Sweep sampling vs learned-proposal Metropolis-Hastings on the tV model
    (t = V = 1): ESS of the energy estimator per second of sampling.  The
    sweep is tetragono's: cached-environment ``replace`` for every bond
    (EnvironmentSampler; ``--ess-diagonal keep`` skips the hole(()) refresh of
    doubly occupied bonds); ``--ess-engine-sweep`` also times the sweep that
    recomputes psi(s') with the amplitude engine.  The learned proposal is
    trained first on hops the sweep tried; one learned sample is K proposals
    per chain.  Local energies (not timed) are tetragono's observer: every
    term's psi(s') by replace over cached environments."""
    import numpy as np

    from tnsp_amd import peps
    from tnsp_amd.engine import AmplitudeEngine
    from tnsp_amd.environment import EnvironmentSampler
    from tnsp_amd.proposal import HopProposalNet, LearnedSampler, collect_training_data, train
    from tnsp_amd.sampler import TVSampler, measure_ess

    spec = getattr(peps, args.ess)
    lat = peps.Lattice(spec)
    # 256 configurations per lane chunk: at C4 the arenas of one chunk hold
    # about 30 GB, so four lanes of 1024 would not fit in HBM
    eng = AmplitudeEngine(lat, chunk=256, lanes=args.sampler_lanes)
    L = spec.L1 * spec.L2
    cand = peps.sample_configurations(spec, 4 * args.ess_chains, seed=77)
    psi = eng.evaluate(cand)
    start = cand[psi != 0][:args.ess_chains]
    sweep = EnvironmentSampler(eng, start, seed=1, diagonal=args.ess_diagonal, lanes=args.sampler_lanes)

    def energy(configs):
        # tetragono's observer: psi(s') by replace over cached environments
        return sweep.local_energy(configs)

    bonds_a = np.array([a[0] * spec.L2 + a[1] for a, _ in sweep.bonds])
    bonds_b = np.array([b[0] * spec.L2 + b[1] for _, b in sweep.bonds])

    def density(configs):
        # the interaction energy V sum n_i n_j: bounded, a second observable
        flat = np.asarray(configs).reshape(len(configs), -1).astype(np.int64)
        return (flat[:, bonds_a] * flat[:, bonds_b]).sum(1).astype(np.float64)

    t0 = time.time()
    occs, hops, targets = collect_training_data(sweep, args.ess_train_sweeps, log=log)
    t_collect = time.time() - t0
    net = HopProposalNet(spec.L1, spec.L2, channels=32, layers=4, seed=3)
    t0 = time.time()
    losses = train(net, occs, hops, targets, epochs=args.ess_epochs, batch=1024, lr=2e-3, log=log)
    t_train = time.time() - t0
    res = {}
    res_sweep = measure_ess(sweep.sweep, lambda: energy(sweep.configs), args.ess_samples, sites=L,
                            log=lambda m: log(f"environment sweep: {m}"), extra=lambda: density(sweep.configs))
    res_sweep["acceptance"] = sweep.accepted / max(1, sweep.proposals)
    res_sweep["diagonal"] = args.ess_diagonal
    res_sweep["reference_chain"] = args.ess_diagonal == "hole"
    res["sweep"] = res_sweep
    if args.ess_engine_sweep:
        plain = TVSampler(eng, sweep.configs.copy(), seed=1)
        r = measure_ess(plain.sweep, lambda: energy(plain.configs), args.ess_samples, sites=L,
                        log=lambda m: log(f"engine sweep: {m}"), extra=lambda: density(plain.configs))
        r["acceptance"] = plain.accepted / max(1, plain.proposals)
        res["engine_sweep"] = r
    ks = [int(x) for x in args.ess_learned_k.split(",") if x] or [max(1, len(sweep.bonds) // 2)]
    res_learned = []
    # the learned chains continue from where the sweep left them; a chain the
    # sweep left at hole(()) = 0 (every sector truncated away) restarts from
    # its initial configuration
    lstart = sweep.configs.copy()
    zero = eng.evaluate(lstart) == 0
    lstart[zero] = start[zero]
    for K in ks:
        learned = LearnedSampler(eng, net, lstart.copy(), seed=2)
        r = measure_ess(lambda: [learned.step() for _ in range(K)],
                        lambda: energy(learned.configs), args.ess_samples, sites=L,
                        log=lambda m: log(f"learned K={K}: {m}"), extra=lambda: density(learned.configs))
        r["acceptance"] = learned.accepted / max(1, learned.proposals)
        r["proposals_per_sample"] = K
        res_learned.append(r)
    best = max(res_learned, key=lambda r: r["ess_per_s"])
    best_x = max(res_learned, key=lambda r: r["ess_extra_per_s"])
    out = {"workload": args.ess, "chains": len(start), "learned": res_learned,
           "gain": best["ess_per_s"] / max(res_sweep["ess_per_s"], 1e-12),
           "gain_density": best_x["ess_extra_per_s"] / max(res_sweep["ess_extra_per_s"], 1e-12),
           "training": {"samples": int(len(hops)), "collect_s": t_collect, "train_s": t_train,
                        "loss_first": losses[0] if losses else None, "loss_last": losses[-1] if losses else None}}
    out.update(res)
    if "engine_sweep" in res:
        out["gain_vs_engine_sweep"] = best["ess_per_s"] / max(res["engine_sweep"]["ess_per_s"], 1e-12)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.hw_queues:
        if not 1 <= args.hw_queues <= 32:
            sys.exit("--hw-queues must be in [1, 32]")
        # read by the HIP runtime when it initialises, which nothing has done yet
        os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)
    import numpy as np
    import torch

    from tnsp_amd import peps
    from tnsp_amd.engine import AmplitudeEngine

    torch.cuda.set_device(0)

    def log(m):
        print(f"[bench] {m}", file=sys.stderr, flush=True)

    if args.sweep_only:
        args.sweep = args.sweep or args.workload
        print(json.dumps({"environment_sweep": sweep_timing(args, log)}), flush=True)
        return
    if args.vmc:
        print(json.dumps({"vmc": vmc_run(args, log)}), flush=True)
        return
    if args.direct:
        print(json.dumps({"direct_sampling": direct_timing(args, log)}), flush=True)
        return
    if args.traffic_calibration:
        src = torch.empty(2**27, dtype=torch.float64, device="cuda")
        torch.empty_like(src).copy_(src)
        torch.cuda.synchronize()
        del src
        log("traffic calibration: 1 GiB device-to-device copy done")
    spec = getattr(peps, args.workload)
    storages = peps.random_storages(spec)
    lat = peps.Lattice(spec, storages)
    threads = args.threads or int(os.environ.get("TNS_THREADS", min(16, os.cpu_count() or 1)))
    eng = AmplitudeEngine(lat, threads=threads, chunk=args.chunk, lanes=args.lanes, precision=args.precision)
    B = args.batch
    total_steps = args.warmup + args.steps
    particles = (24, 40) if spec.kind == "FermiZ2" else None
    configs = peps.sample_configurations(spec, B * total_steps, seed=1234, particles=particles)
    print(f"[bench] workload {args.workload} batch {B} threads {threads} lanes {args.lanes}", file=sys.stderr,
          flush=True)
    tuned = None
    if not args.no_autotune and args.precision in ("fp64", "fp32"):
        t0 = time.time()
        table, ms = eng.autotune(configs[:min(B, args.chunk)])
        tuned = {"table": table, "ms": ms.round(3).tolist()}
        print(f"[bench] GEMM tiles per class (<=16, <=48, larger): {table} ({time.time() - t0:.1f}s)",
              file=sys.stderr, flush=True)
    for w in range(args.warmup):
        t0 = time.time()
        eng.evaluate(configs[w * B:(w + 1) * B])
        print(f"[bench] warmup step {w} {time.time() - t0:.2f}s", file=sys.stderr, flush=True)
    eng.reset_stats()
    torch.cuda.synchronize()
    t0 = time.time()
    psis = []
    for s in range(args.steps):
        i0 = (args.warmup + s) * B
        psis.append(eng.evaluate(configs[i0:i0 + B]))
        print(f"[bench] step {s} done {time.time() - t0:.2f}s", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.time() - t0
    psi = np.concatenate(psis)
    st = eng.stats()
    value = B * args.steps / elapsed
    # kernel durations: one chunk of --timing-chunk configurations on one
    # lane, outside the timed region
    i0 = (args.warmup + args.steps - 1) * B
    n_timing = min(B, args.timing_chunk)
    eng.set_lanes(1)
    eng.set_chunk(n_timing)
    eng.reset_stats()
    eng.set_profile(True)
    eng.evaluate(configs[i0:i0 + n_timing])
    torch.cuda.synchronize()
    eng.set_profile(False)
    ks = eng.kernel_stats()
    eng.set_chunk(args.chunk)
    eng.set_lanes(args.lanes)
    # the legs below run their own samplers and engines: give this engine's
    # working memory back first
    eng.trim()
    print(f"[bench] kernel timing pass: {n_timing} configurations on one lane", file=sys.stderr, flush=True)
    dom = dominant_kernel(ks)
    roof = dom[1] if dom else None
    if roof:
        roof["traffic"] = measured_traffic(dom[0], args)
        roof["traffic_source"] = ("committed rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE profile of this workload "
                                  "(profiles/c4_traffic.json), not measured in this run"
                                  if roof["traffic"] is not None else None)
    plain = None
    gate = None
    if args.plain:
        # the plain path: the drop-in Tensor API, one configuration and one
        # op at a time (tetragono's own loop on the GPU), on configurations of
        # the last timed batch; its psi is the fp64 reference of the gate
        peps.amplitude_plain(lat, configs[0])
        torch.cuda.synchronize()
        i0 = (args.warmup + args.steps - 1) * B
        sel = configs[i0:i0 + args.plain]
        t1 = time.time()
        plain_psi = [peps.amplitude_plain(lat, c)[0] for c in sel]
        torch.cuda.synchronize()
        plain = len(sel) / (time.time() - t1)
        gate = psi_gate(psis[-1][:len(sel)], plain_psi)
        print(f"[bench] plain per-configuration path {plain:.2f} configs/s; gate {gate}", file=sys.stderr,
              flush=True)
    cpu = None
    if not args.no_cpu:
        usable, nproc = host_cores()
        procs = args.cpu_procs or usable
        cpu = cpu_baseline(spec, storages, configs, args.cpu_configs, procs)
        if cpu:
            cpu["nproc"] = nproc
            cpu["sample"] += (f"; {procs} processes = the CPU threads this process may use (cgroup quota / "
                              f"affinity), nproc reports {nproc}")
    nonzero = float(np.mean(psi != 0))
    ess = None
    if args.ess:
        ess = ess_comparison(args, log)
    env_sweep = None
    if args.sweep and not args.no_sweep:
        # outside the timed region; reported, never fatal to the bench line
        try:
            env_sweep = sweep_timing(args, log, eng if args.sweep == args.workload and args.precision == "fp64"
                                     else None)
        except Exception as e:  # noqa: BLE001
            env_sweep = {"error": f"{type(e).__name__}: {e}"}
    ess_leg = None
    if not args.ess and not args.no_ess and args.ess_default:
        # bounded ESS leg outside the timed region: 128 chains, 8 samples
        # after 2 burn-in, learned proposal trained on 2 sweeps' hops, K = 4
        # proposals per learned sample
        try:
            sub = argparse.Namespace(**vars(args))
            sub.ess, sub.ess_chains, sub.ess_samples = args.ess_default, 128, 8
            sub.ess_train_sweeps, sub.ess_epochs, sub.ess_learned_k = 2, 4, "4"
            sub.ess_engine_sweep = False
            t0 = time.time()
            ess_leg = ess_comparison(sub, log)
            ess_leg["wall_s"] = time.time() - t0
        except Exception as e:  # noqa: BLE001
            ess_leg = {"error": f"{type(e).__name__}: {e}"}
    direct = None
    if not args.no_direct:
        # direct sampling on the 4x4 workload, also outside the timed region
        try:
            args.direct = args.direct_default
            direct = direct_timing(args, log)
        except Exception as e:  # noqa: BLE001
            direct = {"error": f"{type(e).__name__}: {e}"}
    if args.workload == "C4" and args.precision == "fp64":
        metric = "configs/sec 8x8 D=8 Dc=64 at |dlogpsi|<=1e-6 vs plain FP64; ESS/sec"
    else:
        metric = (f"configs/sec {spec.L1}x{spec.L2} {spec.kind} Dc={spec.cut}, contraction GEMMs "
                  f"{args.precision} (|dlogpsi| vs plain FP64 in parity_gate)")
    line = {
        "metric": metric,
        "value": value,
        "unit": "configs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else f"f64 (contraction GEMMs {args.precision})",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {spec.kind} {spec.L1}x{spec.L2} PEPS, bond segments "
                               f"{list(spec.segments)} around tV charge backgrounds, Dc={spec.cut}, N={spec.N}, "
                               f"{B} configurations per step drawn uniformly at fixed particle number, "
                               f"batched lockstep engine, contraction GEMMs {args.precision}, QR/SVD fp64",
                   "batch": B, "threads": threads, "lanes": args.lanes,
                   "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))},
        "roofline": roof,
        "cpu_baseline": cpu,
        "parity_gate": gate,
        "ess_per_s": None if not (ess or ess_leg) or "error" in (ess or ess_leg) else {
            "workload": (ess or ess_leg)["workload"], "chains": (ess or ess_leg)["chains"],
            "sweep": (ess or ess_leg)["sweep"]["ess_per_s"],
            "learned": max(r["ess_per_s"] for r in (ess or ess_leg)["learned"]),
            "gain": (ess or ess_leg)["gain"], "gain_second_observable": (ess or ess_leg)["gain_density"],
            "reference_chain": (ess or ess_leg)["sweep"]["reference_chain"],
            "note": "ESS of the tV energy estimator per second of sampling (t = V = 1); the sweep is the "
                    "cached-environment replace sweep" + (
                        ", tetragono's chain (hole(()) refresh of doubly occupied bonds)"
                        if (ess or ess_leg)["sweep"]["reference_chain"] else
                        "; reference_chain false: doubly occupied bonds keep psi instead of tetragono's "
                        "hole(()) refresh (a cheaper, different chain)")},
        "extra": {"kernel_timing": (f"one lane, {n_timing} configurations per launch (the timed steps launch "
                                    f"{min(args.chunk, B)} per lane on {args.lanes} lanes), HIP events on the "
                                    "launch stream"),
                  "kernel_ms": {k: v["ms"] for k, v in ks.items()},
                  "kernel_rates": kernel_rates(ks),
                  "kernel_launches": {k: v["launches"] for k, v in ks.items()},
                  "host_plan_s": st["plan_s"], "host_bind_s": st["bind_s"], "svals_wait_s": st["sync_s"],
                  "engine_total_s": st["total_s"],
                  "gemm_gflop": st["gemm_flops"] / 1e9, "qr_gflop": st["qr_flops"] / 1e9,
                  "svd_gflop": st["svd_flops"] / 1e9, "perm_gbytes": st["perm_elems"] * 16 / 1e9,
                  "device_bytes": st["device_bytes"], "nonzero_fraction": nonzero,
                  "plain_path_configs_per_s": plain,
                  "vs_plain_path": (value / plain) if plain else None, "precision": args.precision, "gemm_tiles": tuned,
                  "ess": ess or ess_leg,
                  "environment_sweep": env_sweep, "direct_sampling": direct},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
