"""bench.py -- end-to-end SKID group finding on one MI355X.

A step is one full pass of the group-finding path over particles already
resident in HBM: scatterer tree + kNN density, the Denmax move loop to
convergence, FOF, micro-steps, group centres, unbinding and kdTooSmall,
ending with the group ids in host memory (main.c:374-471).

Workload: the two-species box (2^24 gas + 2^24 dark particles, 8000 NFW halos
with dN/dM ~ M^-1.9 on [1e-7, 1e-3], gas in cored isothermal profiles),
nSmooth 64, nBucket 16, tau 0.001, -gd, periodic unit box.  The CPU baseline
is the reference SKID (oracle/_ref) timed on this host's cores on a bounded
sample of the same recipe (see --cpu-sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(n_each: int, tau: float):
    """The reference (or, without it, the oracle's restatement) on one core."""
    from oracle import oracle as O
    from skid_amd import gen
    box = gen.config3(n_each=n_each, seed=3, n_halos=max(1, int(8000 * n_each / (1 << 24))))
    kw = dict(nSmooth=64, fTau=tau, fPeriod=1.0, bGasAndDark=1)
    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
    except Exception:
        pass
    t0 = time.perf_counter()
    if O.have_ref():
        O.ref_pipeline(box, **kw)
        kind = "reference"
    else:
        o = O.Oracle(box, **kw)
        o.run()
        kind = "port"
    dt = time.perf_counter() - t0
    return dict(value=box.n / dt, unit="particles/s", cores=1, kind=kind,
                sample=f"same recipe at {n_each}+{n_each} particles ({box.n}), tau scaled by N^-1/3 "
                       f"to {tau:.4g}, nSmooth 64, -gd; end-to-end {dt:.2f}s on 1 core "
                       f"(per-particle cost grows with N, so this overstates the CPU rate at 2^25)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--log2-each", type=int, default=24, help="particles per species = 2^this")
    ap.add_argument("--plain", action="store_true", help="fp32 running sums (reference arithmetic)")
    ap.add_argument("--cpu-sample", type=int, default=16, help="log2 particles per species of the CPU sample; 0 skips")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    from skid_amd import gen, skg
    n_each = 1 << args.log2_each
    tau = 0.001 * (2.0 ** ((24 - args.log2_each) / 3.0))
    t = time.time()
    box = gen.config3(n_each=n_each)
    log(f"generated {box.n} particles in {time.time() - t:.1f}s")
    P = skg.make_params(tau, nSmooth=64, nBucket=16, period=1.0, bGasAndDark=1,
                        accumulate=skg.SKG_ACC_PLAIN if args.plain else skg.SKG_ACC_COMPENSATED)
    ctx = skg.Context(box, P)                       # particles resident in HBM from here on

    def one_pass():
        t0 = time.perf_counter()
        ctx.density()
        tr = ctx.move_loop()
        ctx.fof()
        ctx.microsteps()
        ctx.centers()
        ctx.unbind()
        ctx.too_small()
        g, ng = ctx.groups()
        return time.perf_counter() - t0, len(tr), ng

    for w in range(args.warmup):
        dt, it, ng = one_pass()
        log(f"warmup {w}: {dt:.3f}s ittr={it} groups={ng - 1}")
    ctx.set_profiling(True)
    t_start = ctx.stage_times()
    times = []
    t0 = time.perf_counter()
    for s in range(args.steps):
        dt, it, ng = one_pass()
        times.append(dt)
        log(f"step {s}: {dt:.3f}s ittr={it} groups={ng - 1}")
    wall = time.perf_counter() - t0
    kst = ctx.kernel_stats()
    st = {k: v - t_start[k] for k, v in ctx.stage_times().items()}
    ms_per_step = 1000.0 * wall / args.steps
    value = box.n / (wall / args.steps)
    log("stage ms per step:", {k: round(v / args.steps, 1) for k, v in st.items()})
    log("kernels:", json.dumps({k: {kk: (round(vv, 3) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                                 for k, v in kst.items()}))

    # roofline of the kernel class with the most device time
    dom = max(kst, key=lambda k: kst[k]["ms"])
    d = kst[dom]
    avg_ms = d["ms"] / max(1, d["launches"])
    if dom == "unbind_potential":
        ach = d["work"] / max(1, d["launches"]) / (avg_ms * 1e-3) / 1e12
        roof = dict(bound="valu", achieved=ach, peak=FP32_PEAK_TFLOPS, unit="TFLOP/s", frac=ach / FP32_PEAK_TFLOPS,
                    traffic=None, kernel=dom, avg_launch_ms=avg_ms, launches=d["launches"])
    else:
        ach = d["work"] / max(1, d["launches"]) / (avg_ms * 1e-3) / 1e9
        roof = dict(bound="hbm", achieved=ach, peak=HBM_PEAK_GBS, unit="GB/s", frac=ach / HBM_PEAK_GBS,
                    traffic=None, kernel=dom, avg_launch_ms=avg_ms, launches=d["launches"])
    knn = kst["knn"]
    knn_qps = (knn["work"] / 16.0) / (knn["ms"] * 1e-3) if knn["ms"] > 0 else None

    cpu = None
    if args.cpu_sample > 0:
        ce = 1 << args.cpu_sample
        cpu = cpu_baseline(ce, 0.001 * (2.0 ** ((24 - args.cpu_sample) / 3.0)))
        log("cpu baseline:", cpu)

    out = {
        "metric": "end-to-end particles/s at 2^24 (adjoint+batched vs plain GPU vs CPU); kNN q/s",
        "value": value, "unit": "particles/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"two-species box {n_each}+{n_each} particles (2^{args.log2_each} gas + "
                               f"2^{args.log2_each} dark, 8000 NFW halos, dN/dM~M^-1.9), nSmooth 64, nBucket 16, "
                               f"tau {tau:.4g}, -gd, periodic unit box, end-to-end group finding",
                   "n_particles": box.n, "form": "plain" if args.plain else "compensated",
                   "knn_queries_per_s": knn_qps},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
