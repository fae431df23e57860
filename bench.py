"""bench.py -- end-to-end SKID group finding on one MI355X.

A step is one full pass of the group-finding path over particles already
resident in HBM: scatterer tree + kNN density, the Denmax move loop to
convergence, FOF, micro-steps, group centres, unbinding and kdTooSmall,
ending with the group ids in host memory (main.c:374-471).

Default workload: 256 independent periodic unit boxes with particle counts
log-uniform in [2048, 524288] (seeded draw, 29.1M particles in all), each
NFW halos (N/4000 of them) plus a 25% uniform background, nSmooth 32,
nBucket as tuned, tau = 0.16 N_box^(-1/3) per box.  Three forms are timed:

  value        batched + adjoint + compensated: every box in the same launches
               (skg_load_batch), fp32 terms summed in FP64, and each step also forms
               the nSmooth sensitivity of every density (d rho / d ln k, the
               dh adjoint direction) and the dx/dh adjoint of the total
               density, and copies both to the host; launch shapes from the
               autotuner.  The membership report built from d rho / d ln k
               (host post-processing) is formed once after the timed steps.
               Particles are resident in HBM when a step starts; the same step
               with the upload from host memory is reported beside it
               (config.host_to_ids_particles_per_s);
  plain_gpu    one context per box run in turn, fp32 running sums (the
               reference's arithmetic), no adjoint, default shapes; timed on
               every 4th box in size order (a rate, particles/s);
  cpu_baseline the reference SKID (oracle/_ref) on one core over four boxes
               spread across the size range.

The default run also times the 2^24 gas + 2^24 dark box (nSmooth 64, -gd) for
a few steps and reports it under "two_species" with its own roofline and CPU
baseline (--c3-each 0 skips it).

--workload two-species runs the 2^24 gas + 2^24 dark box instead (one box).
--workload unbind runs config 5: 2^27 particles, 60,000 groups with membership
dN/dn ~ n^-1.9 on [8, 2e6] (NFW spheres, virial ratios in [0.5, 2]) in a uniform
background, flat Lambda (Om 0.3, OL 0.7) at z = 0.5, entered through the -unbind
re-entry (skg_set_groups): a step is the FP64 per-group energies plus the
unbinding and kdTooSmall of every group; CPU baseline: the reference's -unbind
path on the groups of at most 2e4 members of a smaller draw.
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


def pin_one_core():
    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
    except Exception:
        pass


class quiet_stdout:
    """The reference's C code prints its progress on stdout, which here carries
    only the one JSON line: send file descriptor 1 to /dev/null meanwhile, and
    flush C stdio before it comes back."""

    def __enter__(self):
        import ctypes
        self.libc = ctypes.CDLL(None)
        sys.stdout.flush()
        self.libc.fflush(None)
        self.saved = os.dup(1)
        null = os.open(os.devnull, os.O_WRONLY)
        os.dup2(null, 1)
        os.close(null)
        return self

    def __exit__(self, *exc):
        sys.stdout.flush()
        self.libc.fflush(None)
        os.dup2(self.saved, 1)
        os.close(self.saved)
        return False


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_run(boxes, kws):
    """The reference (or, without it, the oracle's restatement) on one core."""
    from oracle import oracle as O
    pin_one_core()
    t0 = time.perf_counter()
    kind = "reference" if O.have_ref() else "port"
    with quiet_stdout():
        for box, kw in zip(boxes, kws):
            if kind == "reference":
                O.ref_pipeline(box, **kw)
            else:
                O.Oracle(box, **kw).run()
    return time.perf_counter() - t0, kind


def copy_calibration():
    """A plain device copy of known size (256 MiB read, 256 MiB written, three
    dispatches of the library's k_copy_calib): what FETCH_SIZE and WRITE_SIZE
    read for it gives the factor from counter KiB to HBM bytes."""
    from skid_amd import gen, skg
    ctx = skg.Context(gen.config1(n=4096, seed=1), skg.make_params(0.01, 32, period=1.0))
    ctx.device_copy(256 << 20, 3)
    ctx.close()
    log(f"copied {256 << 20} bytes x3")


def measured_traffic(name):
    """HBM bytes per launch of the move-step class from the committed counter
    summary (two rocprofv3 --pmc runs, FETCH_SIZE and WRITE_SIZE, calibrated on
    a copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def roofline(kst, traffic=None):
    """Roofline of the kernel class with the most device time (HIP events on
    the library's stream around every launch of the timed steps)."""
    dom = max(kst, key=lambda k: kst[k]["ms"])
    d = kst[dom]
    avg_ms = d["ms"] / max(1, d["launches"])
    if dom == "unbind_potential":
        ach = d["work"] / max(1, d["launches"]) / (avg_ms * 1e-3) / 1e12
        return dict(bound="valu", achieved=ach, peak=FP32_PEAK_TFLOPS, unit="TFLOP/s", frac=ach / FP32_PEAK_TFLOPS,
                    traffic=None, kernel=dom, avg_launch_ms=avg_ms, launches=d["launches"])
    ach = d["work"] / max(1, d["launches"]) / (avg_ms * 1e-3) / 1e9
    out = dict(bound="hbm", achieved=ach, peak=HBM_PEAK_GBS, unit="GB/s", frac=ach / HBM_PEAK_GBS,
               traffic=traffic if dom == "cover_move" else None, kernel=dom, avg_launch_ms=avg_ms,
               launches=d["launches"])
    if out["traffic"]:
        out["traffic_source"] = "profiles/traffic_batched_256_boxes.md (rocprofv3 FETCH_SIZE/WRITE_SIZE, bytes per launch)"
    return out


def run_path(ctx, adjoint=False, gbar=None):
    """main.c:374-471 on a context; with adjoint, also d rho / d ln nSmooth
    (the dh direction, for the membership report) and the vector-Jacobian
    product sum_j gbar_j d rho_j / d x_i, d h_i (the dx adjoint) with the
    given weights.  Returns (trace, nGroup, (rho, drho, groups))."""
    b2, rho, _ = ctx.density()
    drho = None
    if adjoint:
        drho = ctx.density_dlnk()
        ctx.adjoint(gbar, 3)
    tr = ctx.move_loop()
    ctx.fof()
    ctx.microsteps()
    ctx.centers()
    ctx.unbind()
    ctx.too_small()
    g, ng = ctx.groups()
    return tr, ng, (rho, drho, g)


def bench_batched(args):
    import numpy as np
    from skid_amd import gen, skg
    sizes = gen.config4_sizes(args.boxes)
    t = time.time()
    boxes = [gen.config4_box(int(n), b) for b, n in enumerate(sizes)]
    taus = np.array([gen.config4_tau(int(n)) for n in sizes], np.float32)
    N = int(sum(b.n for b in boxes))
    log(f"generated {len(boxes)} boxes, {N} particles in {time.time() - t:.1f}s")
    nb = 16
    if args.shapes:                                  # fixed shapes (e.g. a profiled rerun of tuned ones)
        nb, sub_max, tile = (int(v) for v in args.shapes.split(","))
        os.environ["SKG_COVER_SUB_MAX"] = str(sub_max)
        os.environ["SKG_POT_TILE"] = str(tile)
    P = skg.make_params(0.01, 32, nBucket=nb, period=1.0, accumulate=skg.SKG_ACC_COMPENSATED)
    skg.device_memory(reset_peak=True)
    ctx = skg.BatchContext(boxes, P, taus)           # particles resident in HBM from here on
    t = time.time()
    tuned = ctx.autotune(0 if args.shapes else skg.SKG_TUNE_ALL)
    log(f"autotune {time.time() - t:.1f}s: {tuned}")
    # outputs (and the host copy of the input that reload() uploads) in
    # page-locked arrays the context keeps and overwrites each step
    ctx.keep_outputs(True)
    gbar = skg.pinned_empty(N, np.float32)           # dx adjoint of the total density sum_j rho_j
    gbar[:] = 1.0
    for w in range(args.warmup):
        t0 = time.perf_counter()
        tr, ng, last = run_path(ctx, adjoint=True, gbar=gbar)
        log(f"warmup {w}: {time.perf_counter() - t0:.3f}s groups={ng - 1}")
    tuned = ctx.autotune(0)                          # the tile chosen on the warmup's groups
    ctx.set_profiling(True)
    st0 = ctx.stage_times()
    t0 = time.perf_counter()
    for s in range(args.steps):
        ts = time.perf_counter()
        tr, ng, last = run_path(ctx, adjoint=True, gbar=gbar)
        log(f"step {s}: {time.perf_counter() - ts:.3f}s groups={ng - 1}")
    wall = time.perf_counter() - t0
    ctx.kernel_stats_last = ctx.kernel_stats()
    ctx.stage_times_last = ctx.stage_times()
    ctx.set_profiling(False)
    # the same step from particles in host memory: the upload (skg_load_batch,
    # host SoA to HBM over PCIe) inside the bracket; reported beside value
    pcie = []
    for s in range(args.pcie_steps):
        ts = time.perf_counter()
        ctx.reload()
        run_path(ctx, adjoint=True, gbar=gbar)
        pcie.append(time.perf_counter() - ts)
        log(f"host-to-ids step {s}: {pcie[-1]:.3f}s")
    rho, drho, g = last
    rp = skg.nsmooth_report(rho, drho, g, dens_min=float(np.quantile(rho[rho > 0][::64], 0.25)), rel_k=0.25)
    report = {"dlnrho_dlnk": rp["dlnrho_dlnk"], "dens_min_probe": rp["dens_min"], "rel_k": rp["rel_k"],
              "at_risk": rp["at_risk"], "groups_with_risk": int((rp["group_frac_at_risk"][1:] > 0).sum())}
    log("nSmooth sensitivity report:", report)
    kst = ctx.kernel_stats_last
    dev_peak = skg.device_memory()[1]
    st = {k: (v - st0[k]) / args.steps for k, v in ctx.stage_times_last.items()}
    log("stage ms per step:", {k: round(v, 1) for k, v in st.items()})
    log("kernels:", json.dumps({k: {kk: (round(vv, 3) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                                 for k, v in kst.items()}))
    ctx.close()
    value = N / (wall / args.steps)
    knn = kst["knn"]
    knn_qps = (knn["work"] / 16.0) / (knn["ms"] * 1e-3) if knn["ms"] > 0 else None

    plain = None
    if args.plain_every > 0:
        order = np.argsort(sizes)
        pick = [int(i) for i in order[::args.plain_every]]
        ctxs = [skg.Context(boxes[i], skg.make_params(float(taus[i]), 32, nBucket=16, period=1.0,
                                                     accumulate=skg.SKG_ACC_PLAIN)) for i in pick]
        npl = int(sum(boxes[i].n for i in pick))
        run_path(ctxs[0])                            # warm the kernels
        t0 = time.perf_counter()
        for c in ctxs:
            run_path(c)
        dt = time.perf_counter() - t0
        for c in ctxs:
            c.close()
        plain = dict(value=npl / dt, unit="particles/s", boxes=len(pick), particles=npl, seconds=dt)
        log("plain GPU form:", plain)

    cpu = None
    if args.cpu_boxes > 0:
        order = np.argsort(sizes)
        qs = [0.05 + 0.6 * i / max(1, args.cpu_boxes - 1) for i in range(args.cpu_boxes)]
        pick = [int(order[int(q * (len(order) - 1))]) for q in qs]
        sb = [boxes[i] for i in pick]
        kws = [dict(nSmooth=32, fTau=float(taus[i]), fPeriod=1.0) for i in pick]
        dt, kind = cpu_run(sb, kws)
        ncpu = int(sum(b.n for b in sb))
        cpu = dict(value=ncpu / dt, unit="particles/s", cores=1, kind=kind, cpu_model=cpu_model(),
                   sample=f"{len(sb)} of the 256 boxes ({', '.join(str(b.n) for b in sb)} particles; 5th-65th "
                          f"size percentiles), each run end-to-end as one skid process would, {dt:.2f}s on 1 core")
        log("cpu baseline:", cpu)

    forms = {"batched_adjoint_compensated": value,
             "plain_gpu_per_box": plain["value"] if plain else None,
             "cpu_reference_1core": cpu["value"] if cpu else None}
    if plain:
        forms["batched_over_plain"] = value / plain["value"]
    if cpu:
        forms["gpu_over_cpu"] = value / cpu["value"]
    pcie_rate = N / float(np.median(pcie)) if pcie else None
    config = {"workload": f"{len(boxes)} independent periodic unit boxes, N log-uniform in [2048, 524288] "
                          f"(seeded, {N} particles), NFW halos (N/4000 per box) + 25% background, nSmooth 32, "
                          f"tau 0.16 N^-1/3 per box. A step runs from particles resident in HBM to group ids in "
                          f"host memory: density, move loop, FOF, micro-steps, centres, unbinding, kdTooSmall, "
                          f"plus per particle d rho / d ln nSmooth and the dx/dh adjoint of sum_j rho_j, both "
                          f"copied to the host (the membership report built from d rho / d ln nSmooth is host "
                          f"post-processing, formed once after the timed steps). Host arrays are page-locked "
                          f"and reused from step to step. host_to_ids_particles_per_s adds the upload of the "
                          f"particles from host memory (skg_load_batch) to each step",
              "n_particles": N, "boxes": len(boxes), "form": "batched+adjoint+compensated",
              "host_to_ids_particles_per_s": pcie_rate,
              "autotune": tuned, "knn_queries_per_s": knn_qps, "device_peak_bytes": dev_peak,
              "forms": forms, "plain_gpu": plain,
              "nsmooth_report": report}
    return value, wall, config, roofline(kst, measured_traffic("traffic_batched_256_boxes.md")), cpu


def bench_two_species(args):
    from skid_amd import gen, skg
    le = args.log2_each
    n_each = 1 << le
    tau = 0.001 * (2.0 ** ((24 - le) / 3.0))
    t = time.time()
    box = gen.config3(n_each=n_each)
    log(f"generated {box.n} particles in {time.time() - t:.1f}s")
    P = skg.make_params(tau, nSmooth=64, nBucket=16, period=1.0, bGasAndDark=1, accumulate=skg.SKG_ACC_COMPENSATED)
    skg.device_memory(reset_peak=True)
    ctx = skg.Context(box, P)
    ctx.keep_outputs(True)
    for w in range(args.warmup):
        t0 = time.perf_counter()
        tr, ng, _ = run_path(ctx)
        log(f"warmup {w}: {time.perf_counter() - t0:.3f}s ittr={len(tr)} groups={ng - 1}")
    ctx.set_profiling(True)
    t0 = time.perf_counter()
    for s in range(args.steps):
        ts = time.perf_counter()
        tr, ng, _ = run_path(ctx)
        log(f"step {s}: {time.perf_counter() - ts:.3f}s ittr={len(tr)} groups={ng - 1}")
    wall = time.perf_counter() - t0
    kst = ctx.kernel_stats()
    dev_peak = skg.device_memory()[1]
    log(f"device memory high-water mark {dev_peak / 2**30:.2f} GiB")
    ctx.close()
    knn = kst["knn"]
    cpu = None
    if args.cpu_boxes > 0:
        ce = 1 << 16
        sb = gen.config3(n_each=ce, seed=3, n_halos=max(1, int(8000 * ce / (1 << 24))))
        tc = 0.001 * (2.0 ** ((24 - 16) / 3.0))
        dt, kind = cpu_run([sb], [dict(nSmooth=64, fTau=tc, fPeriod=1.0, bGasAndDark=1)])
        cpu = dict(value=sb.n / dt, unit="particles/s", cores=1, kind=kind, cpu_model=cpu_model(),
                   sample=f"same recipe at 2^16+2^16 particles, tau {tc:.4g}, {dt:.2f}s on 1 core")
    config = {"workload": f"two-species box 2^{le} gas + 2^{le} dark (8000 NFW halos, dN/dM~M^-1.9), nSmooth 64, "
                          f"nBucket 16, tau {tau:.4g}, -gd, periodic unit box, end-to-end group finding",
              "n_particles": box.n, "form": "compensated", "device_peak_bytes": dev_peak,
              "knn_queries_per_s": (knn["work"] / 16.0) / (knn["ms"] * 1e-3) if knn["ms"] > 0 else None}
    return box.n / (wall / args.steps), wall, config, roofline(kst), cpu


def bench_unbind(args):
    import numpy as np
    from skid_amd import gen, skg
    n = 1 << args.log2_n
    t = time.time()
    box, grp = gen.config5(n=n, n_groups=args.groups)
    sz = np.bincount(grp)[1:]
    pairs = float((sz * (sz - 1.0)).sum())
    log(f"generated {n} particles, {args.groups} groups ({sz.sum()} members, largest {sz.max()}, "
        f"sum n(n-1) {pairs:.3g}) in {time.time() - t:.1f}s")
    tau = 0.16 * n ** (-1.0 / 3.0)
    fs, fc = gen.config5_cosmology()
    P = skg.make_params(tau, 32, period=1.0, z=0.5, fShift=fs, fCosmo=fc)
    skg.device_memory(reset_peak=True)
    ctx = skg.Context(box, P)

    # the timed step is the reference's -unbind path: kdInGroup/kdInitpGroup/
    # kdReadCenter (set_groups), kdUnbind, kdTooSmall, kdOutGroup's renumbering
    def one():
        ctx.set_groups(grp)
        nu = ctx.unbind()
        ctx.too_small()
        g, ngf = ctx.groups()
        return nu, ngf

    for w in range(args.warmup):
        t0 = time.perf_counter()
        nu, ngf = one()
        log(f"warmup {w}: {time.perf_counter() - t0:.3f}s unbound={nu} groups={ngf - 1}")
    ctx.set_profiling(True)
    t0 = time.perf_counter()
    for s in range(args.steps):
        ts = time.perf_counter()
        nu, ngf = one()
        log(f"step {s}: {time.perf_counter() - ts:.3f}s unbound={nu} groups={ngf - 1}")
    wall = time.perf_counter() - t0
    kst = ctx.kernel_stats()
    dev_peak = skg.device_memory()[1]
    # per-group FP64 binding energies of the input groups, outside the timed steps
    ts = time.perf_counter()
    ng = ctx.set_groups(grp)
    W, T = ctx.group_energies(ng)
    bound = int(((W[1:] + T[1:]) < 0).sum())
    log(f"group energies: {time.perf_counter() - ts:.3f}s, {bound} of {ng - 1} input groups with T+W < 0")
    log("kernels:", json.dumps({k: {kk: (round(vv, 3) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                                 for k, v in kst.items()}))
    ctx.close()
    members = int(sz.sum())
    # the potential is the VALU-bound kernel; report its roofline
    d = kst["unbind_potential"]
    avg = d["ms"] / max(1, d["launches"])
    ach = d["work"] / max(1, d["launches"]) / (avg * 1e-3) / 1e12
    roof = dict(bound="valu", achieved=ach, peak=FP32_PEAK_TFLOPS, unit="TFLOP/s", frac=ach / FP32_PEAK_TFLOPS,
                traffic=None, kernel="unbind_potential", avg_launch_ms=avg, launches=d["launches"],
                note="20 flops per ordered member pair; the removal loops (FP64 test) are latency-bound")
    cpu = None
    if args.cpu_boxes > 0:
        # the reference's -unbind path (main.c:349-373) on a 2^20 draw with groups capped at 2e4 members
        nb = 1 << 20
        b2, g2 = gen.config5(n=nb, n_groups=500, seed=6, n_range=(8, 20000))
        kw = dict(nSmooth=32, fTau=0.16 * nb ** (-1.0 / 3.0), fPeriod=1.0, z=0.5, Omega0=0.3, Lambda=0.7,
                  H0=2.8944)
        from oracle import oracle as O
        if O.have_ref():
            pin_one_core()
            t1 = time.perf_counter()
            with quiet_stdout():
                O.ref_unbind(b2, g2, **kw)
            dt = time.perf_counter() - t1
            m2 = int((g2 > 0).sum())
            cpu = dict(value=m2 / dt, unit="group members/s", cores=1, kind="reference", cpu_model=cpu_model(),
                       sample=f"reference -unbind path on a 2^20 draw of the same recipe with 500 groups of "
                              f"8..2e4 members ({m2} members), {dt:.2f}s on 1 core")
    value = members / (wall / args.steps)
    config = {"workload": f"unbinding stress: {n} particles, {args.groups} power-law groups ({members} members, "
                          f"largest {int(sz.max())}, sum n(n-1) {pairs:.3g}), z=0.5 flat Lambda, -unbind re-entry, "
                          f"FP32 compensated potentials, FP64 energies and sign test",
              "n_particles": n, "members": members, "form": "unbind", "device_peak_bytes": dev_peak}
    return value, wall, config, roof, cpu


def c3_record(args):
    """The two-species box (2^24 gas + 2^24 dark, nSmooth 64, tau 0.001, -gd) as
    a record of its own inside the default run's line: its particles/s over a
    few steps, the roofline of its dominant kernel, and the reference on one
    core at 2^16 + 2^16 particles."""
    a = argparse.Namespace(**vars(args))
    a.log2_each, a.steps, a.warmup, a.cpu_boxes = args.c3_each, args.c3_steps, 1, args.cpu_boxes
    value, wall, config, roof, cpu = bench_two_species(a)
    return {"metric": "end-to-end particles/s, two-species box", "value": value, "unit": "particles/s",
            "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1000.0 * wall / a.steps, "config": config,
            "roofline": roof, "cpu_baseline": cpu,
            "gpu_over_cpu": (value / cpu["value"]) if cpu else None}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=("batched", "two-species", "unbind"), default="batched")
    ap.add_argument("--log2-n", type=int, default=27, help="unbind: particles = 2^this")
    ap.add_argument("--groups", type=int, default=60000, help="unbind: number of groups")
    ap.add_argument("--boxes", type=int, default=256)
    ap.add_argument("--log2-each", type=int, default=24, help="two-species: particles per species = 2^this")
    ap.add_argument("--plain-every", type=int, default=4, help="plain GPU form on every n-th box (0 skips)")
    ap.add_argument("--cpu-boxes", type=int, default=4, help="boxes in the CPU sample (0 skips)")
    ap.add_argument("--c3-each", type=int, default=24,
                    help="batched: also time the two-species box at 2^this per species (0 skips)")
    ap.add_argument("--c3-steps", type=int, default=2, help="timed steps of the two-species record")
    ap.add_argument("--pcie-steps", type=int, default=1,
                    help="batched: steps timed from particles in host memory (the upload included)")
    ap.add_argument("--shapes", default="", help="nBucket,subMax,potTile: skip the autotuner and use these")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB on the device three times (the reference kernel for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py: the group-finding path runs on one GPU; --gpus {args.gpus} is not supported "
              f"(sharding across GPUs is not implemented)", file=sys.stderr)
        sys.exit(2)
    c3 = None
    if args.workload == "batched":
        value, wall, config, roof, cpu = bench_batched(args)
        if args.c3_each > 0:
            c3 = c3_record(args)
    elif args.workload == "unbind":
        value, wall, config, roof, cpu = bench_unbind(args)
    else:
        value, wall, config, roof, cpu = bench_two_species(args)
    metric = "end-to-end particles/s at 2^24 (adjoint+batched vs plain GPU vs CPU); kNN q/s"
    unit = "particles/s"
    if args.workload == "unbind":
        metric, unit = "-unbind group members/s (power-law groups)", "group members/s"
    out = {
        "metric": metric,
        "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1000.0 * wall / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic", "config": config,
        "roofline": roof, "cpu_baseline": cpu,
    }
    if c3:
        out["two_species"] = c3
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
