#!/usr/bin/env python3
"""cat2amln throughput on one MI355X at N=1e6, kmax=0.1, nside=128.

Workload (C2): 1e6 galaxies uniform in volume in the shell 500 <= r <= 1500 Mpc/h
(r^3 uniform, cos(theta) uniform, phi uniform, seed 0), unit weights, nbar = N/V,
AnlmModes(kmax=0.1, rmin=500, rmax=1500; nside=128) with the potential boundary
condition (lmax=140, nmax=32, 2514 (n,l) maps, 116345 a_nlm), full-sky separable
window (mask = 1, phi(r) = 1, nr = 1000), map2alm niter=3.

`value` is the metric as the reference's users meet it: one sfb_cat2amln call on
the mixed-precision plan, from the host catalogue (3 x N, the reference's layout)
to the host a_nlm vector, catalogue upload and a_nlm download included, timed with
profiling off.  Plan setup (AnlmModes, splines, Legendre tables, tile search) is
outside the timed region and reported under setup_s.  Also in the line:

* the FP64 plan on the same call, and both plans with everything resident in HBM
  (sfb_cat2amln_device);
* per-phase times from a separate profiled pass (HIP events on the plan's stream),
  and the roofline of the dominant kernel from them;
* errors of the mixed path against the FP64 GPU path over every a_nlm, and of both
  against the oracle on a spread sample of maps, per mode and per (n,l) block;
* C3 (4e6 galaxies, masked, nside=256, kmax=0.15), C4 (64 mocks: one batch call,
  eager and as a HIP graph, against 64 sequential calls), C5 (adjoint);
* the CPU baseline: the oracle's C restatement on nproc threads and on one thread.

Prints ONE JSON line.  --gpus N > 1 is refused: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "sphericalfourierbesseldecompositions.jl_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

METRIC = "cat2amln Mgal/s at N=1e6,kmax=0.1,nside=128: mixed vs FP64 GPU, max rel err"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3

# phase -> the kernel it brackets (the roofline is taken for single-kernel phases)
PHASE_KERNEL = {"scatter": "k_scatter_it", "segment": "k_seg_t", "phi": "k_phi_fft", "phi_polar": "k_phi_polar4",
                "analysis": "k_gemm_tall", "synthesis": "k_gemm_tall", "residual": "k_alias"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def nproc():
    """Processing units available to this process, as `nproc` counts them."""
    try:
        return int(subprocess.run(["nproc"], capture_output=True, text=True, timeout=10).stdout.strip())
    except (OSError, ValueError, subprocess.SubprocessError):
        return len(os.sched_getaffinity(0))


def uniform_catalogue(N, rmin, rmax, seed):
    rng = np.random.Generator(np.random.Philox(seed))
    u = rng.random((3, N))
    r = np.cbrt(u[0] * (rmax ** 3 - rmin ** 3) + rmin ** 3)
    th = np.arccos(np.clip(1.0 - 2.0 * u[1], -1.0, 1.0))
    ph = 2 * np.pi * u[2]
    return np.stack([r, th, ph])


def c3_catalogue(N, seed, rmin=500.0, rmax=1500.0, r0=1000.0, sigma=300.0):
    """Northern hemisphere, radial density ~ r^2 exp(-(r - r0)^2 / (2 sigma^2)) on [rmin, rmax]
    (rejection from uniform in volume), FP32-valued weights uniform in [0.5, 1.5]."""
    rng = np.random.Generator(np.random.Philox(seed))
    r = np.empty(0)
    while r.size < N:
        u = rng.random(2 * (N - r.size) + 1000)
        rr = np.cbrt(u * (rmax ** 3 - rmin ** 3) + rmin ** 3)
        r = np.concatenate([r, rr[rng.random(rr.size) < np.exp(-(rr - r0) ** 2 / (2 * sigma ** 2))]])
    th = np.arccos(1.0 - rng.random(N))
    ph = 2 * np.pi * rng.random(N)
    w = rng.uniform(0.5, 1.5, N).astype(np.float32).astype(np.float64)
    return np.stack([r[:N], th, ph]), w


def block_table(am):
    """(n, l, first, end) of every (n, l) block of a_nlm, in map order."""
    out, i = [], 0
    for n in range(1, am.nmax + 1):
        for l in range(int(am.lmax_n[n - 1]) + 1):
            out.append((n, l, i, i + l + 1))
            i += l + 1
    return out


def errors(a, ref, am, qs=None):
    """Per mode: max and RMS of |a - ref| / |ref|.  Per (n, l) block: max_blk = max over blocks
    of max_m |a - ref| / RMS_m |ref|; rms_blk = RMS over modes of |a - ref| / RMS_block |ref|.
    Over the maps qs (map indices), or all."""
    rel, blk_max, blk_all, worst = [], [], [], None
    for q, (n, l, i0, i1) in enumerate(block_table(am)):
        if qs is not None and q not in qs:
            continue
        r, d = ref[i0:i1], np.abs(a[i0:i1] - ref[i0:i1])
        rms = max(float(np.sqrt(np.mean(np.abs(r) ** 2))), 1e-300)
        rel.append(d / np.maximum(np.abs(r), 1e-300))
        blk_all.append(d / rms)
        e = float(d.max() / rms)
        if worst is None or e > worst[0]:
            worst = (e, n, l)
        blk_max.append(e)
    rel, blk_all = np.concatenate(rel), np.concatenate(blk_all)
    return {"max_rel": float(rel.max()), "rms_rel": float(np.sqrt(np.mean(rel ** 2))),
            "max_blk": float(max(blk_max)), "rms_blk": float(np.sqrt(np.mean(blk_all ** 2))),
            "worst_block_nl": [worst[1], worst[2]], "blocks": len(blk_max), "modes": int(rel.size)}


def algorithmic_work(am, N, f32=False, ngroup=None):
    """Per-launch algorithmic bytes / flops of each single-kernel phase (f32: the mixed path's
    FP32 delta maps and ring coefficients)."""
    L, nside = am.lmax, am.nside
    npix, K, nmaps = 12 * nside * nside, 2 * nside, am.nmaps
    M2 = 2 * nmaps
    vb = 4 if f32 else 8
    nlm_pairs = (L + 1) * (L + 2) // 2  # sum over (m, parity) of the l count
    delta_bytes = nmaps * npix * vb
    x_bytes = (L + 1) * 2 * K * M2 * vb
    a_bytes = nlm_pairs * M2 * 8
    # rings with 4 nside pixels: ring pairs nside-1 .. 2 nside-1 (2 nside + 1 rings)
    k_eq0 = nside - 1
    npix_eq = (2 * nside + 1) * 4 * nside
    x_eq = x_bytes * (K - k_eq0) // K
    if ngroup is None:
        ngroup = len({am.knot_x[am.knot_off[q]:am.knot_off[q + 1]].tobytes() for q in range(am.nmaps)})
    return {
        # delta maps written once; per pixel its slice entry and window factor read once; per
        # galaxy its radius read once (mixed path: segment and t found from it on the uniform
        # knots) or its 16-bit segment and t (FP64 path)
        "scatter": {"bytes": delta_bytes + N * (8 if f32 else 2 + vb) + (npix + 1) * 4 + npix * 8,
                    "flops": 10.0 * N * nmaps},
        # per knot group and galaxy: r read, (16-bit segment, t) written
        "segment": {"bytes": ngroup * N * (8 + 2 + vb), "flops": 0.0},
        "phi": {"bytes": nmaps * npix_eq * vb + x_eq, "flops": 0.0},
        "phi_polar": {"bytes": nmaps * (npix - npix_eq) * vb + (x_bytes - x_eq), "flops": 0.0},
        # a = dOmega * X Lambda with N/S ring-pair symmetry: 2 flop x rows x K x sum n_b
        "analysis": {"bytes": x_bytes + a_bytes, "flops": 2.0 * M2 * K * nlm_pairs},
        "synthesis": {"bytes": x_bytes + a_bytes, "flops": 2.0 * M2 * K * nlm_pairs},
        "residual": {"bytes": 3 * x_bytes, "flops": 0.0},
    }


def timed(fn, steps, warmup, sync):
    for _ in range(warmup):
        fn()
    sync()
    t = time.perf_counter()
    for _ in range(steps):
        fn()
    sync()
    return (time.perf_counter() - t) / steps


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--N", type=int, default=1_000_000)
    ap.add_argument("--cpu-maps", type=int, default=480,
                    help="maps of the CPU baseline sample on nproc threads (8 runs spread over the map list)")
    ap.add_argument("--cpu-maps-1t", type=int, default=32, help="maps of the one-thread CPU sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--mocks", type=int, default=64,
                    help="C4: catalogues with N_b ~ Poisson(mu_b), mu_b ~ U[1e5, 1e6] (0 skips it)")
    ap.add_argument("--no-c3", action="store_true", help="skip the C3 (4e6 galaxies, nside 256) timing")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-host-path", action="store_true",
                    help="time only the device-resident calls (profiling runs); `value` is then the "
                         "device-resident figure and value_path says so")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one elementwise copy of 512 MiB (torch.neg, FP64) before the steps, as the "
                         "known-size kernel that scales FETCH_SIZE / WRITE_SIZE counter readings")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the (n,l) maps would shard "
              "across GPUs (each GPU a share of the maps, the sorted catalogue broadcast once, a_nlm "
              "gathered), but no multi-GPU path is built.", file=sys.stderr)
        sys.exit(2)

    import torch
    import sfb_amd as S
    from sfb_amd import _lib as L

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    dev = torch.device("cuda:0")
    sync = torch.cuda.synchronize
    rmin, rmax, kmax, nside, niter = 500.0, 1500.0, 0.1, 128, 3
    N = args.N

    t0 = time.perf_counter()
    am = S.AnlmModes(kmax, rmin, rmax, nside=nside)
    t_modes = time.perf_counter() - t0
    nlm = S.getnlmsize(am)
    log(f"C2 AnlmModes: lmax={am.lmax} nmax={am.nmax} maps={am.nmaps} nlm={nlm} ({t_modes:.2f}s)")
    wmodes = S.ConfigurationSpaceModes(rmin, rmax, 1000, nside)
    npix = 12 * nside * nside
    t0 = time.perf_counter()
    win = S.win_rhat_ln(S.SeparableArray(np.ones(1000), np.ones(npix), "phi", "mask"), wmodes, am)
    t_win = time.perf_counter() - t0
    cat = uniform_catalogue(N, rmin, rmax, seed=0)
    cat_cm = np.asfortranarray(cat)  # 3 x N column-major, as the reference passes it
    nbar = N / (4 * np.pi / 3 * (rmax ** 3 - rmin ** 3))

    d_rtp = torch.from_numpy(np.ascontiguousarray(cat.T)).to(dev)
    d_mask = torch.from_numpy(np.ascontiguousarray(win.mask, dtype=np.float64)).to(dev)
    d_wln = torch.from_numpy(np.ascontiguousarray(win.w_ln.reshape(-1, order="F"))).to(dev)
    d_out = torch.empty(2 * nlm, dtype=torch.float64, device=dev)

    t0 = time.perf_counter()
    plan = S.Plan(am, niter=niter, precision="fp64")
    t_plan = time.perf_counter() - t0
    tiles, tune_ms = plan.tiles()
    tiles["search_ms"] = tune_ms
    log(f"plan: {t_plan:.2f}s (phi tile search {tune_ms:.0f} ms: {tiles}), window: {t_win:.2f}s")

    if args.calibrate_copy:
        src = torch.ones(64 << 20, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        sync()
        torch.neg(src, out=dst)
        sync()
        del src, dst

    precs = ("fp64", "mixed")

    def dev_call():
        plan.cat2amln_device(N, d_rtp, None, nbar, L.SFB_WIN_SEPARABLE, d_mask, d_wln, d_out)

    # -- C2 from host arrays (the metric) and device-resident, profiling off ----------------
    host_ms, dev_ms, anlm = {}, {}, {}
    for prec in precs:
        plan.set_precision(prec)
        if not args.no_host_path:
            host_ms[prec] = 1e3 * timed(lambda: plan.cat2amln(cat_cm, nbar, win, None), args.steps,
                                        max(1, args.warmup), sync)
            anlm[prec] = plan.cat2amln(cat_cm, nbar, win, None).copy()
            log(f"[C2 {prec} host arrays] {host_ms[prec]:.2f} ms/call ({N / host_ms[prec] / 1e3:.1f} Mgal/s)")
        dev_ms[prec] = 1e3 * timed(dev_call, args.steps, args.warmup, sync)
        if prec not in anlm:
            anlm[prec] = d_out.cpu().numpy().view(np.complex128).copy()
        log(f"[C2 {prec} device-resident] {dev_ms[prec]:.2f} ms/call ({N / dev_ms[prec] / 1e3:.1f} Mgal/s)")

    # -- per-phase times: a separate profiled pass --------------------------------------------
    phases = {}
    for prec in precs:
        plan.set_precision(prec)
        dev_call()
        plan.profile(True)
        for _ in range(args.steps):
            dev_call()
        sync()
        ph, calls = plan.phase_times()
        plan.profile(False)
        phases[prec] = {k: {"ms_per_call": v[0] / max(calls, 1), "launches_per_call": v[1] / max(calls, 1),
                            "ms_per_launch": v[0] / max(v[1], 1)} for k, v in ph.items() if v[1] > 0}
        log(f"[C2 {prec} phases] " + "  ".join(f"{k}={v['ms_per_call']:.3f}" for k, v in phases[prec].items()))

    # -- accuracy of the mixed path against the FP64 GPU path over every a_nlm ---------------
    err_mixed_vs_fp64 = errors(anlm["mixed"], anlm["fp64"], am)
    log(f"mixed vs FP64 GPU over {err_mixed_vs_fp64['blocks']} blocks: {err_mixed_vs_fp64}")

    # -- split-BF16 Jacobi GEMMs (hi + mid + lo and hi + lo bf16 terms), device-resident -----
    split_bf16 = {}
    for prec in ("bf16x3", "bf16x2"):
        plan.set_precision(prec)
        ms = 1e3 * timed(dev_call, args.steps, args.warmup, sync)
        e = errors(d_out.cpu().numpy().view(np.complex128).copy(), anlm["fp64"], am)
        split_bf16[prec] = {"device_resident_ms": ms, "max_blk_vs_fp64": e["max_blk"], "rms_blk_vs_fp64": e["rms_blk"]}
        log(f"[C2 {prec} device-resident] {ms:.2f} ms/call, max_blk vs FP64 GPU {e['max_blk']:.2e}")

    # -- C5: adjoint (dL/dr, dL/dw for a random cotangent), device-resident ------------------
    adjoint = None
    if not args.no_adjoint:
        adjoint = {}
        d_cot = torch.randn(2 * nlm, dtype=torch.float64, device=dev)
        d_dr = torch.empty(N, dtype=torch.float64, device=dev)
        d_dw = torch.empty(N, dtype=torch.float64, device=dev)
        for prec in precs:
            plan.set_precision(prec)
            adjoint[prec + "_ms"] = 1e3 * timed(
                lambda: plan.cat2amln_vjp_device(N, d_rtp, None, nbar, d_cot, d_dr, d_dw), args.steps,
                max(1, args.warmup), sync)
            log(f"[C5 adjoint {prec}] {adjoint[prec + '_ms']:.2f} ms/call")
        del d_cot, d_dr, d_dw

    # -- C4: batched mocks in the C2 geometry, generated on the device -----------------------
    batch = None
    if args.mocks > 0:
        B = args.mocks
        rng = np.random.default_rng(3)
        N_b = rng.poisson(rng.uniform(1e5, 1e6, B)).astype(np.int64)
        offs = np.concatenate([[0], np.cumsum(N_b)])
        tot = int(N_b.sum())
        gen = torch.Generator(device=dev)
        gen.manual_seed(2)
        u = torch.rand(3, tot, dtype=torch.float64, device=dev, generator=gen)
        d_mock = torch.stack([torch.pow(u[0] * (rmax ** 3 - rmin ** 3) + rmin ** 3, 1.0 / 3.0),
                              torch.acos(1.0 - 2.0 * u[1]), 2.0 * np.pi * u[2]], dim=1).contiguous()
        del u
        nbars = N_b / (4 * np.pi / 3 * (rmax ** 3 - rmin ** 3))
        d_seq = torch.empty(B * 2 * nlm, dtype=torch.float64, device=dev)
        d_bat = torch.empty_like(d_seq)
        reps = 2

        def sequential():
            for b in range(B):
                plan.cat2amln_device(int(N_b[b]), d_mock[offs[b]:], None, float(nbars[b]), L.SFB_WIN_SEPARABLE,
                                     d_mask, d_wln, d_seq[b * 2 * nlm:])

        def batched():
            plan.cat2amln_batch_device(B, N_b, d_mock, None, nbars, L.SFB_WIN_SEPARABLE, d_mask, d_wln, d_bat)

        batch = {"mocks": B, "galaxies": tot, "mu_range": [1e5, 1e6], "reps": reps}
        for prec in precs:
            plan.set_precision(prec)
            t_seq = timed(sequential, reps, 1, sync)
            ref = d_seq.cpu().numpy()
            plan.set_batch_graph(False)
            t_eager = timed(batched, reps, 1, sync)
            same_eager = bool(np.array_equal(d_bat.cpu().numpy(), ref))
            plan.set_batch_graph(True)
            batched()  # eager run of the largest mock, capture, first launch
            sync()
            cap_ms = plan.set_batch_graph(None)
            t_graph = timed(batched, reps, 0, sync)
            same_graph = bool(np.array_equal(d_bat.cpu().numpy(), ref))
            plan.set_batch_graph(False)
            batch[prec] = {"sequential_ms": t_seq * 1e3, "batch_eager_ms": t_eager * 1e3,
                           "batch_graph_ms": t_graph * 1e3, "graph_capture_ms": cap_ms,
                           "Mgal_s": tot / min(t_eager, t_graph) / 1e6,
                           "speedup_vs_64_sequential": t_seq / min(t_eager, t_graph),
                           "bit_identical_to_sequential": same_eager and same_graph}
            log(f"[C4 {prec}] {B} mocks, {tot} galaxies: sequential {t_seq * 1e3:.1f} ms, batch eager "
                f"{t_eager * 1e3:.1f} ms, graph {t_graph * 1e3:.1f} ms (capture {cap_ms:.0f} ms), "
                f"identical {same_eager and same_graph}")
        del d_mock, d_seq, d_bat

    # -- roofline of the dominant kernel (profiled pass, headline precision) -----------------
    head = "mixed"
    work = algorithmic_work(am, N, f32=(head == "mixed"))
    ph = phases[head]
    single = {k: v for k, v in ph.items() if k in work and abs(v["launches_per_call"] - round(v["launches_per_call"])) < 1e-9}
    dom = max(single, key=lambda k: single[k]["ms_per_launch"])
    w = work[dom]
    t_launch = ph[dom]["ms_per_launch"] * 1e-3
    if dom in ("analysis", "synthesis"):
        achieved = w["flops"] / t_launch / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None}
    else:
        achieved = w["bytes"] / t_launch / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None}
    roof.update({"kernel": PHASE_KERNEL.get(dom, dom), "phase": dom, "algorithmic_bytes_per_launch": w["bytes"],
                 "launch_ms": t_launch * 1e3})
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f).get(head, {}).get(dom)
        if tr:
            roof["traffic"] = tr["fetch_bytes"] + tr["write_bytes"]
            roof["traffic_unit"] = "B/launch"

    # -- CPU baseline and errors against the oracle on a spread sample of maps ---------------
    cpu, err_oracle = None, {}
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        om = O.OracleModes.__new__(O.OracleModes)
        om.nside, om.lmax, om.nmax, om.lmax_n = am.nside, am.lmax, am.nmax, am.lmax_n
        om.knot_off, om.knot_x, om.knot_y, om.knot_ypp = am.knot_off, am.knot_x, am.knot_y, am.knot_ypp
        wln = win.w_ln.reshape(-1, order="F")

        def sample(nmaps_total, threads):
            """8 runs of maps spread evenly over the map list (low to high n and l)."""
            per = max(1, nmaps_total // 8)
            starts = [int(k * (am.nmaps - per) / 7) for k in range(8)]
            ref = np.full(nlm, np.nan, dtype=complex)
            qs = set()
            t = time.perf_counter()
            for q0 in starts:
                r = O.cat2amln(om, cat, nbar, 1, win.mask, wln, q_range=(q0, q0 + per), nthreads=threads)
                qs.update(range(q0, q0 + per))
                ok = np.isfinite(r)
                ref[ok] = r[ok]
            return time.perf_counter() - t, ref, qs, starts, per

        nth = nproc()
        tc, ref, qs, starts, per = sample(args.cpu_maps, nth)
        t1, _, qs1, starts1, per1 = sample(args.cpu_maps_1t, 1)
        full_n, full_1 = tc * am.nmaps / len(qs), t1 * am.nmaps / len(qs1)
        cpu = {"value": N / full_n / 1e6, "unit": "Mgal/s", "cores": nth, "kind": "port",
               "value_1_thread": N / full_1 / 1e6, "nproc": nth, "os_cpu_count": os.cpu_count(),
               "sample": (f"the oracle's FP64 C restatement of cat2amln (OpenMP over maps; direct per-ring DFT in "
                          f"long double, so slower than the reference's libsharp FFT path: a lower bound on the "
                          f"reference's speed) on the same 1e6-galaxy catalogue: {len(qs)} maps in 8 runs of {per} "
                          f"starting at {starts} on {nth} threads (nproc) took {tc:.1f} s, and {len(qs1)} maps in "
                          f"8 runs of {per1} on 1 thread took {t1:.1f} s; each scaled to all {am.nmaps} maps")}
        for prec in precs:
            err_oracle[prec] = errors(anlm[prec], ref, am, qs)
        log(f"cpu baseline: {cpu['value']:.5f} Mgal/s on {nth} threads, {cpu['value_1_thread']:.5f} on 1; "
            f"errors vs oracle {err_oracle}")

    # -- C3: masked, selected survey at nside 256 -------------------------------------------
    c3 = None
    if not args.no_c3:
        plan.close()
        del d_rtp, d_mask, d_wln, d_out
        torch.cuda.empty_cache()
        t0 = time.perf_counter()
        am3 = S.AnlmModes(0.15, 500.0, 1500.0, nside=256)
        nlm3 = S.getnlmsize(am3)
        wm3 = S.ConfigurationSpaceModes(500.0, 1500.0, 1000, 256)
        phi3 = np.exp(-(wm3.r - 1000.0) ** 2 / (2 * 300.0 ** 2))
        # pixels with theta < pi/2: the ring scheme numbers rings from the north pole, and the
        # rings above the equator ring (z = 0) hold 2 ns (ns - 1) + 4 ns^2 pixels
        ns3 = 256
        mask3 = np.zeros(12 * ns3 ** 2)
        mask3[:2 * ns3 * (ns3 - 1) + 4 * ns3 * ns3] = 1.0
        sep3 = S.SeparableArray(phi3, mask3, name1="phi", name2="mask")
        win3 = S.win_rhat_ln(sep3, wm3, am3)
        N3 = 4_000_000
        cat3, w3 = c3_catalogue(N3, seed=3)
        cat3 = np.asfortranarray(cat3)
        nbar3 = N3 / S.integrate_window(sep3, wm3)
        plan3 = S.Plan(am3, niter=niter, precision="fp64")
        setup3 = time.perf_counter() - t0
        log(f"C3 setup: lmax={am3.lmax} nmax={am3.nmax} maps={am3.nmaps} ({setup3:.1f}s)")
        d3_rtp = torch.from_numpy(np.ascontiguousarray(cat3.T)).to(dev)
        d3_w = torch.from_numpy(w3).to(dev)
        d3_mask = torch.from_numpy(np.ascontiguousarray(win3.mask, dtype=np.float64)).to(dev)
        d3_wln = torch.from_numpy(np.ascontiguousarray(win3.w_ln.reshape(-1, order="F"))).to(dev)
        d3_out = torch.empty(2 * nlm3, dtype=torch.float64, device=dev)
        c3 = {"workload": "N=4e6 northern hemisphere, Gaussian radial profile (r0 1000, sigma 300), FP32-valued "
                          "weights, separable window (hemisphere mask x profile), kmax=0.15, nside=256, lmax=214, "
                          "nmax=48, 5645 maps, niter=3", "setup_s": setup3}
        a3 = {}
        steps3 = max(2, args.steps // 2)
        for prec in precs:
            plan3.set_precision(prec)
            c3[prec + "_host_ms"] = 1e3 * timed(lambda: plan3.cat2amln(cat3, nbar3, win3, w3), steps3, 1, sync)
            a3[prec] = plan3.cat2amln(cat3, nbar3, win3, w3).copy()
            c3[prec + "_device_ms"] = 1e3 * timed(
                lambda: plan3.cat2amln_device(N3, d3_rtp, d3_w, nbar3, L.SFB_WIN_SEPARABLE, d3_mask, d3_wln, d3_out),
                steps3, 1, sync)
            log(f"[C3 {prec}] host arrays {c3[prec + '_host_ms']:.1f} ms, device-resident "
                f"{c3[prec + '_device_ms']:.1f} ms")
        c3["speedup_mixed_over_fp64_host"] = c3["fp64_host_ms"] / c3["mixed_host_ms"]
        c3["mixed_vs_fp64_gpu"] = errors(a3["mixed"], a3["fp64"], am3)
        log(f"C3 mixed vs FP64 GPU: {c3['mixed_vs_fp64_gpu']}")
        plan3.close()

    # -- the line ----------------------------------------------------------------------------
    if args.no_host_path:
        ms_head, value_path = dev_ms[head], "device-resident (sfb_cat2amln_device)"
    else:
        ms_head, value_path = host_ms[head], "host arrays (sfb_cat2amln), catalogue upload and a_nlm download included"
    line = {
        "metric": METRIC, "value": N / ms_head / 1e3, "unit": "Mgal/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_head, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32/f64 mixed", "data": "synthetic",
        "config": {"workload": "C2: N=1e6 uniform shell 500-1500 Mpc/h, kmax=0.1, nside=128, lmax=140, nmax=32, "
                               "2514 maps, niter=3, separable full-sky window",
                   "path": "mixed", "value_path": value_path},
        "host_arrays_Mgal_s": {p: N / host_ms[p] / 1e3 for p in host_ms},
        "device_resident_Mgal_s": {p: N / dev_ms[p] / 1e3 for p in dev_ms},
        "host_arrays_ms": host_ms, "device_resident_ms": dev_ms,
        "speedup_mixed_over_fp64": (host_ms["fp64"] / host_ms["mixed"]) if host_ms else None,
        "speedup_mixed_over_fp64_device_resident": dev_ms["fp64"] / dev_ms["mixed"],
        "max_rel_err": err_oracle.get("mixed", {}).get("max_rel"),
        "max_blk_err": err_oracle.get("mixed", {}).get("max_blk"),
        "errors_vs_oracle": err_oracle,
        "errors_mixed_vs_fp64_gpu_all_maps": err_mixed_vs_fp64,
        "split_bf16_jacobi": split_bf16,
        "error_definitions": "max_rel / rms_rel: per mode |a - ref| / |ref|; max_blk: max over (n,l) blocks of "
                             "max_m |a - ref| / RMS_m |ref|; rms_blk: RMS over modes of |a - ref| / RMS_block |ref|",
        "phases_ms": phases,
        "setup_s": {"AnlmModes": t_modes, "plan": t_plan, "window": t_win},
        "phi_tiles": tiles,
        "C3": c3,
        "C4_batched_mocks": batch,
        "C5_adjoint": adjoint,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
