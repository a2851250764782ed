#!/usr/bin/env python3
"""cat2amln throughput on one MI355X at N=1e6, kmax=0.1, nside=128.

Workload: 1e6 galaxies uniform in volume in the shell 500 <= r <= 1500 Mpc/h
(r^3 uniform, cos(theta) uniform, phi uniform, seed 0), unit weights,
nbar = N/V, AnlmModes(kmax=0.1, rmin=500, rmax=1500; nside=128) with the
potential boundary condition (lmax=140, nmax=32, 2514 (n,l) maps, 116345 a_nlm),
full-sky separable window (mask = 1, phi(r) = 1, nr = 1000), map2alm niter=3.

One step = one sfb_cat2amln_device call with the catalogue, the window and the
output already in HBM: ang2pix, both sorts, the spline scatter, the phi-transform,
all Legendre passes and the packing.  Plan setup (AnlmModes, splines, lambda
tables) is outside the timed region and reported separately.

Prints ONE JSON line (rank 0).  --gpus N > 1 is refused: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def catalogue(N, rmin, rmax, seed):
    rng = np.random.Generator(np.random.Philox(seed))
    u = rng.random((3, N))
    r = np.cbrt(u[0] * (rmax ** 3 - rmin ** 3) + rmin ** 3)
    th = np.arccos(np.clip(1.0 - 2.0 * u[1], -1.0, 1.0))
    ph = 2 * np.pi * u[2]
    return np.stack([r, th, ph])


def algorithmic_work(am, N, niter, f32=False, ngroup=None):
    """Per-launch algorithmic bytes / flops of every phase (f32: the mixed path's
    FP32 delta maps and ring coefficients)."""
    L, nside = am.lmax, am.nside
    npix, K, nmaps = 12 * nside * nside, 2 * nside, am.nmaps
    M2 = 2 * nmaps
    vb = 4 if f32 else 8
    nlm_pairs = (L + 1) * (L + 2) // 2  # sum over (m, parity) of the l count
    delta_bytes = nmaps * npix * vb
    x_bytes = (L + 1) * 2 * K * M2 * vb
    a_bytes = nlm_pairs * M2 * 8
    if ngroup is None:
        ngroup = len({am.knot_x[am.knot_off[q]:am.knot_off[q + 1]].tobytes() for q in range(am.nmaps)})
    return {
        # delta maps written once; per galaxy its weight, 16-bit segment and t read once
        "scatter": {"bytes": delta_bytes + N * (vb + 2 + vb) + (npix + 1) * 4 + npix * 8,
                    "flops": 10.0 * N * nmaps},
        # per knot group and galaxy: r read, (16-bit segment, t) written
        "segment": {"bytes": ngroup * N * (8 + 2 + vb), "flops": 0.0},
        "phi": {"bytes": delta_bytes + x_bytes, "flops": 0.0},
        # a = dOmega * X Lambda with N/S ring-pair symmetry: 2 flop x rows x K x sum n_b
        "analysis": {"bytes": x_bytes + a_bytes, "flops": 2.0 * M2 * K * nlm_pairs},
        "synthesis": {"bytes": x_bytes + a_bytes, "flops": 2.0 * M2 * K * nlm_pairs},
        "residual": {"bytes": 3 * x_bytes, "flops": 0.0},
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--N", type=int, default=1_000_000)
    ap.add_argument("--precision", choices=["fp64", "mixed", "both"], default="both")
    ap.add_argument("--cpu-maps", type=int, default=480, help="maps the CPU baseline sample runs (about 13 s on 16 threads)")
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "16") or 16))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--mocks", type=int, default=64,
                    help="batched-mock measurement: catalogues with N_b ~ Poisson(mu_b), mu_b ~ U[1e5, 1e6] "
                         "(0 skips it)")
    ap.add_argument("--no-host-path", action="store_true",
                    help="skip the timing of the host-array entry point (sfb_cat2amln)")
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
    rmin, rmax, kmax, nside, niter = 500.0, 1500.0, 0.1, 128, 3
    N = args.N

    t0 = time.perf_counter()
    am = S.AnlmModes(kmax, rmin, rmax, nside=nside)
    t_modes = time.perf_counter() - t0
    log(f"AnlmModes: lmax={am.lmax} nmax={am.nmax} maps={am.nmaps} nlm={S.getnlmsize(am)} ({t_modes:.2f}s)")
    wmodes = S.ConfigurationSpaceModes(rmin, rmax, 1000, nside)
    npix = 12 * nside * nside
    t0 = time.perf_counter()
    win = S.win_rhat_ln(S.SeparableArray(np.ones(1000), np.ones(npix), "phi", "mask"), wmodes, am)
    t_win = time.perf_counter() - t0
    cat = catalogue(N, rmin, rmax, seed=0)
    nbar = N / (4 * np.pi / 3 * (rmax ** 3 - rmin ** 3))

    d_rtp = torch.from_numpy(np.ascontiguousarray(cat.T)).to(dev)
    d_mask = torch.from_numpy(np.ascontiguousarray(win.mask, dtype=np.float64)).to(dev)
    d_wln = torch.from_numpy(np.ascontiguousarray(win.w_ln.reshape(-1, order="F"))).to(dev)
    nlm = S.getnlmsize(am)
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
        torch.cuda.synchronize()
        torch.neg(src, out=dst)
        torch.cuda.synchronize()
        del src, dst

    def run(prec):
        plan.set_precision(prec)
        for _ in range(args.warmup):
            plan.cat2amln_device(N, d_rtp, None, nbar, L.SFB_WIN_SEPARABLE, d_mask, d_wln, d_out)
        plan.profile(True)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            plan.cat2amln_device(N, d_rtp, None, nbar, L.SFB_WIN_SEPARABLE, d_mask, d_wln, d_out)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / args.steps
        phases, calls = plan.phase_times()
        plan.profile(False)
        out = d_out.cpu().numpy()
        log(f"[{prec}] {dt * 1e3:.2f} ms/step  " +
            "  ".join(f"{k}={v[0] / calls:.2f}ms/{v[1] // max(calls, 1)}" for k, v in phases.items()))
        return dt, phases, calls, out[0::2] + 1j * out[1::2]

    results = {}
    precs = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    for prec in precs:
        results[prec] = run(prec)

    # the same call from host arrays (sfb_cat2amln: catalogue upload and a_nlm download
    # over PCIe inside the call), per precision; reported beside `value`, never as it
    # (the catalogue in the reference's layout, 3 x N column-major, which the wrapper passes as is)
    host_path = {}
    cat_cm = np.asfortranarray(cat)
    for prec in ([] if args.no_host_path else precs):
        plan.set_precision(prec)
        for _ in range(max(1, args.warmup)):
            plan.cat2amln(cat_cm, nbar, win, None)
        t = time.perf_counter()
        for _ in range(args.steps):
            plan.cat2amln(cat_cm, nbar, win, None)
        hdt = (time.perf_counter() - t) / args.steps
        host_path[prec] = {"ms": hdt * 1e3, "Mgal_s": N / hdt / 1e6}
        log(f"[{prec} host arrays] {hdt * 1e3:.2f} ms/step ({N / hdt / 1e6:.1f} Mgal/s)")

    # adjoint (sfb_cat2amln_vjp_device, FP64) on the same catalogue: dL/dr, dL/dw for a
    # random cotangent, inputs resident in HBM
    plan.set_precision("fp64")
    d_cot = torch.randn(2 * nlm, dtype=torch.float64, device=dev)
    d_dr = torch.empty(N, dtype=torch.float64, device=dev)
    d_dw = torch.empty(N, dtype=torch.float64, device=dev)
    for _ in range(max(1, args.warmup)):
        plan.cat2amln_vjp_device(N, d_rtp, None, nbar, d_cot, d_dr, d_dw)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        plan.cat2amln_vjp_device(N, d_rtp, None, nbar, d_cot, d_dr, d_dw)
    torch.cuda.synchronize()
    adj_dt = (time.perf_counter() - t) / args.steps
    log(f"[adjoint fp64] {adj_dt * 1e3:.2f} ms/step")

    # batched mocks in the same geometry, generated on the device, one
    # sfb_cat2amln_batch_device call over all of them (FP64, then mixed)
    batch = None
    if args.mocks > 0:
        rng = np.random.default_rng(3)
        N_b = rng.poisson(rng.uniform(1e5, 1e6, args.mocks)).astype(np.int64)
        tot = int(N_b.sum())
        gen = torch.Generator(device=dev)
        gen.manual_seed(2)
        u = torch.rand(3, tot, dtype=torch.float64, device=dev, generator=gen)
        d_mock = torch.stack([torch.pow(u[0] * (rmax ** 3 - rmin ** 3) + rmin ** 3, 1.0 / 3.0),
                              torch.acos(1.0 - 2.0 * u[1]), 2.0 * np.pi * u[2]], dim=1).contiguous()
        del u
        nbars = N_b / (4 * np.pi / 3 * (rmax ** 3 - rmin ** 3))
        d_mout = torch.empty(args.mocks * 2 * nlm, dtype=torch.float64, device=dev)
        batch = {}
        for prec in ("fp64", "mixed"):
            plan.set_precision(prec)
            plan.cat2amln_batch_device(args.mocks, N_b, d_mock, None, nbars, L.SFB_WIN_SEPARABLE, d_mask, d_wln,
                                       d_mout)
            torch.cuda.synchronize()
            t = time.perf_counter()
            plan.cat2amln_batch_device(args.mocks, N_b, d_mock, None, nbars, L.SFB_WIN_SEPARABLE, d_mask, d_wln,
                                       d_mout)
            torch.cuda.synchronize()
            bdt = time.perf_counter() - t
            b = {"mocks": args.mocks, "galaxies": tot, "ms": bdt * 1e3, "ms_per_mock": bdt * 1e3 / args.mocks,
                 "Mgal_s": tot / bdt / 1e6, "path": prec}
            if prec == "fp64":
                batch = b
            else:
                batch["mixed"] = b
            log(f"[batch {prec}] {args.mocks} mocks, {tot} galaxies: {bdt * 1e3:.1f} ms ({tot / bdt / 1e6:.1f} Mgal/s)")
        del d_mock, d_mout

    head = "mixed" if "mixed" in results else "fp64"
    dt, phases, calls, anlm = results[head]
    work = algorithmic_work(am, N, niter, f32=(head == "mixed"))

    # dominant kernel of the headline path and its roofline: the phase with the largest time
    # per kernel (the phi phase brackets two kernels back to back: the equatorial FFT and the
    # polar rings; every other phase launch is one kernel)
    per_launch = {k: (v[0] / max(v[1], 1)) for k, v in phases.items() if v[1] > 0}
    kernels_per_launch = {"phi": 2}
    total = {k: v[0] / kernels_per_launch.get(k, 1) for k, v in phases.items()}
    dom = max((k for k in total if k in work), key=lambda k: total[k])
    w = work[dom]
    t_launch = per_launch[dom] * 1e-3
    if dom in ("analysis", "synthesis"):
        peak = FP64_PEAK_TFLOPS if head == "fp64" else FP32_PEAK_TFLOPS
        achieved = w["flops"] / t_launch / 1e12
        roof = {"kernel": dom, "bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": None}
    else:
        achieved = w["bytes"] / t_launch / 1e9
        roof = {"kernel": dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None}

    # HBM bytes per launch of the dominant kernel from the FETCH_SIZE / WRITE_SIZE
    # counter passes (scaled by the same counters on the calibration copy)
    roof["algorithmic_bytes_per_launch"] = w["bytes"]
    roof["launch_ms"] = t_launch * 1e3
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f).get(head, {}).get(dom)
        if tr:
            roof["traffic"] = tr["fetch_bytes"] + tr["write_bytes"]
            roof["traffic_unit"] = "B/launch"

    # CPU baseline + accuracy on a bounded sample of maps
    cpu = None
    errs = {}
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        om = O.OracleModes.__new__(O.OracleModes)
        om.nside, om.lmax, om.nmax, om.lmax_n = am.nside, am.lmax, am.nmax, am.lmax_n
        om.knot_off, om.knot_x, om.knot_y, om.knot_ypp = am.knot_off, am.knot_x, am.knot_y, am.knot_ypp
        q1 = min(args.cpu_maps, am.nmaps)
        t = time.perf_counter()
        ref = O.cat2amln(om, cat, nbar, 1, win.mask, win.w_ln.reshape(-1, order="F"), q_range=(0, q1),
                         nthreads=args.cpu_threads)
        tc = time.perf_counter() - t
        t_full = tc * am.nmaps / q1
        cpu = {"value": N / t_full / 1e6, "unit": "Mgal/s", "cores": args.cpu_threads, "kind": "port",
               "sample": f"oracle C restatement (OpenMP over maps) on maps 0..{q1 - 1} of {am.nmaps} of the same "
                         f"1e6-galaxy catalogue: {tc:.1f}s, scaled by {am.nmaps}/{q1}"}
        sel = np.zeros(nlm, dtype=bool)
        i = 0
        for n in range(1, am.nmax + 1):
            for l in range(int(am.lmax_n[n - 1]) + 1):
                if am.map_index(n, l) < q1:
                    sel[i:i + l + 1] = True
                i += l + 1
        for prec, (_, _, _, a) in results.items():
            d = np.abs(a[sel] - ref[sel])
            per_mode = d / np.abs(ref[sel])
            errs[prec] = {"max_rel": float(per_mode.max()), "rms_rel": float(np.sqrt(np.mean(per_mode ** 2))),
                          "max_rel_to_max": float(d.max() / np.abs(ref[sel]).max())}
        log(f"cpu baseline: {cpu['value']:.5f} Mgal/s ({tc:.1f}s for {q1} maps); errors {errs}")

    value = N / dt / 1e6
    line = {
        "metric": METRIC, "value": value, "unit": "Mgal/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if head == "fp64" else "f32/f64 mixed", "data": "synthetic",
        "config": {"workload": "N=1e6 uniform shell 500-1500 Mpc/h, kmax=0.1, nside=128, lmax=140, nmax=32, "
                               "2514 maps, niter=3, separable full-sky window",
                   "path": head},
        "fp64_Mgal_s": N / results["fp64"][0] / 1e6 if "fp64" in results else None,
        "mixed_Mgal_s": N / results["mixed"][0] / 1e6 if "mixed" in results else None,
        "speedup_mixed_over_fp64": (results["fp64"][0] / results["mixed"][0])
        if ("mixed" in results and "fp64" in results) else None,
        "max_rel_err": errs.get(head, {}).get("max_rel"),
        "errors_vs_oracle": errs,
        "phases_ms": {k: v[0] / max(calls, 1) for k, v in phases.items()},
        "setup_s": {"AnlmModes": t_modes, "plan": t_plan},
        "phi_tiles": tiles,
        "adjoint_fp64_ms": adj_dt * 1e3,
        "host_arrays_path": host_path,
        "batched_mocks": batch,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
