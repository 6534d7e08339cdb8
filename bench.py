"""Throughput of the derivative-carrying forward pass + exact adjoint on one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A step is one pass of the hot path over the workload's whole batch, inputs
resident in HBM: input layer, forward (linear map + Faa di Bruno step per
layer), collocation seed of the mean-square residual loss, adjoint
(pair sums, weight-gradient GEMM, transposed linear map per layer) and
accumulation of the batch gradient.  It runs once in FP64 (the drop-in
parity mode) and once per mixed-precision mode; the headline value is the
mixed-precision (BF16 hi+lo split MFMA) points/s, reported with its error
against the FP64 run per derivative order.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "fwd+adjoint points/s, mixed-prec vs FP64 on GPU, at stated err vs FP64 oracle"
PEAK = {"hbm": 8000.0, "fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}
DTYPE = {"fp32": "f32 (FP64-accumulated forward map)",
         "bf16x3": "f32 storage, bf16 hi+lo split MFMA",
         "bf16x6": "f32 storage, bf16 three-way split MFMA (FP64-accumulated forward map)"}
TRAFFIC = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "hbm_traffic.json")


def traffic_of(workload, mode, category):
    """HBM bytes per launch of a kernel category, as measured by the
    FETCH_SIZE / WRITE_SIZE counter passes summarised in profiles/ (scaled
    by the same counters' reading of a plain copy); None if not measured."""
    try:
        with open(TRAFFIC) as f:
            t = json.load(f)
        return t["workloads"][workload][mode][category]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def algorithmic(wl, na, F, Q, P, N, wbytes, keep_s):
    """Per-step algorithmic work of each kernel category: (flops, bytes).

    K1 and K5 are 2 n_l (n_{l-1}+1) |A| flops per point and layer, K6
    2 n_l n_{l-1} |A| for l > 2; K1 reads T of the layer below and writes S,
    K5 reads S-bar and T of the layer below, K6 reads S-bar and writes T-bar
    (|A| words per row and point).  Keeping h, the Faa di Bruno forward is
    n (3F + 4Q + 3(|A|-1)) flops and reads S, writes T and h; the adjoint is
    n (3P + 2|A|) flops and reads T-bar and h, writes S-bar.  Keeping S
    (keep_s), the forward is n (3F + 2Q + (|A|-1)) flops and reads S, writes
    T; the adjoint recomputes h (n (3F + 2Q + (|A|-1)) flops more) and reads
    T-bar and S, writes S-bar."""
    d = wl.dims
    L = len(d)
    w = float(wbytes) * na * N
    lin = sum(2.0 * d[l] * (d[l - 1] + 1) * na for l in range(1, L)) * N
    lin_b = sum(d[l - 1] + d[l] for l in range(1, L)) * w
    lin_t = sum(2.0 * d[l] * d[l - 1] * na for l in range(2, L)) * N
    lin_t_b = sum(d[l] + d[l - 1] for l in range(2, L)) * w
    hid = sum(d[l] for l in range(1, L - 1))
    bb_flops = hid * (3.0 * P + 2.0 * na) * N
    if keep_s:
        h_flops = hid * (3.0 * F + 2.0 * Q + (na - 1)) * N
        bell_fwd, bell_bwd = (h_flops, 2.0 * hid * w), (bb_flops + h_flops, 3.0 * hid * w)
    else:
        bell_fwd = (hid * (3.0 * F + 4.0 * Q + 3.0 * (na - 1)) * N, 3.0 * hid * w)
        bell_bwd = (bb_flops, 3.0 * hid * w)
    return {"linear": (lin, lin_b), "linear_t": (lin_t, lin_t_b), "grad": (lin, lin_b),
            "bell_fwd": bell_fwd, "bell_bwd": bell_bwd}


def rel_by_order(x, ref, deg):
    out = {}
    for p in sorted(set(int(v) for v in deg)):
        sel = np.asarray(deg) == p
        r = float(np.abs(ref[..., sel]).max())
        e = np.abs(x[..., sel] - ref[..., sel])
        out[str(p)] = {"max": float(e.max() / max(r, 1e-300)), "median": float(np.median(e) / max(r, 1e-300))}
    return out


def bench_ngd(args, D, torch, wl, net, xd, res, N, dev, mu=1e-3):
    """One natural-gradient step of the NGD workload: Gauss-Newton rows dR/dw
    of every point (FP64 adjoint, rows stored FP32), the loss gradient b,
    G = J^T J on FP32 MFMA and the FP64 Cholesky solve of (G/N + mu I) u = b/N
    (app/optimizer_module.f90:221-469 with ABS damping)."""
    bw = D.BatchWork(net, N, D.PREC_FP64)
    g = D.Grad(net)
    s = torch.cuda.Stream(device=dev)
    bw.set_stream(s.cuda_stream)
    ld = ((N + 63) // 64) * 64
    rows = torch.zeros((net.nw, ld), dtype=torch.float32, device=dev)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    acc = np.zeros(4)
    with torch.cuda.stream(s):
        for it in range(args.warmup + args.steps):
            ev[0].record(s)
            bw.residual_rows(xd, res, rows, ld)
            ev[1].record(s)
            g.zero()
            bw.residual_step(xd, res, 1.0, g, want_rsq=False)
            ev[2].record(s)
            b = torch.from_numpy(g.get()).to(dev)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            u = bw.ngd_solve(rows, ld, N, b, N, mu, 1)
            t_solve = time.perf_counter() - t0
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ud = bw.ngd_solve(rows, ld, N, b, N, mu, 1, dual=True)
            t_dual = time.perf_counter() - t0
            if it >= args.warmup:
                acc += [ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2]), t_solve * 1e3, t_dual * 1e3]
    acc /= args.steps
    ms = float(acc[:3].sum())
    nw = net.nw
    gram_flop = 2.0 * nw * nw * N
    chol_flop = nw ** 3 / 3.0
    line = {"metric": "NGD step rows/s (Gauss-Newton rows + gradient + J^T J + FP64 Cholesky solve)",
            "value": N / (ms * 1e-3), "unit": "rows/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64 rows/solve, f32 metric", "data": "synthetic",
            "config": {"workload": f"{wl.name}: MLP {wl.dims} tanh, D0={wl.d0} K={wl.k}, N={N}, nw={nw}, "
                                   f"rows of {wl.note}, ABS damping {mu}"},
            "breakdown_ms": {"rows": acc[0], "gradient": acc[1], "metric_and_solve": acc[2]},
            "dual_route": {"metric_and_solve_ms": acc[3], "ms_per_step": float(acc[0] + acc[1] + acc[3]),
                           "rows_per_s": N / (float(acc[0] + acc[1] + acc[3]) * 1e-3),
                           "rel_diff_to_primal": float((torch.as_tensor(ud) - torch.as_tensor(u)).norm() /
                                                       torch.as_tensor(u).norm())},
            "metric_solve_flops": {"gram": gram_flop, "cholesky": chol_flop},
            "u_norm": float(torch.as_tensor(u).norm())}
    print(json.dumps(line), flush=True)
    bw.set_stream(None)
    return 0


def bench_kf(args, D, torch, wl, net, w, x, res, terms, dev):
    """The extended Kalman filter over the weights: one step presents the
    workload's npt collocation points in order as observations -- forward
    and adjoint of the point, R and the seed dR/dT on the device, the
    covariance pass (pending downdate + P j), gain and weight step
    (app/sgd_batch_module.f90:547-550, lib/kalman_module.f90:240-493) --
    and returns the weights to the host once."""
    p0, lam_ini, lam0, dec = wl.kf
    nobs = min(wl.npt, args.kf_obs) if args.kf_obs else wl.npt
    x = x[:nobs]
    nw = net.nw
    xd = torch.from_numpy(x).to(dev)

    def fresh():
        net.weights = w
        return D.KalmanFilter(net, p0, lam_ini, lam0, dec)

    bw = D.BatchWork(net, 1, D.PREC_FP64)
    s = torch.cuda.Stream(device=dev)
    bw.set_stream(s.cuda_stream)
    kf = fresh()
    with torch.cuda.stream(s):
        for _ in range(args.warmup):
            kf.update_residual(bw, xd, res)
        torch.cuda.synchronize()
        # inside the timed region the roofline's kernel (the covariance pass)
        # is bracketed by HIP events on the work space's stream in the last
        # step
        bw.kernel_ms()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(s)
        for i in range(args.steps):
            if i == args.steps - 1:
                bw.profile(True, only="kf_pass")
            kf.update_residual(bw, xd, res)
        e1.record(s)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / args.steps
        ms = e0.elapsed_time(e1) / args.steps
        bw.profile(False)
        kms_timed = bw.kernel_ms()
        # one more step with every kernel category bracketed by HIP events
        bw.profile(True)
        kf.update_residual(bw, xd, res)
        torch.cuda.synchronize()
        bw.profile(False)
        kms = bw.kernel_ms()
        kms["kf_pass"] = kms_timed["kf_pass"]   # one step, as the breakdown
    st = kf.state()
    kf.free()

    # the dominant kernel: the covariance pass, one launch per observation
    tot, cnt = kms["kf_pass"]
    launch_ms = tot / max(cnt, 1)
    tri = nw * (nw + 1) / 2.0 if not dec else float(sum(
        wl.dims[l] * (wl.dims[l - 1] + 1) ** 2 for l in range(1, len(wl.dims))))
    byts = 2.0 * 8.0 * tri          # read and write of the stored covariance
    roof = {"bound": "hbm", "achieved": byts / (launch_ms * 1e-3) / 1e9, "peak": PEAK["hbm"], "unit": "GB/s",
            "frac": byts / (launch_ms * 1e-3) / 1e9 / PEAK["hbm"], "traffic": traffic_of(wl.name, "fp64", "kf_pass"),
            "kernel": "k_kf_tiles" if not dec else "k_kf_blocks", "launch_ms": launch_ms,
            "algorithmic_bytes_per_launch": byts}

    # parity on the first observations against the oracle, and the CPU
    # baseline: the reference's own kalman_module (oracle/_ref) on a sample
    check, cpu = None, None
    try:
        import oracle
        if not os.path.exists(oracle.LIB):
            oracle.build(ref=False)
        o = oracle.Oracle(wl.d0, wl.k, wl.seeds, wl.act)
        m = 40 if nw < 2000 else 3
        kf = fresh()
        kf.update_residual(bw, x[:m], res)
        wg = net.weights
        Pg = kf.covariance() if nw < 2000 else None
        kf.free()
        seeds0 = np.zeros((m, 1, o.na))
        wr, Pr, _, _ = o.kf_seq(wl.dims, w, p0, lam_ini, lam0, dec, 0.0, 0.0, 3, x[:m], seeds0, np.zeros(m),
                                terms=terms)
        check = {"observations": m, "weights_max_rel": float(np.abs(wg - wr).max() / np.abs(wr).max())}
        if Pg is not None:
            check["cov_max_rel"] = float(np.abs(Pg - Pr).max() / np.abs(Pr).max())
        if not args.no_cpu:
            ref = oracle.RefLib.available()

            def cpu_run(k):
                if ref:
                    # the reference's Fortran runtime writes to stdout at
                    # exit: run it in a child process whose output is dropped
                    import subprocess
                    import tempfile
                    with tempfile.TemporaryDirectory() as td:
                        out = os.path.join(td, "t.json")
                        subprocess.run([sys.executable, os.path.abspath(__file__), "--workload", wl.name,
                                        "--kf-ref-timing", str(k), "--out", out], check=True,
                                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
                        with open(out) as f:
                            return json.load(f)["seconds"]
                t0 = time.perf_counter()
                o.kf_seq(wl.dims, w, p0, lam_ini, lam0, dec, 0.0, 0.0, 3, x[:k], np.zeros((k, 1, o.na)),
                         np.zeros(k), terms=terms)
                return time.perf_counter() - t0
            k0 = 2
            dt = cpu_run(k0)
            k = int(min(nobs, max(k0, k0 * args.cpu_seconds / max(dt, 1e-6))))
            dt = cpu_run(k)
            cpu = {"value": k / dt, "unit": "observations/s", "cores": 1, "kind": "reference" if ref else "port",
                   "sample": f"first {k} of the {nobs} observations, "
                             + ("the reference's kalman_module (lib/ compiled with amdflang -O3) through "
                                "ref_driver.f90" if ref else "the oracle's C restatement")
                             + f", dense covariance, 1 thread, {dt:.1f} s (filter set-up included)"}
    except Exception as exc:   # a reported baseline, never the product
        cpu = cpu or {"error": repr(exc)}
    bw.set_stream(None)
    bw.free()
    cov_bytes = st["bytes"]
    line = {"metric": "Kalman filter observations/s (collocation residual rows, dense covariance update)",
            "value": nobs / (ms * 1e-3), "unit": "observations/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"{wl.name}: MLP {wl.dims} tanh, D0={wl.d0} K={wl.k}, nw={nw}, {nobs} "
                                   f"observations of {wl.note} per step, p0={p0} lam={lam_ini} lam0={lam0}, "
                                   f"{'decoupled' if dec else 'dense'} covariance {cov_bytes / 1e9:.3f} GB"},
            "roofline": roof, "cpu_baseline": cpu, "vs_oracle": check,
            "wall_ms_per_step": wall * 1e3,
            "kernels": {c: {"ms_per_step": t, "launches_per_step": n} for c, (t, n) in kms.items() if n},
            "us_per_observation": ms * 1e3 / nobs}
    print(json.dumps(line), flush=True)
    return 0


def kf_ref_timing(args):
    """CPU baseline leg of the filter workloads (a child process, no GPU):
    the first k observations through the reference's own kalman_module,
    compiled from its sources into oracle/_ref; the seconds go to --out."""
    import oracle
    from dnnf90_amd.workloads import WORKLOADS
    wl = WORKLOADS[args.workload]
    p0, lam_ini, lam0, dec = wl.kf
    w, x = wl.inputs()
    o = oracle.Oracle(wl.d0, wl.k, wl.seeds, wl.act)
    k = args.kf_ref_timing
    alpha = o.alpha()
    terms = wl.terms([list(a) for a in alpha])
    runner = oracle.RefNet(wl.d0, wl.k, wl.seeds, wl.act, wl.dims, w)
    t0 = time.perf_counter()
    runner.kf_seq(p0, lam_ini, lam0, dec, 0.0, 0.0, 3, x[:k], np.zeros((k, 1, o.na)), np.zeros(k), terms=terms)
    dt = time.perf_counter() - t0
    with open(args.out, "w") as f:
        json.dump({"seconds": dt, "observations": k}, f)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="dense4_k3")
    ap.add_argument("--mixed", default="bf16x6", choices=["bf16x6", "fp32", "bf16x3"])
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only copy a buffer of this many MiB 5 times (HBM counter calibration)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kf-obs", type=int, default=0, help="filter workloads: observations per step")
    ap.add_argument("--kf-ref-timing", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--out", default="", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.kf_ref_timing:
        return kf_ref_timing(args)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the gradient all-reduce across GPUs is a design note, not a measured path)")
        return 2

    import torch
    if args.copy_mib:
        n = args.copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_mib": args.copy_mib, "launches": 5}), flush=True)
        return 0
    import dnnf90_amd as D
    from dnnf90_amd.workloads import WORKLOADS
    from dnnf90_amd._lib import lib

    wl = WORKLOADS[args.workload]
    N = wl.npt
    ts = D.TableSet(wl.d0, wl.k, wl.seeds, wl.act)
    net = D.Net(ts, wl.dims)
    w, x = wl.inputs()
    net.weights = w
    alpha = [a[:wl.d0] for a in D.current_alpha()]
    terms = wl.terms(alpha)
    res = D.Residual(terms)
    deg = np.array([sum(a) for a in alpha])
    na = len(alpha)
    counts = np.zeros(4, np.int32)
    lib().dnnf90_tabset_counts(ts.id, counts.ctypes.data)
    F, P = int(counts[2]), int(counts[1])
    Q = int(sum(max(int(p) - 1, 0) for p in deg))
    dev = torch.device("cuda:0")
    xd = torch.from_numpy(x).to(dev)
    fac = 2.0 / N

    if args.workload == "ngd2_k4":
        return bench_ngd(args, D, torch, wl, net, xd, res, N, dev)
    if wl.kf is not None:
        return bench_kf(args, D, torch, wl, net, w, x, res, terms, dev)

    def run_mode(prec, name, wbytes, mma_peak, budget=0):
        bw = D.BatchWork(net, N, prec, budget=budget)
        info = bw.info()
        kinfo = np.zeros(6, np.int32)
        lib().dnnf90_net_kernel_info(net.id, prec, kinfo.ctypes.data)
        keep_s = int(kinfo[0]) == 2
        g = D.Grad(net)
        s = torch.cuda.Stream(device=dev)
        bw.set_stream(s.cuda_stream)
        with torch.cuda.stream(s):
            for _ in range(args.warmup):
                bw.residual_step(xd, res, fac, g, want_rsq=False)
            torch.cuda.synchronize()
            # inside the timed region the roofline's kernel (the Faa di Bruno
            # adjoint) is bracketed by HIP events on the work space's stream
            # in the last step (the markers add launch gaps; one step's worth
            # spread over K steps)
            bw.kernel_ms()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            t0 = time.perf_counter()
            e0.record(s)
            for i in range(args.steps):
                if i == args.steps - 1:
                    bw.profile(True, only="bell_bwd")
                bw.residual_step(xd, res, fac, g, want_rsq=False)
            e1.record(s)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
            ms = e0.elapsed_time(e1) / args.steps
            bw.profile(False)
            kms_timed = bw.kernel_ms()
            # the per-kernel breakdown: the same steps again, untimed, with
            # every kernel category bracketed (the markers add launch gaps,
            # so the timed steps carry only the roofline kernel's)
            bw.profile(True)
            for _ in range(args.steps):
                bw.residual_step(xd, res, fac, g, want_rsq=False)
            torch.cuda.synchronize()
            bw.profile(False)
            kms = bw.kernel_ms()
            kms["bell_bwd"] = tuple(v * args.steps for v in kms_timed["bell_bwd"])   # one step -> K
            tout = bw.forward(xd)
            torch.cuda.synchronize()
        g.zero()
        rsq = bw.residual_step(xd, res, fac, g)
        grad = g.get()
        t_host = tout.cpu().numpy()
        bw.set_stream(None)
        bw.free()
        g.free()
        work = algorithmic(wl, na, F, Q, P, N, wbytes, keep_s)
        kern = {}
        for c, (tot, cnt) in kms.items():
            ent = {"ms_per_step": tot / args.steps, "launches_per_step": cnt / args.steps}
            if c in work and tot > 0:
                # judged against the roof that binds: MFMA for the linear
                # maps unless their bytes take longer at 8 TB/s than their
                # flops at the mode's MFMA peak; HBM for the Faa di Bruno kernels
                flops, byts = work[c]
                sec = tot / args.steps * 1e-3
                mfma_bound = c in ("linear", "linear_t", "grad") and flops / (mma_peak * 1e12) > byts / (PEAK["hbm"] * 1e9)
                if mfma_bound:
                    ent.update(bound="mfma", achieved=flops / sec / 1e12, unit="TFLOP/s", peak=mma_peak)
                else:
                    ent.update(bound="hbm", achieved=byts / sec / 1e9, unit="GB/s", peak=PEAK["hbm"])
                ent["frac"] = ent["achieved"] / ent["peak"]
                ent["algorithmic_flops_per_step"] = flops
                ent["algorithmic_bytes_per_step"] = byts
            kern[c] = ent
        info["keep_s"] = keep_s
        return {"name": name, "ms_per_step": ms, "wall_ms_per_step": wall * 1e3 / args.steps, "info": info,
                "points_per_s": N / (ms * 1e-3), "kernels": kern, "T": t_host, "grad": grad, "rsq": rsq}

    fp64 = run_mode(D.PREC_FP64, "fp64", 8, PEAK["fp64"])
    others = {}
    modes = (("fp32", D.PREC_FP32, PEAK["fp32"]), ("bf16x3", D.PREC_BF16X3, PEAK["bf16"] / 3.0),
             ("bf16x6", D.PREC_BF16X6, PEAK["bf16"] / 6.0))
    for name, prec, peak in modes:
        r = run_mode(prec, name, 4, peak)
        r["err_T"] = rel_by_order(r["T"], fp64["T"], deg)
        r["err_g"] = float(np.abs(r["grad"] - fp64["grad"]).max() / np.abs(fp64["grad"]).max())
        others[name] = r
    mixed = others[args.mixed]
    err_T, err_g = mixed["err_T"], mixed["err_g"]
    # the same mixed step with per-layer recompute of h in the reverse pass
    mprec = {m[0]: m[1] for m in modes}[args.mixed]
    mpeak = {m[0]: m[2] for m in modes}[args.mixed]
    rc_run = run_mode(mprec, args.mixed + "+recompute", 4, mpeak, budget=1)
    recompute = {"mode": args.mixed, "ms_per_step": rc_run["ms_per_step"],
                 "points_per_s": rc_run["points_per_s"],
                 "identical_grad": bool(np.array_equal(rc_run["grad"], mixed["grad"])),
                 "device_bytes": rc_run["info"]["bytes"], "device_bytes_all_kept": rc_run["info"]["bytes_stored"]}

    # FP64 GPU against the oracle on a sample of points
    oracle_check = None
    cpu = None
    try:
        import oracle
        if not os.path.exists(oracle.LIB):
            oracle.build(ref=False)
        o = oracle.Oracle(wl.d0, wl.k, wl.seeds, wl.act)
        ns = min(N, 256)
        to = o.eval(wl.dims, w, x[:ns])
        oracle_check = {"points": ns, "forward_max_rel": max(v["max"] for v in rel_by_order(fp64["T"][:ns], to, deg).values())}
        if not args.no_cpu:
            ncores = os.cpu_count() or 1
            n0 = 64
            t0 = time.perf_counter()
            o.pinn_step(wl.dims, w, x[:n0], terms, fac)
            dt = time.perf_counter() - t0
            ns = int(min(N, max(n0, n0 * args.cpu_seconds / max(dt, 1e-6))))
            t0 = time.perf_counter()
            o.pinn_step(wl.dims, w, x[:ns], terms, fac)
            dt = time.perf_counter() - t0
            cpu = {"value": ns / dt, "unit": "points/s", "cores": ncores, "kind": "port",
                   "sample": f"first {ns} of the {N} points, oracle FP64 fwd+seed+adjoint+accumulate, "
                             f"OpenMP over points, {ncores} threads, {dt:.1f} s"}
    except Exception as exc:  # the CPU leg is a reported baseline, never the product
        cpu = {"error": repr(exc)}

    # the roofline line names one kernel: the dominant category whose
    # launches in a step are all the same kernel (the generated Faa di Bruno
    # kernels; the linear-map and gradient categories mix instantiations for
    # the input, hidden and output layers -- every category is in "kernels")
    kk = mixed["kernels"]
    dom = max((c for c in ("bell_fwd", "bell_bwd") if "frac" in kk.get(c, {})), key=lambda c: kk[c]["ms_per_step"])
    dk = kk[dom]
    roof = {"bound": dk["bound"], "achieved": dk["achieved"], "peak": dk["peak"], "unit": dk["unit"],
            "frac": dk["frac"], "traffic": traffic_of(wl.name, args.mixed, dom), "kernel": dom,
            "launch_ms": dk["ms_per_step"] / max(dk["launches_per_step"], 1e-9)}

    line = {
        "metric": METRIC, "value": mixed["points_per_s"], "unit": "points/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": mixed["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPE[args.mixed], "data": "synthetic",
        "config": {"workload": f"{wl.name}: MLP {wl.dims} {'tanh' if wl.act == 0 else 'erf-GELU'}, "
                               f"D0={wl.d0} K={wl.k} |A|={na}, N={N}, loss mean square of {wl.note}",
                   "precision": args.mixed},
        "roofline": roof,
        "cpu_baseline": cpu,
        "fp64": {"value": fp64["points_per_s"], "ms_per_step": fp64["ms_per_step"], "kernels": fp64["kernels"],
                 "vs_oracle": oracle_check},
        "speedup_vs_fp64": mixed["points_per_s"] / fp64["points_per_s"],
        "err_vs_fp64": {"forward_by_order": err_T, "grad_max_rel": err_g},
        "kernels": mixed["kernels"],
        "kept_plane": {"fp64": "S" if fp64["info"]["keep_s"] else "h",
                       args.mixed: "S" if mixed["info"]["keep_s"] else "h"},
        "recompute": recompute,
        "modes": {k: {"points_per_s": r["points_per_s"], "ms_per_step": r["ms_per_step"],
                      "speedup_vs_fp64": r["points_per_s"] / fp64["points_per_s"],
                      "max_err_by_order": {o: v["max"] for o, v in r["err_T"].items()},
                      "grad_max_rel": r["err_g"],
                      "kernel_ms": {c: v["ms_per_step"] for c, v in r["kernels"].items()}}
                  for k, r in others.items()},
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
