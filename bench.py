"""Benchmark: the n=8000 FP64 two-stage HT reduction on one MI355X.

A step is one full reduction (stage 1 ab2rht_simple + stage 2) of one
synthetic FP64 pencil (A ~ N(0,1), B = R of a QR of an N(0,1) matrix,
Q = Z = I), generated on the GPU and kept resident in HBM; each step restores
A, B, Q, Z from resident copies and reduces them in place.

Prints one JSON line: seconds per reduction (value), GFLOP/s and % of the FP64
peak by the 14 n^3 convention, per-stage times, backward errors of the last
step, a roofline entry for the dominant throughput kernel measured with HIP
events over the timed steps, and the CPU baseline (the C oracle's
restatement of ParaHT, timed on a bounded sample on the host cores).
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "n=8000 FP64 HT-reduce s & %FP64 peak; batched pencils/s; fwd+bwd/fwd ratio"
FP64_PEAK = 78.6e12  # MI355X dense FP64 (matrix and vector), FLOP/s
HBM_PEAK = 8.0e12    # bytes/s

# kernel classes timed by the library -> (kernel, bound).  The throughput
# kernels are bound by FP64 MFMA (the compact-WY applier) or by HBM (the
# stage-2 far updates); the generators are chains of dependent steps, bound
# by latency, and are reported per step instead of against a roofline.
KERNEL_CLASSES = {
    "wy_apply": ("k_apply_blocks_lds (compact-WY block reflectors, FP64 MFMA)", "mfma"),
    "s1_panel_qr": ("k_s1_panel_qr_tau + k_s1_panel_t (panel QR chain + T)", "latency"),
    "s1_rq_gen": ("k_s1_rq_fast (RQ + LQ right generator + T)", "latency"),
    "s2_gen": ("k_hh_sweep (pipelined Householder bulge chase of one sweep)", "latency"),
    "s2_rows": ("k_hh_far (compact-WY reflectors on rows of Q, Z and of A, B above the "
                "window, FP64 MFMA)", "hbm"),
    "s2_cols": ("k_hh_far<cols> (compact-WY reflectors on columns of A, B right of the "
                "window, FP64 MFMA)", "hbm"),
}
# the same classes for the Givens stage 2 (rht2ht_givens)
KERNEL_CLASSES_GIVENS = dict(KERNEL_CLASSES, **{
    "s2_gen": ("k_s2_sweep (pipelined Givens bulge chase of one sweep)", "latency"),
    "s2_rows": ("k_s2_rows_w2 (rotations on rows of Q, Z and of A, B above the window)", "hbm"),
    "s2_cols": ("k_s2_cols_w2 (rotations on columns of A, B right of the window)", "hbm"),
})


def chase_steps(n, r, nq):
    """Dependent steps of the pipelined sweeps of one stage 2: a sweep of nq
    columns over nblocks bulge blocks takes nblocks + 3 (nq - 1) steps (each
    column trails the previous one by the two-step lag and its own step)."""
    steps = 0
    for s in range(1, n - 1, nq):
        nbl = (n - s) // r + ((n - s) % r != 0)
        steps += nbl + 3 * (min(nq, n - 1 - s) - 1)
    return steps


def traffic_per_launch(cls, args):
    """HBM bytes per launch of a kernel class from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this workload (scaled by the copy-kernel
    calibration), or None when no summary exists."""
    path = os.path.join(ROOT, "profiles",
                        f"traffic_{cls}_n{args.n}_r{args.r}_p{args.p}_nq{args.nq}.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_dispatch:"):
                return float(line.split(":", 1)[1])
    return None


def cpu_baseline(args, gpu_cfg0_s):
    """CPU baselines on the host cores, timed in this run on the same
    distribution of pencils (A ~ N(0,1), B = R of a QR of an N(0,1) matrix):
    the oracle's restatement of ParaHT (stage 1 + the Householder stage 2 the
    reference driver calls, OpenMP on all host cores) and LAPACK DGGHD3
    (scipy's OpenBLAS, and MKL where the host has it).  Measured at n = 1000
    (the reference README's configuration, r = 16, p = 8, nq = 8, next to this
    build's GPU time for it) and at n = cpu_n with the benchmark's (r, p, nq);
    `value` is the restatement's n = cpu_n time extrapolated by (n/cpu_n)^3 to
    the benchmark's n (labelled as such)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from paraht_amd.pencils import random_pencil

    def timed(f, *a):
        t0 = time.perf_counter()
        f(*a)
        return time.perf_counter() - t0

    measured = {}
    A0, B0 = random_pencil(1000)
    measured["n1000_paraht_restatement_s"] = timed(O.reduce, A0, B0, 16, 8, 8, "householder")
    measured["n1000_dgghd3_openblas_s"] = timed(O.dgghd3, A0, B0)
    t0 = time.perf_counter()
    if O.dgghd3_mkl(A0, B0) is not None:  # None: no MKL on this host
        measured["n1000_dgghd3_mkl_s"] = time.perf_counter() - t0
    ns = args.cpu_n
    A0, B0 = random_pencil(ns)
    t = timed(O.reduce, A0, B0, args.r, args.p, args.nq, "householder")
    measured[f"n{ns}_paraht_restatement_s"] = t
    measured[f"n{ns}_dgghd3_openblas_s"] = timed(O.dgghd3, A0, B0)
    print(f"cpu baseline: restatement n={ns} {t:.2f} s", file=sys.stderr, flush=True)
    scaled = t * (args.n / ns) ** 3
    return {"value": scaled, "unit": "s", "cores": O.num_threads(), "kind": "port",
            "extrapolated": True,
            "sample": f"oracle ab2rht_simple + rht2ht_householder at n={ns}, r={args.r}, "
                      f"p={args.p}, nq={args.nq}: {t:.3f} s measured on {O.num_threads()} "
                      f"host threads; value = that time x (n/{ns})^3 for n={args.n}",
            "measured": measured,
            "cfg0_gpu_s": gpu_cfg0_s,
            "cfg0_reference_readme_s": {"paraht": 1.208396, "dgghd3": 3.635724}}


def bench_cfg0(args):
    """This build's GPU time for the reference README's configuration: one
    pencil of n = 1000, r = 16, p = 8, nq = 8, resident in HBM (best of 3 after
    one untimed call), with the Householder stage 2 (the variant the README
    times) and with the Givens one."""
    import torch

    import paraht_amd as ph
    n = 1000
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(1302)
    A0 = ph.column_major(torch.randn(n, n, dtype=torch.float64, device=dev, generator=g))
    R = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device=dev, generator=g),
                        mode="r")[1]
    B0 = ph.column_major(torch.triu(R))
    ctx = ph.Context(0)
    out = {}
    for name, givens in (("householder", False), ("givens", True)):
        best = None
        for _ in range(4):
            A, B = A0.clone(), B0.clone()
            Q = ph.column_major(torch.eye(n, dtype=torch.float64, device=dev))
            Z = Q.clone()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ph.reduce(A, B, 16, 8, 8, Q, Z, ctx=ctx, givens=givens)
            t = time.perf_counter() - t0
            best = t if best is None else min(best, t)
        out[name] = best
    return out


def bench_batched(args):
    """Batched path: 2048 FP64 pencils with n uniform in [32, 512] (A ~ N(0,1),
    B = R of the QR of an N(0,1) matrix, Q = Z = I), reduced in one
    HIP-graph-captured persistent launch with the size-bucket (r, p, nq);
    pencils/s over 3 timed launches, against a loop of the single-pencil
    path over the first `batched_loop_sample` pencils of the same draw."""
    import numpy as np
    import torch

    import paraht_amd as ph
    from paraht_amd.batched import Batch, defaults, reduce_batch
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(1302)
    sizes = [int(x) for x in rng.integers(32, 513, size=2048)]
    g = torch.Generator(device=dev)
    g.manual_seed(1302)
    As, Bs = [], []
    for n in sizes:
        As.append(torch.randn(n, n, dtype=torch.float64, device=dev, generator=g))
        R = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device=dev, generator=g),
                            mode="r")[1]
        Bs.append(torch.triu(R))
    bt = Batch(sizes)
    desc = bt.descriptors()
    bt.load(As, Bs)
    reduce_batch(bt, desc=desc, graph=True)  # warmup (and graph capture)
    ts = []
    for _ in range(3):
        bt.load(As, Bs)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        reduce_batch(bt, desc=desc, graph=True)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    t_b = sum(ts) / len(ts)
    worst = 0.0
    for i in range(0, len(sizes), 64):  # backward errors of every 64th pencil
        H, T, Q, Z = (bt.view(M, i) for M in (bt.A, bt.B, bt.Q, bt.Z))
        ea = (torch.linalg.norm(Q.T @ As[i] @ Z - H) / torch.linalg.norm(As[i])).item()
        eb = (torch.linalg.norm(Q.T @ Bs[i] @ Z - T) / torch.linalg.norm(Bs[i])).item()
        worst = max(worst, ea, eb)
    k = min(args.batched_loop_sample, len(sizes))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(k):
        r, p, nq = defaults(sizes[i])
        if r == 1:  # the one-stage path exists only in the batched kernel
            r, p, nq = 8, 2, 4
        ph.reduce(ph.column_major(As[i].clone()), ph.column_major(Bs[i].clone()), r, p, nq)
    torch.cuda.synchronize()
    t_l = time.perf_counter() - t0
    # the loop's rate scaled by the sample's share of the 14 n^3 work
    work_share = sum(n ** 3 for n in sizes[:k]) / (sum(n ** 3 for n in sizes) / len(sizes) * k)
    loop_rate = k / t_l * work_share
    rate = len(sizes) / t_b
    print(f"batched: {len(sizes)} pencils in {t_b:.3f} s; loop {k} pencils in {t_l:.3f} s",
          file=sys.stderr, flush=True)
    return {"metric": "batched pencils/s", "pencils": len(sizes), "n_range": [32, 512],
            "s": t_b, "pencils_per_s": rate, "gflops": sum(14.0 * n ** 3 for n in sizes) / t_b / 1e9,
            "params": "size-bucket defaults " + str(sorted({defaults(n) for n in sizes})),
            "launch": "one HIP-graph-captured persistent launch",
            "loop_baseline": {"pencils": k, "s": t_l, "pencils_per_s": loop_rate,
                              "note": "single-pencil path per pencil, rate scaled to the "
                                      "batch's mean n^3"},
            "speedup_vs_loop": rate / loop_rate,
            "max_backward_error_every_64th": worst}


def bench_adjoint(args):
    """Reverse mode: 64 FP64 pencils with n=256 (batched forward, one
    workgroup per pencil in the adjoint) plus one pencil with n=2048
    (single-pencil forward, cooperative-grid adjoint), A ~ N(0,1), B = R of
    the QR of an N(0,1) matrix; loss = sum W*H + V*T + P*Q + S*Z with W, V,
    P, S ~ N(0,1).  Forward time and forward + backward time (best of 3)."""
    import torch

    from paraht_amd.adjoint import ht_reduce
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(4)

    def pencils(b, n):
        A = torch.randn(b, n, n, dtype=torch.float64, device=dev, generator=g)
        R = torch.linalg.qr(torch.randn(b, n, n, dtype=torch.float64, device=dev, generator=g),
                            mode="r")[1]
        return A, torch.triu(R)

    parts = [pencils(64, 256), pencils(1, 2048)]
    parts = [parts[0], (parts[1][0][0], parts[1][1][0])]
    weights = [[torch.randn(A.shape, dtype=torch.float64, device=dev, generator=g)
                for _ in range(4)] for A, _ in parts]

    def run(backward):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for (A0, B0), (W, V, P, S) in zip(parts, weights):
            A = A0.detach().requires_grad_(backward)
            B = B0.detach().requires_grad_(backward)
            H, T, Q, Z = ht_reduce(A, B)
            if backward:
                loss = ((torch.triu(W, -1) * H).sum() + (torch.triu(V) * T).sum()
                        + (P * Q).sum() + (S * Z).sum())
                torch.autograd.grad(loss, (A, B))
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    run(False)
    run(True)
    t_f = min(run(False) for _ in range(3))
    t_fb = min(run(True) for _ in range(3))
    print(f"adjoint: forward {t_f:.3f} s, forward+backward {t_fb:.3f} s",
          file=sys.stderr, flush=True)
    return {"workload": "64 pencils n=256 + 1 pencil n=2048; loss sum W*H+V*T+P*Q+S*Z",
            "forward_s": t_f, "forward_backward_s": t_fb, "ratio": t_fb / t_f}


def bench_qz(args):
    """The QZ iteration after the reduction (paraht_hgeqz*): generalized
    eigenvalues of (1) single pencils of order 1000 and 2000 (A ~ N(0,1),
    B = R of the QR of an N(0,1) matrix, reduced on the GPU first, untimed),
    eigenvalues only and, at 1000, the Schur form with Q, Z; (2) the batched draw of
    2048 pencils with n uniform in [32, 512] after their batched reduction,
    eigenvalues only, in one launch.  CPU baseline: LAPACK DHGEQZ (scipy's
    OpenBLAS) on the order-1000 pencil and on every 64th pencil of the
    batch."""
    import numpy as np
    import torch

    import paraht_amd as ph
    from paraht_amd.batched import Batch, reduce_batch
    from paraht_amd.qz import hgeqz_batch
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(77)
    out = {"single": []}
    host = {}
    for n in (1000, 2000):
        A = ph.column_major(torch.randn(n, n, dtype=torch.float64, device=dev, generator=g))
        B = ph.column_major(torch.triu(torch.linalg.qr(
            torch.randn(n, n, dtype=torch.float64, device=dev, generator=g), mode="r")[1]))
        H, T, Q, Z = ph.reduce(A, B, 16, 4, 8, givens=True)
        host[n] = (H.cpu().numpy(), T.cpu().numpy())
        rec = {"n": n}
        for mode in ("eigenvalues", "schur_qz") if n <= 1000 else ("eigenvalues",):
            h, t = ph.column_major(H.clone()), ph.column_major(T.clone())
            q, z = ph.column_major(Q.clone()), ph.column_major(Z.clone())
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if mode == "eigenvalues":
                *_, info = ph.hgeqz(h, t)
            else:
                *_, info = ph.hgeqz(h, t, q, z, schur=True)
            torch.cuda.synchronize()
            rec[mode + "_s"] = time.perf_counter() - t0
            rec[mode + "_info"] = info
        print(f"qz: n={n} " + ", ".join(f"{k[:-2]} {v:.3f} s" for k, v in rec.items()
                                         if k.endswith("_s")), file=sys.stderr, flush=True)
        out["single"].append(rec)
    rng = np.random.default_rng(1302)
    sizes = [int(x) for x in rng.integers(32, 513, size=2048)]
    bt = Batch(sizes)
    As = [torch.randn(n, n, dtype=torch.float64, device=dev, generator=g) for n in sizes]
    Bs = [torch.triu(torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device=dev,
                                                 generator=g), mode="r")[1]) for n in sizes]
    bt.load(As, Bs)
    reduce_batch(bt)
    H0, T0 = bt.A.clone(), bt.B.clone()
    ts = []
    for _ in range(2):
        bt.A.copy_(H0)
        bt.B.copy_(T0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        *_, info = hgeqz_batch(bt)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    t_b = min(ts)
    out["batched"] = {"pencils": len(sizes), "n_range": [32, 512], "s": t_b,
                      "pencils_per_s": len(sizes) / t_b, "nonconverged": sum(1 for x in info if x)}
    print(f"qz: batched {len(sizes)} pencils in {t_b:.3f} s", file=sys.stderr, flush=True)
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle"))
        import oracle
        from threadpoolctl import threadpool_limits
        Hh, Th = host[1000]
        sample = list(range(0, len(sizes), 64))
        pencils = [(bt.view(H0, i).cpu().numpy(), bt.view(T0, i).cpu().numpy()) for i in sample]
        with threadpool_limits(1):
            t0 = time.perf_counter()
            oracle.dhgeqz(Hh, Th)
            t_c = time.perf_counter() - t0
            t0 = time.perf_counter()
            for Hs, Ts in pencils:
                oracle.dhgeqz(Hs, Ts)
            t_s = time.perf_counter() - t0
        share = sum(sizes[i] ** 3 for i in sample) / (sum(n ** 3 for n in sizes) / len(sizes)
                                                      * len(sample))
        out["cpu_dhgeqz"] = {"routine": "LAPACK DHGEQZ JOB='E' (scipy OpenBLAS)", "cores": 1,
                             "n1000_s": t_c,
                             "batched_sample": len(sample),
                             "batched_pencils_per_s": len(sample) / t_s * share}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=8000)
    ap.add_argument("--r", type=int, default=64)
    ap.add_argument("--p", type=int, default=0,
                    help="blocks per stage-1 reflector block (0: the tuned table, else 2)")
    ap.add_argument("--nq", type=int, default=0,
                    help="columns per stage-2 sweep (0: the tuned table, else 16)")
    ap.add_argument("--stage2", choices=("householder", "givens"), default="givens",
                    help="stage-2 variant of the timed steps: rht2ht_givens (the faster chase "
                         "here) or rht2ht_householder (the reference driver's); the other is "
                         "timed once beside it")
    ap.add_argument("--cpu-n", type=int, default=2000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched section (2048 pencils, n uniform in [32, 512])")
    ap.add_argument("--batched-loop-sample", type=int, default=48,
                    help="pencils timed through the single-pencil path for the loop baseline")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the forward + backward measurement")
    ap.add_argument("--no-qz", action="store_true",
                    help="skip the QZ iteration measurements")
    ap.add_argument("--qz-only", action="store_true",
                    help="only the QZ iteration measurements (one JSON object)")
    ap.add_argument("--no-other", action="store_true",
                    help="skip the reduction with the other stage-2 variant")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the HT reduction of one pencil is "
              "not sharded across GPUs in this build (--gpus 1).", file=sys.stderr)
        sys.exit(2)
    if args.qz_only:
        print(json.dumps({"qz": bench_qz(args)}), flush=True)
        return

    import torch

    import paraht_amd as ph
    from paraht_amd.tune import tuned_params
    tuned = tuned_params(args.n, args.r)
    if not args.p:
        args.p = tuned[0] if tuned else 2
    if not args.nq:
        args.nq = tuned[1] if tuned else 16
    n, r, p, nq = args.n, args.r, args.p, args.nq
    givens = args.stage2 == "givens"
    classes = KERNEL_CLASSES_GIVENS if givens else KERNEL_CLASSES
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(1302)
    A0 = ph.column_major(torch.randn(n, n, dtype=torch.float64, device=dev, generator=g))
    R = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device=dev, generator=g),
                        mode="r")[1]
    B0 = ph.column_major(torch.triu(R))
    del R
    A, B = torch.empty_like(A0), torch.empty_like(B0)
    Q, Z = torch.empty_like(A0), torch.empty_like(A0)
    eye = ph.column_major(torch.eye(n, dtype=torch.float64, device=dev))
    ctx = ph.Context(0)

    def step():
        A.copy_(A0)
        B.copy_(B0)
        Q.copy_(eye)
        Z.copy_(eye)
        torch.cuda.synchronize()
        ph.reduce(A, B, r, p, nq, Q, Z, ctx=ctx, givens=givens)

    for i in range(args.warmup):
        step()
        print(f"warmup {i + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t1s, t2s = [], []
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        lt = ctx.last_times()
        t1s.append(lt["stage1_ms"] / 1e3)
        t2s.append(lt["stage2_ms"] / 1e3)
        print(f"step {i + 1}/{args.steps}: stage1 {t1s[-1]:.3f} s, stage2 {t2s[-1]:.3f} s",
              file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # the per-kernel breakdown comes from one more step with every launch
    # bracketed by HIP events (kept out of the timed steps: the events cost
    # about 3%)
    kstats = {}
    prof_steps = 1
    if not args.no_profile:
        ctx.profile(True)
        for _ in range(prof_steps):
            step()
        torch.cuda.synchronize()
        kstats = ctx.kernel_stats()
        ctx.profile(False)
        print("profiled step done", file=sys.stderr, flush=True)

    # correctness of the last step (FP64 GEMMs on the GPU)
    ea = (torch.linalg.norm(Q.T @ A0 @ Z - A) / torch.linalg.norm(A0)).item()
    eb = (torch.linalg.norm(Q.T @ B0 @ Z - B) / torch.linalg.norm(B0)).item()
    zH = int(torch.count_nonzero(torch.tril(A, -2)).item())
    zT = int(torch.count_nonzero(torch.tril(B, -1)).item())

    # the other stage-2 variant on the same pencil: one untimed call (its
    # workspace), then one timed reduction
    variants = {args.stage2: {"stage1_s": sum(t1s) / args.steps,
                              "stage2_s": sum(t2s) / args.steps,
                              "backward_errors": [ea, eb]}}
    other = "householder" if givens else "givens"
    if not args.no_other:
        for rep in range(2):
            A.copy_(A0)
            B.copy_(B0)
            Q.copy_(eye)
            Z.copy_(eye)
            torch.cuda.synchronize()
            ph.reduce(A, B, r, p, nq, Q, Z, ctx=ctx, givens=not givens)
        lt = ctx.last_times()
        oa = (torch.linalg.norm(Q.T @ A0 @ Z - A) / torch.linalg.norm(A0)).item()
        ob = (torch.linalg.norm(Q.T @ B0 @ Z - B) / torch.linalg.norm(B0)).item()
        variants[other] = {"stage1_s": lt["stage1_ms"] / 1e3, "stage2_s": lt["stage2_ms"] / 1e3,
                           "backward_errors": [oa, ob]}
        print(f"{other}: stage2 {variants[other]['stage2_s']:.3f} s", file=sys.stderr,
              flush=True)

    ms = wall / args.steps * 1e3
    t_red = (sum(t1s) + sum(t2s)) / args.steps
    flops = 14.0 * n ** 3
    t1 = sum(t1s) / args.steps
    t2 = sum(t2s) / args.steps
    out = {
        "metric": METRIC, "value": t_red, "unit": "s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"one FP64 pencil n={n}, r={r}, p={p}, nq={nq}; "
                               "A~N(0,1), B=R of QR of N(0,1), Q=Z=I",
                   "n": n, "r": r, "p": p, "nq": nq, "stage2": args.stage2,
                   "p_nq_from": "tuned table paraht_amd/tuned.json" if tuned else "fixed"},
        "gflops": flops / t_red / 1e9,
        "pct_fp64_peak": 100.0 * flops / t_red / FP64_PEAK,
        "stage1_s": t1, "stage2_s": t2,
        "stage1_pct_fp64_peak": 100.0 * (28.0 / 3.0) * n ** 3 / t1 / FP64_PEAK,
        "stage2_hbm_gbs": 32.0 * n ** 3 / nq / t2 / 1e9,
        "backward_errors": [ea, eb], "nonzeros_below_structure": [zH, zT],
        "stage2_variants": variants,
    }
    if kstats:
        tot = sum(v[0] for v in kstats.values()) or 1.0
        out["kernels"] = {k: {"ms": v[0] / prof_steps, "launches": v[1] // prof_steps,
                              "avg_us": 1e3 * v[0] / v[1], "share": v[0] / tot}
                          for k, v in kstats.items() if v[1]}
        dom = max(kstats, key=lambda k: kstats[k][0])
        out["dominant_kernel"] = dom
        if "s2_gen" in kstats and kstats["s2_gen"][1]:
            ms_g, cnt_g = kstats["s2_gen"][0], kstats["s2_gen"][1]
            steps = chase_steps(n, r, nq)
            out["latency_bound"] = {
                "kernel": classes["s2_gen"][0],
                "launches": cnt_g // prof_steps, "avg_launch_us": 1e3 * ms_g / cnt_g,
                "chase_steps": steps, "us_per_step": 1e3 * ms_g / prof_steps / steps}
        # the roofline entry: the throughput kernel class with the most time
        thr = [k for k in kstats if classes.get(k, ("", ""))[1] in ("mfma", "hbm")
               and kstats[k][1]]
        if thr:
            dom = max(thr, key=lambda k: kstats[k][0])
            ms_k, cnt, fl, by = kstats[dom]
            name, bound = classes[dom]
            t_launch = ms_k / cnt / 1e3  # average launch duration (HIP events), s
            if bound == "mfma":
                ach = fl / cnt / t_launch / 1e12
                peak, unit = FP64_PEAK / 1e12, "TFLOP/s"
            else:
                ach = by / cnt / t_launch / 1e9
                peak, unit = HBM_PEAK / 1e9, "GB/s"
            out["roofline"] = {"kernel": name, "bound": bound, "achieved": ach,
                               "peak": peak, "unit": unit, "frac": ach / peak,
                               "traffic": traffic_per_launch(dom, args),
                               "avg_launch_us": t_launch * 1e6, "launches": cnt // prof_steps,
                               "algorithmic_per_launch": (fl if bound == "mfma" else by) / cnt}
    if not args.no_batched:
        out["batched"] = bench_batched(args)
    if not args.no_adjoint:
        out["adjoint"] = bench_adjoint(args)
    if not args.no_qz:
        out["qz"] = bench_qz(args)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args, bench_cfg0(args))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
