"""Benchmark of the Schur-complement hot path on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Prints one JSON line.  Inputs are generated on the host, copied to HBM
before the timed region, and every timed step is one full pass of the
workload's hot path on device-resident data.  --gpus above 1 is refused:
this build targets one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "sdpx.rs_amd")]

import numpy as np  # noqa: E402

METRIC = json.load(open(os.path.join(ROOT, "BASELINE.json")))["metric"]
INT8_PEAK_OPS = 5.0e15      # MI355X dense INT8 MFMA
FP64_PEAK = 78.6e12         # MI355X dense FP64 MFMA
HBM_PEAK = 8.0e12           # bytes/s


def product_traffic(workload):
    """HBM bytes per product-kernel launch of this workload from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (counters scaled by a copy
    kernel of known size), or None."""
    try:
        d = json.load(open(os.path.join(ROOT, "profiles", "product_traffic.json")))[workload]
        return d["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo) or None."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def progress(msg):
    """A progress line on stderr (the JSON result stays alone on stdout)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def to_device(rec):
    import torch

    t = torch.from_numpy(rec.view(np.uint8).reshape(-1).copy())
    return t.to("cuda")


def rns_gemm_workload(args):
    """Exact 256-bit GEMM 4096 x 4096 x 4096, fixed-point entries uniform in
    [-1, 1) with 256 fraction bits (the RNS micro-workload)."""
    import torch

    import sdpx_gpu
    from sdpx_gpu import mp

    nl, m, n, k = 4, 4096, 4096, 4096
    rng = np.random.default_rng(42)
    a = mp.random_fixed(m * k, nl, 256, rng)
    b = mp.random_fixed(k * n, nl, 256, rng)
    da, db = to_device(a), to_device(b)
    dout = torch.empty(m * n * mp.dtype(nl).itemsize, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    L = sdpx_gpu.lib()
    sdpx_gpu.check(L.sdpx_gpu_init(0))

    def step():
        code = sdpx_gpu.rns_gemm_dev("N", "N", m, n, k, da.data_ptr(), m, db.data_ptr(), k, nl,
                                     False, dout.data_ptr(), m * n)
        assert code == sdpx_gpu.OK

    for _ in range(args.warmup):
        step()
    # per-stage device times on the library's stream (HIP events)
    sdpx_gpu.set_profiling(True)
    step()
    stages = sdpx_gpu.last_stage_times()
    sdpx_gpu.set_profiling(False)
    moduli, planes, bits = sdpx_gpu.last_plan()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    int8_ops = 2.0 * m * n * k * planes_ops(moduli, planes)
    product_s = stages[2] / 1e3
    achieved = int8_ops / product_s
    result = {
        "value": dt, "unit": "s/GEMM", "ms_per_step": dt * 1e3, "dtype": "int8->int32 (u64 limbs)",
        "config": {"workload": "rns_gemm 4096x4096x4096 MP256 fixed-point (exact, bit-exact RNS)",
                   "m": m, "n": n, "k": k, "precision_bits": 256, "moduli": moduli,
                   "int8_planes": planes, "log2_modulus": round(bits, 1),
                   "stage_ms": {"scan": stages[0], "encode": stages[1], "product": stages[2],
                                "crt": stages[3]},
                   "mp_macs_per_s": m * n * k / dt},
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": INT8_PEAK_OPS / 1e12,
                     "unit": "TOP/s (INT8)", "frac": achieved / INT8_PEAK_OPS,
                     "traffic": product_traffic("rns_gemm"),
                     "kernel": "product_kernel (all moduli)", "int8_ops_per_step": int8_ops,
                     "share_of_step": product_s / dt},
    }
    if args.cpu_baseline:
        result["cpu_baseline"] = cpu_baseline_gemm(a, b, m, n, k, nl)
    return result


def planes_ops(moduli, planes):
    """INT8 GEMMs per exact product: one per single-digit modulus, four per
    two-digit modulus (planes = single + 2 * double)."""
    double = planes - moduli
    single = moduli - double
    return single + 4 * double


def cpu_baseline_gemm(a, b, m, n, k, nl):
    """The reference's CPU exact product restated in C (residues over 18-26
    bit primes as binary64 products, CRT, one rounding; OpenMP over all
    host cores) on a bounded sample: the leading sub x sub output block with
    the full inner dimension, scaled to the full product."""
    sys.path[:0] = [os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle")]
    from helpers import oracle_gemm_rns
    from sdpx_gpu import mp

    cores = os.cpu_count() or 1
    sub = 512
    dt = mp.dtype(nl)
    a_sub = np.ascontiguousarray(a.view(dt).reshape(k, m)[:, :sub]).reshape(-1)  # column-major m x k
    b_sub = np.ascontiguousarray(b.view(dt).reshape(n, k)[:sub]).reshape(-1)      # column-major k x n
    t0 = time.perf_counter()
    assert oracle_gemm_rns("N", "N", sub, sub, k, a_sub, sub, b_sub, k, threads=cores) is not None
    el = time.perf_counter() - t0
    return {"value": el * (m * n) / (sub * sub), "unit": "s/GEMM", "cores": cores, "kind": "port",
            "cpu": cpu_model(),
            "sample": f"the residue product (the reference's CPU algorithm, rns_blas.rs) of the leading "
                      f"{sub}x{sub} outputs with the full inner dimension {k}: {el:.2f} s, scaled by "
                      f"{m * n}/{sub * sub}"}


def sampled_workload(args):
    """One IPM iteration's Schur build + arrow factor of the sampled
    bootstrap-style SDP at 256 bits (40 PSD blocks of side U{10..60},
    K = 2h - 1 Chebyshev samples, log-uniform weights in [1e-6, 1], 2000
    border rows), with every exact product as INT8-MFMA RNS products; the
    same problem in float64 on the same GPU beside it."""
    import torch

    import sdpx_gpu
    from sdpx_gpu import mp, workloads

    nl = 4
    sdpx_gpu.check(sdpx_gpu.lib().sdpx_gpu_init(0))
    progress(f"building the sampled problem ({args.border} border rows)")
    prob = workloads.sampled_problem(nl, seed=7, t=args.border)
    t = prob["t"]
    progress("problem built")

    run = sdpx_gpu.SampledDev(prob, nl)   # inputs resident in HBM

    def step():
        assert run.run()

    for _ in range(args.warmup):
        step()
    progress("warmup iteration done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    progress(f"{args.steps} timed iterations: {dt:.3f} s each")
    phases = sdpx_gpu.sampled_stats()
    # the host-buffer entry point (records staged over PCIe both ways)
    h0 = time.perf_counter()
    ok, *_ = sdpx_gpu.sampled_arrow(prob, nl)
    assert ok
    host_dt = time.perf_counter() - h0
    # one profiled pass for the kernel roofline (HIP events on the library stream)
    sdpx_gpu.set_profiling(True)
    sdpx_gpu.rns_totals(reset=True)
    sdpx_gpu.panel_phase_times(reset=True)
    step()
    tot = sdpx_gpu.rns_totals(reset=True)
    pms, plaunch = sdpx_gpu.panel_phase_times(reset=True)
    sdpx_gpu.set_profiling(False)
    # the per-iteration solves: one right-hand side through the factor and
    # stationary refinement against the exact-rows residual (tolerances
    # eps^(3/4) = 2^-191.25, the reference's MPFR default, stop ratio 5)
    run.schur()
    n_kkt = sum(b["k"] for b in prob["blocks"]) + t
    rhs = mp.from_float64(np.random.default_rng(5).standard_normal(n_kkt), nl)
    run.solve_refined(rhs, 1, 2.0 ** -191.25, 2.0 ** -191.25)
    torch.cuda.synchronize()
    r0 = time.perf_counter()
    _, rok, rnorm, rcount = run.solve_refined(rhs, 1, 2.0 ** -191.25, 2.0 ** -191.25)
    refine_dt = time.perf_counter() - r0
    progress(f"refined solve: {refine_dt * 1e3:.1f} ms, {rcount} corrections")
    # float64 on the same GPU, same problem (dynamic pivot rule: the 256-bit
    # leaves are singular in binary64)
    dev = sdpx_gpu.SampledF64(prob)
    for _ in range(max(1, args.warmup)):
        dev.run(dynamic=(1e-13, 2e-7))
    torch.cuda.synchronize()
    f0 = time.perf_counter()
    fsteps = max(5, args.steps)
    for _ in range(fsteps):
        info, ftimes = dev.run(dynamic=(1e-13, 2e-7))
    torch.cuda.synchronize()
    fdt = (time.perf_counter() - f0) / fsteps
    product_s = tot["stage_ms"][2] / 1e3
    achieved = tot["int8_ops"] / product_s if product_s > 0 else 0.0
    kmax = max(b["k"] for b in prob["blocks"])
    chain = t + kmax  # sequential pivot columns: the border's, after the leaves' (one launch per leaf column)
    phase_table = {
        "ldl_dots": {"ms": pms["ldl_dots"], "bound": "latency: a chain of dependent 256-bit operations per "
                     "pivot column", "columns": chain, "us_per_column": 1e3 * pms["ldl_dots"] / chain},
        "trsm_base": {"ms": pms["trsm_base"], "bound": "latency: one dependent step per solved column",
                      "launch_groups": plaunch["trsm_base"]},
        "forward_rows": {"ms": pms["forward_rows"], "bound": "latency/VALU: K dependent exact-dot rows per "
                         "column, all leaves and border columns in one launch", "steps": kmax,
                         "us_per_step": 1e3 * pms["forward_rows"] / kmax},
        "exact_products": {"ms": pms["exact_products"], "bound": "mfma (INT8) for the product stage, hbm for "
                           "scan / encode / CRT", "stage_ms": dict(zip(("scan", "encode", "product", "crt"),
                                                                       tot["stage_ms"])),
                           "product_int8_tops": achieved / 1e12, "product_frac_of_int8_peak":
                           achieved / INT8_PEAK_OPS},
        "other_ms": dt * 1e3 - sum(pms.values()),
    }
    result = {
        "value": dt, "unit": "s/IPM-iter", "ms_per_step": dt * 1e3,
        "dtype": "mp256 (int8 MFMA residues, int32 accumulation)",
        "config": {"workload": "sampled bootstrap SDP, 40 PSD blocks h~U{10..60}, K=2h-1, "
                               f"{t} border rows, MPFR-256: Schur build + arrow LDL^T",
                   "precision_bits": 256, "blocks": len(prob["blocks"]), "border": t,
                   "samples": int(sum(b["k"] for b in prob["blocks"])),
                   "phase_s": {k: phases[k] for k in ("build_s", "leaves_s", "border_s")},
                   "gpu_products": phases["gpu_products"], "host_products": phases["host_products"],
                   "rns_stage_ms": dict(zip(("scan", "encode", "product", "crt"), tot["stage_ms"])),
                   "rns_calls": tot["calls"], "mp_macs": tot["mp_macs"],
                   "fp64_gpu_s_per_iter": fdt, "fp64_phase_ms": ftimes,
                   "mp256_over_fp64": dt / fdt,
                   "phase_table_ms": phase_table,
                   "host_entry_s_per_iter": host_dt,
                   "solve_refined_1rhs_s": refine_dt, "solve_refined_corrections": rcount,
                   "solve_refined_residual_log2": float(np.log2(max(mp.to_float64(rnorm)[0], 1e-300)))},
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": INT8_PEAK_OPS / 1e12,
                     "unit": "TOP/s (INT8)", "frac": achieved / INT8_PEAK_OPS,
                     "traffic": product_traffic("sampled") if t == 2000 else None,
                     "kernel": "product_kernel (INT8 MFMA residue products)",
                     "int8_ops_per_step": tot["int8_ops"], "share_of_step": product_s / dt,
                     "note": "the largest share of the step is the latency-bound exact-dot panel "
                             "kernels (ldl_col_wave_kernel, trsm_rl_kernel), which have no MFMA or "
                             "HBM roofline"},
    }
    if args.target and t != 8192:
        result["config"]["target_8192"] = sampled_target_point(args)
    if args.cpu_baseline and t <= 2000:  # larger borders: minutes of CPU time
        sys.path[:0] = [os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle")]
        from helpers import oracle_sampled_arrow, oracle_set_rns_products

        cores = os.cpu_count() or 1
        oracle_set_rns_products(True)
        try:
            c0 = time.perf_counter()
            ok, *_ = oracle_sampled_arrow(prob, nl, threads=cores)
            cdt = time.perf_counter() - c0
        finally:
            oracle_set_rns_products(False)
        result["cpu_baseline"] = {"value": cdt, "unit": "s/IPM-iter", "cores": cores, "kind": "port",
                                  "cpu": cpu_model(),
                                  "sample": "the whole workload (one iteration) through the C restatement, "
                                            "the reference's CPU algorithm: residue products (binary64 "
                                            "residues over 18-26 bit primes, CRT) where it takes them "
                                            "(mpfr.rs:349-361), exact dots elsewhere; OpenMP on all cores"}
    return result


def sampled_target_point(args):
    """The same workload at an 8192-row Schur border (the order the target
    is stated at): seconds per iteration at 256 bits and of the Float64
    form on the same GPU, and their ratio (2 timed iterations each)."""
    import torch

    import sdpx_gpu
    from sdpx_gpu import workloads

    nl = 4
    progress("target point: building the 8192-row problem")
    prob = workloads.sampled_problem(nl, seed=7, t=8192)
    run = sdpx_gpu.SampledDev(prob, nl)
    assert run.run()  # warmup
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    steps = 2
    for _ in range(steps):
        assert run.run()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    phases = sdpx_gpu.sampled_stats()
    del run
    dev = sdpx_gpu.SampledF64(prob)
    dev.run(dynamic=(1e-13, 2e-7))
    torch.cuda.synchronize()
    f0 = time.perf_counter()
    for _ in range(3):
        dev.run(dynamic=(1e-13, 2e-7))
    torch.cuda.synchronize()
    fdt = (time.perf_counter() - f0) / 3
    progress(f"target point: {dt:.3f} s MP256, {fdt * 1e3:.2f} ms Float64 per iteration")
    return {"border": 8192, "mp256_s_per_iter": dt, "fp64_gpu_s_per_iter": fdt, "mp256_over_fp64": dt / fdt,
            "phase_s": {k: phases[k] for k in ("build_s", "leaves_s", "border_s")},
            "samples": int(sum(b["k"] for b in prob["blocks"]))}


def chol_workload(args):
    """Blocked MPFR-512 LDL^T of a synthetic SPD Schur matrix of order 8192,
    A = Q diag(s) Q^T with Q a seeded Haar orthogonal matrix (binary64,
    lifted exactly) and s = 2^-100 ... 1 log-spaced (powers of two, so
    Q diag(s) is exact in binary64 and A is formed exactly rounded at 512
    bits by one RNS product on the GPU), device-resident; the same matrix's
    float64 Cholesky on the same GPU beside it."""
    import torch

    import sdpx_gpu
    from sdpx_gpu import mp

    nl, n = 8, args.order
    rng = np.random.default_rng(11)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = 2.0 ** np.round(np.linspace(-100, 0, n))
    bq = q * s  # exact: power-of-two scaling
    sdpx_gpu.check(sdpx_gpu.lib().sdpx_gpu_init(0))
    dq = to_device(mp.from_float64(q.T.reshape(-1), nl))       # Q column-major
    dbq = to_device(mp.from_float64(bq.T.reshape(-1), nl))     # Q diag(s)
    vb = mp.dtype(nl).itemsize
    da = torch.empty(n * n * vb, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    code = sdpx_gpu.rns_gemm_dev("N", "T", n, n, n, dbq.data_ptr(), n, dq.data_ptr(), n, nl, False,
                                 da.data_ptr(), n * n)
    assert code == sdpx_gpu.OK
    del dq, dbq
    work = torch.empty_like(da)
    dinv = torch.empty(n * vb, dtype=torch.uint8, device="cuda")
    L = sdpx_gpu.lib()
    import ctypes

    def step():
        work.copy_(da)
        torch.cuda.synchronize()
        ok = ctypes.c_int32()
        t0 = time.perf_counter()
        sdpx_gpu.check(L.sdpx_gpu_mp_ldl_dev(nl, n, work.data_ptr(), n, None, 0, None, None,
                                             dinv.data_ptr(), 0, ctypes.byref(ok)))
        dt = time.perf_counter() - t0
        assert ok.value == 1
        return dt

    progress(f"order {n}: operands built")
    for i in range(args.warmup):
        progress(f"warmup factor {i} ({step():.2f} s)")
    sdpx_gpu.set_profiling(True)
    sdpx_gpu.rns_totals(reset=True)
    times = []
    for i in range(args.steps):
        times.append(step())
        progress(f"factor {i}: {times[-1]:.2f} s")
    tot = sdpx_gpu.rns_totals(reset=True)
    sdpx_gpu.set_profiling(False)
    dt = sum(times) / len(times)
    st = sdpx_gpu.mp_ldl_stats()
    # 16 Gaussian right-hand sides: L D L^T x = b on the device factor, then
    # the residual b - A x with A x exactly rounded at 1024 bits (records
    # widened exactly; one RNS product) - the 1024-bit reference residual
    nrhs = 16
    bmat = rng.standard_normal((n, nrhs))
    db = to_device(mp.from_float64(bmat.T.reshape(-1), nl))
    dx = db.clone()
    torch.cuda.synchronize()
    s0 = time.perf_counter()
    sdpx_gpu.check(L.sdpx_gpu_mp_ldl_solve_dev(nl, n, work.data_ptr(), n, dinv.data_ptr(), nrhs,
                                               dx.data_ptr(), n))
    torch.cuda.synchronize()
    solve_s = time.perf_counter() - s0
    progress(f"16 right-hand sides solved ({solve_s:.2f} s)")

    def widen(t, count):  # 8-limb records -> 16-limb records, same values
        r = t.view(count, vb)
        return torch.cat([torch.zeros(count, 64, dtype=torch.uint8, device="cuda"), r], dim=1).reshape(-1)

    a16 = widen(da, n * n)
    x16 = widen(dx, n * nrhs)
    vb16 = mp.dtype(16).itemsize
    ax16 = torch.empty(n * nrhs * vb16, dtype=torch.uint8, device="cuda")
    code = sdpx_gpu.rns_gemm_dev("N", "N", n, nrhs, n, a16.data_ptr(), n, x16.data_ptr(), n, 16, False,
                                 ax16.data_ptr(), n * nrhs)
    assert code == sdpx_gpu.OK
    del a16, x16
    b16 = mp.from_float64(bmat.T.reshape(-1), 16)
    resid = mp.sub_exact_log2(b16, ax16.cpu().numpy().view(mp.dtype(16)))
    xv = mp.to_float64(dx.cpu().numpy().view(mp.dtype(nl)))
    rel = resid - np.log2(max(np.abs(bmat).max(), 1e-300))
    progress(f"residual max |b - A x| = 2^{resid:.1f} (1024-bit A x)")
    # float64 Cholesky of the same matrix (dynamic pivot rule: cond 2^100)
    a64 = torch.tensor((q * s) @ q.T, dtype=torch.float64, device="cuda")
    f = sdpx_gpu.DenseFactor(n)
    f.set_dynamic_pivot(1e-13, 2e-7)
    f.refactor_dev(a64.data_ptr(), n)
    torch.cuda.synchronize()
    f0 = time.perf_counter()
    for _ in range(3):
        f.refactor_dev(a64.data_ptr(), n)
    fdt = (time.perf_counter() - f0) / 3
    product_s = tot["stage_ms"][2] / 1e3
    achieved = tot["int8_ops"] / product_s if product_s > 0 else 0.0
    result = {
        "value": dt, "unit": "s/factor", "ms_per_step": dt * 1e3,
        "dtype": "mp512 (int8 MFMA residues, int32 accumulation)",
        "config": {"workload": f"MPFR-512 LDL^T of Q diag(s) Q^T, order {n}, s in [2^-100, 1]",
                   "order": n, "precision_bits": 512, "gpu_products": st["gpu_products"],
                   "rns_stage_ms_per_factor": {k: v / args.steps for k, v in
                                               zip(("scan", "encode", "product", "crt"),
                                                   tot["stage_ms"])},
                   "mp_macs_per_factor": tot["mp_macs"] / args.steps,
                   "fp64_gpu_s_per_factor": fdt, "mp512_over_fp64": dt / fdt,
                   "solve_16rhs_s": solve_s,
                   "residual_log2_max_abs": resid if np.isfinite(resid) else None,
                   "residual_log2_rel_to_max_b": rel if np.isfinite(rel) else None,
                   "solution_max_abs": float(np.abs(xv).max())},
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": INT8_PEAK_OPS / 1e12,
                     "unit": "TOP/s (INT8)", "frac": achieved / INT8_PEAK_OPS, "traffic": None,
                     "kernel": "product_kernel (INT8 MFMA residue products)",
                     "int8_ops_per_step": tot["int8_ops"] / args.steps},
    }
    if args.cpu_baseline:
        sys.path[:0] = [os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle")]
        from helpers import oracle_ldl, oracle_set_rns_products

        sub = 2048
        a_sub = mp.from_float64(((q[:sub] * s) @ q[:sub].T).T.reshape(-1), nl)
        cores = os.cpu_count() or 1
        oracle_set_rns_products(True)
        try:
            c0 = time.perf_counter()
            oracle_ldl(a_sub, sub, threads=cores)
            cdt = time.perf_counter() - c0
        finally:
            oracle_set_rns_products(False)
        result["cpu_baseline"] = {"value": cdt * (n / sub) ** 3, "unit": "s/factor", "cores": cores,
                                  "kind": "port", "cpu": cpu_model(),
                                  "sample": f"C restatement of ldl_rec on the leading {sub}x{sub} block, "
                                            f"trailing products by the reference's CPU residue product, "
                                            f"base cases by exact dots ({cdt:.1f} s), scaled by (n/{sub})^3"}
    return result


def rns_syrk_workload(args):
    """Exact 256-bit SYRK B^T B, B 16384 x 4096 fixed-point (packed upper)."""
    import torch

    import sdpx_gpu
    from sdpx_gpu import mp

    nl, k, n = 4, 16384, 4096
    rng = np.random.default_rng(43)
    b = mp.random_fixed(k * n, nl, 256, rng)
    db = to_device(b)
    out_len = n * (n + 1) // 2
    dout = torch.empty(out_len * mp.dtype(nl).itemsize, dtype=torch.uint8, device="cuda")
    sdpx_gpu.check(sdpx_gpu.lib().sdpx_gpu_init(0))
    torch.cuda.synchronize()

    def step():
        code = sdpx_gpu.rns_gemm_dev("T", "N", n, n, k, db.data_ptr(), k, db.data_ptr(), k, nl,
                                     True, dout.data_ptr(), out_len)
        assert code == sdpx_gpu.OK

    for _ in range(args.warmup):
        step()
    sdpx_gpu.set_profiling(True)
    sdpx_gpu.rns_totals(reset=True)
    step()
    tot = sdpx_gpu.rns_totals(reset=True)
    sdpx_gpu.set_profiling(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    product_s = tot["stage_ms"][2] / 1e3
    achieved = tot["int8_ops"] / product_s
    return {"value": dt, "unit": "s/SYRK", "ms_per_step": dt * 1e3, "dtype": "int8->int32",
            "config": {"workload": "rns_syrk B^T B, B 16384x4096 MP256 fixed-point",
                       "stage_ms": dict(zip(("scan", "encode", "product", "crt"), tot["stage_ms"])),
                       "mp_macs": tot["mp_macs"]},
            "roofline": {"bound": "mfma", "achieved": achieved / 1e12,
                         "peak": INT8_PEAK_OPS / 1e12, "unit": "TOP/s (INT8)",
                         "frac": achieved / INT8_PEAK_OPS, "traffic": None,
                         "kernel": "product_kernel", "int8_ops_per_step": tot["int8_ops"]}}


def fp64_workload(args):
    """Float64 forms: one IPM iteration's Schur build (M = svec(Rinv A_i
    Rinv^T) . svec(Rinv A_j Rinv^T) on FP64 MFMA) + dense Cholesky of the
    dense-SDP configurations, n = 200, m = 1000 (timed) and n = 50, m = 100,
    and the adjoint with 64 cotangents through the same factor, all on
    device-resident data.  Rinv = I + G / (4 sqrt n) stands for the NT
    scaling factor of an iterate."""
    import ctypes

    import torch

    import sdpx_gpu
    from sdpx_gpu import ipm

    L = sdpx_gpu.lib()
    sdpx_gpu.check(L.sdpx_gpu_init(0))
    vp = ctypes.c_void_p

    def setup(n, m, seed):
        prob = ipm.dense_problem(n, m, seed)
        rng = np.random.default_rng(seed + 1)
        rinv = np.eye(n) + rng.standard_normal((n, n)) / (4.0 * np.sqrt(n))
        d_r = torch.tensor(np.asfortranarray(rinv).T.copy().reshape(-1), dtype=torch.float64, device="cuda")
        d_a = torch.tensor(prob["a_svec"].T.copy().reshape(-1), dtype=torch.float64, device="cuda")
        d_s = torch.zeros(m * m, dtype=torch.float64, device="cuda")
        f = sdpx_gpu.DenseFactor(m)
        return prob, rinv, d_r, d_a, d_s, f

    def iteration(n, m, d_r, d_a, d_s, f):
        sdpx_gpu.check(L.sdpx_gpu_schur_psd_f64_dev(n, m, vp(d_r.data_ptr()), vp(d_a.data_ptr()),
                                                    vp(d_s.data_ptr()), m))
        assert f.refactor_dev(d_s.data_ptr(), m)

    def timed(fn, steps):
        for _ in range(max(1, args.warmup)):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    out = {}
    for n, m, seed in ((200, 1000, 4), (50, 100, 0)):
        prob, rinv, d_r, d_a, d_s, f = setup(n, m, seed)
        dt = timed(lambda: iteration(n, m, d_r, d_a, d_s, f), max(5, args.steps))
        out[(n, m)] = (dt, prob, rinv, d_r, d_a, f)
        progress(f"fp64 n={n} m={m}: {dt * 1e3:.3f} ms per Schur build + factor")
    dt, prob, rinv, d_r, d_a, f = out[(200, 1000)]
    n, m, s_len, r = 200, 1000, 200 * 201 // 2, 64
    ybar = torch.tensor(np.random.default_rng(1).standard_normal((r, m)).reshape(-1), dtype=torch.float64,
                        device="cuda")
    d_bg = torch.zeros(m * r, dtype=torch.float64, device="cuda")
    d_cg = torch.zeros(s_len * r, dtype=torch.float64, device="cuda")

    def adjoint():
        sdpx_gpu.check(L.sdpx_gpu_adjoint_psd_f64_dev(f._h, n, vp(d_r.data_ptr()), vp(d_a.data_ptr()), r,
                                                      vp(ybar.data_ptr()), vp(d_bg.data_ptr()),
                                                      vp(d_cg.data_ptr())))
    adt = timed(adjoint, max(5, args.steps))
    progress(f"adjoint, 64 cotangents: {adt * 1e3:.3f} ms")

    # NT scaling of the PSD blocks (the step before the Schur build): one
    # block of order 200, one of order 50, and 40 blocks of orders 10..60
    # in one launch; iterates S, Z = Q diag(logspace(0, -2)) Q^T
    def nt_case(dims, seed):
        rng = np.random.default_rng(seed)

        def spd(k):
            q, _ = np.linalg.qr(rng.standard_normal((k, k)))
            return (q * np.logspace(0, -2, k)) @ q.T

        nt = sdpx_gpu.NtScaling(dims)
        nt.s.copy_(torch.from_numpy(np.concatenate([sdpx_gpu.svec_f64(spd(k)) for k in dims])))
        nt.z.copy_(torch.from_numpy(np.concatenate([sdpx_gpu.svec_f64(spd(k)) for k in dims])))
        assert not any(nt.run())
        return timed(nt.run, max(5, args.steps))

    sampled_dims = [int(v) for v in np.random.default_rng(7).integers(10, 61, 40)]
    nt_ms = {"n200": nt_case([200], 1) * 1e3, "n50": nt_case([50], 2) * 1e3,
             "sampled_40_blocks_10_60": nt_case(sampled_dims, 3) * 1e3}
    progress(f"NT scaling: {nt_ms}")
    flops = m * 4.0 * n ** 3 + m * (m + 1) * s_len + m ** 3 / 3.0
    result = {
        "value": dt, "unit": "s/IPM-iter", "ms_per_step": dt * 1e3, "dtype": "f64",
        "config": {"workload": "Float64 dense SDP, one PSD block n=200, m=1000: Schur build (FP64 MFMA "
                               "congruences + SYRK) + Cholesky per IPM iteration",
                   "n": n, "m": m, "cfg_n50_m100_s_per_iter": out[(50, 100)][0],
                   "adjoint_64_cotangents_s": adt, "flops_per_iter": flops,
                   "nt_scaling_ms": nt_ms},
        "roofline": {"bound": "mfma", "achieved": flops / dt / 1e12, "peak": FP64_PEAK / 1e12,
                     "unit": "TFLOP/s (FP64)", "frac": flops / dt / FP64_PEAK, "traffic": None,
                     "kernel": "schur_psd_f64 + dense Cholesky (whole iteration)"},
    }
    if args.cpu_baseline:
        A = prob["A"]
        cores = os.cpu_count() or 1
        c0 = time.perf_counter()
        Ms = np.stack([ipm.svec(rinv @ A[i] @ rinv.T) for i in range(m)], axis=1)
        M = Ms.T @ Ms
        np.linalg.cholesky(M)
        cdt = time.perf_counter() - c0
        result["cpu_baseline"] = {"value": cdt, "unit": "s/IPM-iter", "cores": cores, "kind": "port",
                                  "cpu": cpu_model(),
                                  "sample": "the same iteration (n=200, m=1000) in numpy/LAPACK, the "
                                            "reference formula of kkt/condensed/psd.rs:1299-1456"}
    return result


WORKLOADS = {"rns_gemm": rns_gemm_workload, "sampled": sampled_workload,
             "chol": chol_workload, "rns_syrk": rns_syrk_workload, "fp64": fp64_workload}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="sampled", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", dest="cpu_baseline", action="store_false")
    ap.add_argument("--no-target", dest="target", action="store_false",
                    help="skip the 8192-row measurement of --workload sampled")
    ap.add_argument("--order", type=int, default=8192, help="matrix order of --workload chol")
    ap.add_argument("--border", type=int, default=2000,
                    help="border rows (free variables) of --workload sampled")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run one device-to-device copy of MIB MiB (a known-traffic kernel "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if args.copy_calibration:
        import torch

        src = torch.ones(args.copy_calibration << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        progress(f"copy calibration: {args.copy_calibration} MiB read, {args.copy_calibration} MiB written")
        del src, dst
    r = WORKLOADS[args.workload](args)
    line = {"metric": METRIC, "value": r["value"], "unit": r["unit"], "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["ms_per_step"],
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
            "dtype": r["dtype"], "data": "synthetic", "config": r["config"],
            "roofline": r["roofline"]}
    if "cpu_baseline" in r:
        line["cpu_baseline"] = r["cpu_baseline"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
