#!/usr/bin/env python3
"""Benchmark of the MI355X genotype hot path.

Headline workload (one GPU): the LOCO matrix-free LMM genotypes, n = 200,000 samples x
m = 600,000 SNPs, HWE genotypes with p ~ U(0.01, 0.5) and 0.5 % missing, generated on the device
(seed 20261205).  A step is one K.V apply with k = 16 right-hand sides over all m SNPs of the
resident packed matrix (both genotype products and all epilogues).  value = genotypes/s =
n*m / (time per step), inputs resident in HBM.

The same line carries the other legs: the LOCO PCG + scan over 22 chromosomes on these genotypes
(16 phenotype replicates, relative residual 1e-6, exact GLS stage), the plain FP32 GPU
comparator, K.V / exact and standardised GRM / projected scan at n = 50,000 x m = 500,000,
a Poisson stream of serving requests over n = 100,000 x m = 1,000,000 resident genotypes, the
roofline of the dominant kernel (HIP events on the operator's own stream) and the CPU rate of
the reference's blocked K.V algorithm on a bounded slice of the headline workload
(cpu_baseline, kind "port").

Usage: python bench.py [--gpus 1] [--steps 20] [--warmup 3]
"""
from __future__ import annotations

import argparse
from concurrent.futures import ThreadPoolExecutor
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
INT8_PEAK_TOPS = 5000.0
FP16_PEAK_TFLOPS = 2500.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def profiled_traffic(key, n, m, k):
    """HBM bytes per launch of a kernel from the committed rocprofv3 PMC summary of this workload
    (FETCH_SIZE + WRITE_SIZE, calibrated against a plain copy kernel), or None if absent."""
    path = os.path.join(ROOT, "profiles", f"kv_apply_n{n // 1000}k_m{m // 1000}k_k{k}_hbm_traffic.md")
    try:
        for line in open(path):
            if line.startswith(f"traffic_bytes_per_launch {key} "):
                return float(line.rsplit(":", 1)[1])
    except OSError:
        pass
    return None


def copy_rate(csa, gib=4, reps=5):
    """On-box HBM reference: a plain device copy of `gib` GiB (k_copy16), read + write bytes over
    the fastest of `reps` synchronous calls, in GB/s."""
    nbytes = gib << 30
    src = csa.DeviceArray(nbytes // 8)
    dst = csa.DeviceArray(nbytes // 8)
    best = float("inf")
    for _ in range(reps + 1):
        t0 = time.perf_counter()
        csa._check(csa.lib().cs_dev_copy(dst.ptr, src.ptr, nbytes))
        best = min(best, time.perf_counter() - t0)
    src.close()
    dst.close()
    return 2.0 * nbytes / best / 1e9


def measured_traffic(args):
    """HBM bytes per launch of the two genotype products of this K.V workload, measured in this run:
    two rocprofv3 --pmc child runs of `bench.py --kv-only` (FETCH_SIZE, then WRITE_SIZE), each
    scaled by what the same counter reads for the plain 2 GiB copy kernel (`--copy-probe`, two
    more child runs).  Returns {"k_geno1s": bytes, "k_geno2s": bytes, "fetch_scale": .,
    "write_scale": .} or None when rocprofv3 is absent or a run fails (the caller then falls back
    to the committed profile)."""
    import csv
    import shutil
    import subprocess
    import tempfile
    exe = shutil.which("rocprofv3") or ("/opt/rocm/bin/rocprofv3" if os.path.exists("/opt/rocm/bin/rocprofv3") else None)
    if exe is None:
        return None
    tmp = tempfile.mkdtemp(prefix="csamd_pmc_", dir="/tmp")
    env = dict(os.environ, TMPDIR="/tmp")

    def run(counter, extra, tag):
        out = os.path.join(tmp, tag)
        cmd = [exe, "--pmc", counter, "--output-format", "csv", "-d", out, "-o", tag, "--",
               sys.executable, os.path.abspath(__file__)] + extra
        subprocess.run(cmd, cwd="/tmp", env=env, timeout=240, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL, check=True)
        vals = {}
        for root, _, files in os.walk(out):
            for fn in files:
                if fn.endswith("counter_collection.csv"):
                    for r in csv.DictReader(open(os.path.join(root, fn))):
                        if r.get("Counter_Name") == counter:
                            vals.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
        return {kname: sum(v) / len(v) for kname, v in vals.items()}

    try:
        kv = ["--kv-only", "--steps", "1", "--warmup", "0", "--no-cpu", "--n", str(args.n), "--m", str(args.m),
              "--k", str(args.k), "--prec", args.prec, "--missing-mode", str(args.missing_mode)]
        f, w = run("FETCH_SIZE", kv, "f"), run("WRITE_SIZE", kv, "w")
        cf, cw = run("FETCH_SIZE", ["--copy-probe"], "cf"), run("WRITE_SIZE", ["--copy-probe"], "cw")
        copy_kib = (2 << 30) / 1024.0
        cfk = [v for kname, v in cf.items() if "k_copy16" in kname]
        cwk = [v for kname, v in cw.items() if "k_copy16" in kname]
        fs, ws = copy_kib / cfk[0], copy_kib / cwk[0]
        out = dict(fetch_scale=round(fs, 4), write_scale=round(ws, 4))
        for key in ("k_geno1s", "k_geno2s"):
            fk = [v for kname, v in f.items() if f"csamd::{key}<" in kname and ", false, " in kname]
            wk = [v for kname, v in w.items() if f"csamd::{key}<" in kname and ", false, " in kname]
            if fk and wk:
                out[key] = (max(fk) * fs + max(wk) * ws) * 1024.0
        return out if "k_geno1s" in out and "k_geno2s" in out else None
    except Exception as e:   # a missing tool or a failed counter run leaves the committed profile
        log(f"traffic measurement skipped: {e}")
        return None
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def cpu_baseline(N, M_slice, k, seed, pmin=0.01, miss=0.005):
    """The reference's K.V algorithm (kinship_operator.cpp:225-299: 512-SNP blocks decoded to FP64,
    z^T X, then z (z^T X); the oracle's cache-blocked, threaded restatement) on a slice of the
    headline workload: all N samples, the first M_slice SNPs of the same generator."""
    import oracle
    threads = min(os.cpu_count() or 1, 16)
    body = oracle.gen_hwe_bed(seed, N, M_slice, pmin, 0.5, miss, nthreads=threads)
    means, inv = oracle.stats(body, N, M_slice, nthreads=threads)
    X = np.random.default_rng(1).standard_normal((N, k))
    t0 = time.perf_counter()
    oracle.kinship_apply_blocked(body, N, means, inv, None, None, X, nthreads=threads)
    dt = time.perf_counter() - t0
    return dict(value=N * M_slice / dt, unit="genotypes/s", cores=threads, kind="port",
                sample=f"FP64 CPU K.V apply in the reference's 512-SNP block form, k={k}, all {N} samples x "
                       f"first {M_slice} SNPs of the same generator ({dt:.1f} s)")


def kv_leg(csa, reader, N, M, k, steps, warmup, prec="int8_l3", missing_mode=0, extras=True):
    """K.V applies with k device-resident right-hand sides over all SNPs: time per step, phases
    (HIP events on the operator's stream), roofline of the dominant genotype product, and -- with
    extras -- the host-buffer (PCIe) rate and the few-column applies."""
    op = csa.PlinkKinshipOperator(reader)
    op.set_precision(csa.CS_PREC_INT8_L3 if prec == "int8_l3" else csa.CS_PREC_INT8_L2)
    op.set_missing_mode(missing_mode)
    X = csa.DeviceArray.from_host(np.random.default_rng(7).standard_normal((N, k)))
    Y = csa.DeviceArray(N, k)
    for _ in range(warmup):
        op.apply_device(X.ptr, N, k, Y.ptr, N)
    csa.synchronize()
    phase = np.zeros(5)
    t0 = time.perf_counter()
    for _ in range(steps):
        op.apply_device(X.ptr, N, k, Y.ptr, N)
        phase += op.last_timing_ms()
    csa.synchronize()
    dt = time.perf_counter() - t0
    ms_step = dt * 1e3 / steps
    phase /= steps
    value = N * M / (dt / steps)
    log(f"K.V n={N} m={M} k={k} {prec}: {ms_step:.3f} ms/step, phases ms {np.round(phase, 4).tolist()}")
    # dominant kernel: the larger of the two genotype products (Z^T V on the SNP-major copy,
    # Z W on the sample-major copy); algorithmic bytes = the packed matrix read once.
    gt_bytes = M * ((N + 3) // 4)
    g_bytes = N * ((M + 3) // 4)
    if phase[1] >= phase[3]:
        name, kms, kbytes, key = "k_geno1s (Z^T V, SNP-major)", phase[1], gt_bytes, "k_geno1s"
    else:
        name, kms, kbytes, key = "k_geno2s (Z W, sample-major)", phase[3], g_bytes, "k_geno2s"
    achieved = kbytes / (kms * 1e-3) / 1e9
    # int8 MFMA work the exact formulation issues in that kernel: 2 products (x~, o) x L limbs of
    # 16 x 16 x 64 MACs per 16 rows x 64 k per group, i.e. 2 * 2 L * 16 * rows * k_len ops
    limbs = 3 if prec == "int8_l3" else 2
    groups = (k + 15) // 16
    rows_pad = ((M if key == "k_geno1s" else N) + 255) // 256 * 256
    k_pad = ((N if key == "k_geno1s" else M) + 255) // 256 * 256
    ops_issued = 2.0 * 2 * limbs * 16 * groups * float(rows_pad) * float(k_pad)
    mfma_rate = ops_issued / (kms * 1e-3) / 1e12
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=profiled_traffic(key, N, M, k), kernel=name,
                    kernel_ms=round(kms, 4), bytes_per_launch=kbytes,
                    apply_frac=round(gt_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                    apply_frac_two_reads=round(2 * gt_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                    int8_mfma_tops_issued=round(mfma_rate, 1), int8_mfma_frac=round(mfma_rate / INT8_PEAK_TOPS, 4))
    out = dict(ms_per_step=ms_step, genotypes_per_s=value, phases_ms=np.round(phase, 4).tolist(), roofline=roofline)
    if extras:
        # the same apply through the host-buffer entry point (PCIe transfers of X and Y included)
        Xh = np.asfortranarray(np.random.default_rng(7).standard_normal((N, k)))
        op.apply(Xh)
        t1 = time.perf_counter()
        for _ in range(3):
            op.apply(Xh)
        out["host_buffer_genotypes_per_s"] = N * M / ((time.perf_counter() - t1) / 3)
        # few right-hand sides (limbs packed into one MFMA tile when k L <= 16), device buffers
        small_k = {}
        for ks in (1, 4):
            if ks >= k:
                continue
            for _ in range(2):
                op.apply_device(X.ptr, N, ks, Y.ptr, N)
            csa.synchronize()
            t1 = time.perf_counter()
            for _ in range(10):
                op.apply_device(X.ptr, N, ks, Y.ptr, N)
            csa.synchronize()
            small_k[f"k{ks}_ms"] = (time.perf_counter() - t1) / 10 * 1e3
        out["small_k_apply_ms"] = small_k or None
    X.close()
    Y.close()
    op.close()
    return out


def fp32_leg(csa, reader, N, M, k, reps=2):
    """The plain FP32 GPU comparator: the same K.V with genotypes decoded to FP32 values and FP32
    FMAs on the vector ALUs (no MFMA, no split): per-SNP value tables, right-hand sides staged in
    LDS and reused over four rows per lane, one FMA per genotype and column in Z^T V, two in Z W.
    Reported with its FP32 rate (3 n m k FMAs per apply)."""
    op = csa.PlinkKinshipOperator(reader)
    op.set_precision(csa.CS_PREC_FP32)
    X = csa.DeviceArray.from_host(np.random.default_rng(7).standard_normal((N, k)))
    Y = csa.DeviceArray(N, k)
    op.apply_device(X.ptr, N, k, Y.ptr, N)
    csa.synchronize()
    t1 = time.perf_counter()
    for _ in range(reps):
        op.apply_device(X.ptr, N, k, Y.ptr, N)
    csa.synchronize()
    dt = (time.perf_counter() - t1) / reps
    X.close()
    Y.close()
    op.close()
    tflops = 2.0 * 3.0 * N * M * k / dt / 1e12
    log(f"K.V plain FP32: {dt * 1e3:.2f} ms/step, {tflops:.1f} TFLOP/s")
    return dict(ms_per_step=dt * 1e3, genotypes_per_s=N * M / dt, fp32_tflops=tflops)


def grm_leg(csa, reader, N, M):
    """Exact centred GRM (alg 1) of the whole workload into a device-resident GCTA triangle.

    tops_full_square = 2 n^2 m / wall time (the customary full-square op count; the build computes
    only the lower triangle).  The MFMA roofline counts the int8 MFMA ops actually issued by the
    product kernels (T1 plus the digit products that carry missing genotypes) over their time."""
    nt = N * (N + 1) // 2
    d_tri = csa.DeviceArray((nt + 1) // 2)
    csa.synchronize()
    t0 = time.perf_counter()
    s2pq, nv = csa.grm_build_device(reader, d_tri.ptr)
    csa.synchronize()
    dt = time.perf_counter() - t0
    ms_prod, ms_comb, launches, ops_issued = csa.GrmBuilder().last_timing_ms()
    achieved = ops_issued / (ms_prod * 1e-3) / 1e12
    d_tri.close()
    return dict(ms=dt * 1e3, ms_products=ms_prod, ms_combine=ms_comb, launches=launches,
                int8_ops_issued=ops_issued,
                sum_2pq=s2pq, n_valid=nv, tops_full_square=2.0 * N * N * M / dt / 1e12,
                roofline=dict(bound="mfma", achieved=round(achieved, 1), peak=INT8_PEAK_TOPS,
                              unit="TOP/s (int8)", frac=round(achieved / INT8_PEAK_TOPS, 4),
                              kernel="k_grm_xpass (all launches)"))


def grm_std_leg(csa, reader, N, M):
    """Standardised GRM (alg 0: FP16 hi/lo split products on MFMA, FP32 accumulation flushed to
    FP64) of the whole workload into a device-resident GCTA triangle.  The roofline counts the
    FP16 MFMA flops issued (3 split products per lower-triangle wave tile) over the product time."""
    nt = N * (N + 1) // 2
    d_tri = csa.DeviceArray((nt + 1) // 2)
    csa.synchronize()
    t0 = time.perf_counter()
    s2pq, nv = csa.grm_build_device(reader, d_tri.ptr, alg=0)
    csa.synchronize()
    dt = time.perf_counter() - t0
    ms_prod, ms_fin, launches, flops = csa.GrmBuilder().last_timing_ms()
    achieved = flops / (ms_prod * 1e-3) / 1e12
    d_tri.close()
    return dict(ms=dt * 1e3, ms_products=ms_prod, ms_final=ms_fin, launches=launches,
                fp16_flops_issued=flops, n_valid=nv, tops_full_square=2.0 * N * N * M / dt / 1e12,
                roofline=dict(bound="mfma", achieved=round(achieved, 1), peak=FP16_PEAK_TFLOPS,
                              unit="TFLOP/s (fp16)", frac=round(achieved / FP16_PEAK_TFLOPS, 4),
                              kernel="k_grm_std (all launches)"))


def scan_leg(csa, reader, N, M, T=16):
    """PCG-projected LMM scan of T phenotype replicates over the whole workload (one multi-RHS
    PCG solve of T + 4 covariate + 16 calibration columns, one fused scan pass, the exact GLS stage
    at the library's default level), and the fused scan pass alone with R and the statistics in HBM
    (cs_lmm_scan_device).  genotypes/s = n m / time of one pass over all SNPs (all T traits); the
    roofline of the genotype product of the pass (packed SNP-major bytes / its event time)."""
    rng = np.random.default_rng(17)
    X = np.column_stack([np.ones(N), rng.standard_normal((N, 3))]).astype(np.float32)
    Y = rng.standard_normal((N, T))
    cal = np.sort(rng.choice(M, 16, replace=False)).astype(np.int32)
    op = csa.PlinkKinshipOperator(reader)
    csa.synchronize()
    t0 = time.perf_counter()
    out = csa.pcg_projected_scan(op, reader, None, Y, X, 0.4, 0.6, cal, rel_tol=1e-6, max_iter=200)
    csa.synchronize()
    total = time.perf_counter() - t0
    op.close()
    # the fused scan pass alone: R = [P y_1..T, X] on the device, statistics left in HBM
    dR = csa.DeviceArray.from_host(rng.standard_normal((N, T + 4)))
    cov_C = np.eye(4) * 1e-6
    kw = dict(n_score=T, alpha=1.0, cov=(T, T + 4), cov_C=cov_C)
    res, _ = csa.lmm_scan_device(reader, None, dR, **kw)
    csa.synchronize()
    t1 = time.perf_counter()
    res, (ms_counts, ms_prod, ms_epi, packed) = csa.lmm_scan_device(reader, None, dR, out=res, **kw)
    csa.synchronize()
    scan_s = time.perf_counter() - t1
    for a in [dR] + list(res.values()):
        a.close()
    gbytes = M * ((N + 3) // 4)
    achieved = gbytes / (ms_prod * 1e-3) / 1e9
    return dict(traits=T, covariates=4, pcg_iterations=out["pcg_iterations"], gamma=out["gamma"],
                max_rel_residual=out["max_rel_residual"], n_exact=out["n_exact"],
                calibration_rel_sd=out["calibration_rel_sd"], dlog10p_bound_below=out["dlog10p_bound_below"],
                total_ms=total * 1e3, scan_pass_ms=scan_s * 1e3,
                scan_device_ms=dict(class_counts=ms_counts, products=ms_prod, epilogue=ms_epi),
                scan_genotypes_per_s=N * M / scan_s, scan_trait_genotypes_per_s=N * M * T / scan_s,
                end_to_end_trait_genotypes_per_s=N * M * T / total,
                roofline=dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                              frac=round(achieved / HBM_PEAK_GBS, 4), kernel="k_geno<EPI_PQ> (g^T R)",
                              kernel_ms=round(ms_prod, 4), bytes_per_launch=gbytes),
                # kernel (a): the 2-bit decode and per-SNP class counts of the pass, one read of the
                # packed SNP-major matrix
                decode_roofline=dict(bound="hbm", achieved=round(gbytes / (ms_counts * 1e-3) / 1e9, 1),
                                     peak=HBM_PEAK_GBS, unit="GB/s",
                                     frac=round(gbytes / (ms_counts * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                                     kernel="k_wcounts_pc (decode + class counts)", kernel_ms=round(ms_counts, 4),
                                     bytes_per_launch=gbytes))


# GRCh38 autosome lengths (Mb), the chromosome proportions of the LOCO workload
CHROM_MB = [248.96, 242.19, 198.30, 190.22, 181.54, 170.81, 159.35, 145.14, 138.39, 133.80, 135.09,
            133.28, 114.36, 107.04, 101.99, 90.34, 83.26, 80.37, 58.62, 64.44, 46.71, 50.82]


def c3_setup(csa, n, m, T, seed):
    """The C3 genotypes and phenotypes: m SNPs over 22 chromosomes in proportion to their
    lengths, 20 causal SNPs per chromosome, h2 = 0.4, T phenotype replicates with independent
    effects and noise; X = [1, 2 N(0,1)]."""
    rng = np.random.default_rng(seed)
    reader = csa.PlinkReader.synthetic(seed, n, m, 0.01, 0.5, 0.005)
    bounds = np.concatenate([[0], np.round(np.cumsum(CHROM_MB) / np.sum(CHROM_MB) * m).astype(np.int64)])
    bounds[-1] = m
    causal = np.concatenate([rng.choice(np.arange(bounds[c], bounds[c + 1]), 20, replace=False)
                             for c in range(22)]).astype(np.int32)
    causal.sort()
    dZ = csa.DeviceArray(n, causal.size)
    csa.markers_device(reader, None, causal, dZ, min_maf=0.0)
    Z = dZ.download()
    dZ.close()
    Z /= np.maximum(Z.std(axis=0), 1e-12)
    X = np.column_stack([np.ones(n), rng.standard_normal((n, 2))]).astype(np.float32)
    G = Z @ rng.standard_normal((causal.size, T))
    G *= np.sqrt(0.4) / G.std(axis=0)
    Y = G + np.sqrt(0.6) * rng.standard_normal((n, T))
    del Z, G
    means, inv_stds = reader.getMeans(), reader.getInvStds()
    f = means / 2.0
    grm_ok = (inv_stds != 0) & (np.minimum(f, 1 - f) >= 0.01)
    return dict(rng=rng, reader=reader, bounds=bounds, X=X, Y=Y, grm_ok=grm_ok, n=n, m=m)


def loco_leg(csa, ctx, T=16, exact_log10p=0.0):
    """LOCO matrix-free PCG + LMM scan (c3_setup's genotypes and phenotypes): per chromosome one
    kinship operator without that chromosome, one multi-RHS PCG solve (T traits + 3 covariates +
    16 calibration SNPs, relative residual 1e-6), the fused scan of that chromosome's SNPs, and the
    exact GLS stage for every SNP whose smallest P over the T traits, times T, reaches
    10^-exact_log10p (0: the library default 5).  genotypes/s = n m / total time."""
    rng, reader, bounds, X, Y, grm_ok = ctx["rng"], ctx["reader"], ctx["bounds"], ctx["X"], ctx["Y"], ctx["grm_ok"]
    n, m = ctx["n"], ctx["m"]
    csa.synchronize()
    t0 = time.perf_counter()
    iters, resid, gammas, rsd, bound, tested, n_exact, ex_it = [], [], [], [], [], 0, 0, []
    for c in range(22):
        lo, hi = bounds[c], bounds[c + 1]
        idx = np.arange(m)
        loco = idx[grm_ok & ((idx < lo) | (idx >= hi))].astype(np.int32)
        snps = np.arange(lo, hi, dtype=np.int32)
        cal = np.sort(rng.choice(snps, 16, replace=False)).astype(np.int32)
        op = csa.PlinkKinshipOperator(reader, None, loco)
        out = csa.pcg_projected_scan(op, reader, None, Y[:, :T], X, 0.4, 0.6, cal, snps=snps, rel_tol=1e-6,
                                     max_iter=200, exact_log10p=exact_log10p)
        op.close()
        iters.append(out["pcg_iterations"])
        resid.append(out["max_rel_residual"])
        gammas.append(out["gamma"])
        rsd.append(out["calibration_rel_sd"])
        bound.append(out["dlog10p_bound_below"])
        tested += int(np.sum(out["passes"]))
        n_exact += int(out["n_exact"])
        ex_it.append(out["exact_pcg_iterations"])
    csa.synchronize()
    total = time.perf_counter() - t0
    res = dict(workload=f"n={n} x m={m}, 22 LOCO chromosomes, {T} traits, PCG rel. residual 1e-6, "
                        f"exact GLS at level {exact_log10p or 5:g} (+ log10 {T})",
               total_s=total, genotypes_per_s=n * m / total,
               trait_genotypes_per_s=n * m * T / total, pcg_iterations=iters,
               max_rel_residual=float(np.max(resid)), gamma_range=[float(np.min(gammas)), float(np.max(gammas))],
               n_tested=tested, exact_log10p=exact_log10p, n_exact=n_exact, exact_pcg_iterations=ex_it,
               calibration_rel_sd_range=[float(np.nanmin(rsd)), float(np.nanmax(rsd))],
               dlog10p_bound_below_max=float(np.nanmax(bound)))
    log(f"LOCO leg: {total:.1f} s for 22 chromosomes, {res['genotypes_per_s']:.3e} genotypes/s, "
        f"PCG iterations {min(iters)}-{max(iters)}, {n_exact} SNPs exact, calibration spread "
        f"{res['calibration_rel_sd_range']}")
    return res


def lowrank_leg(csa, ctx, rank=256, oversample=10):
    """The reference's live matrix-free GWAS path (--fastlmm-matrix-free, low-rank mode) on the C3
    genotypes (c3_setup, first phenotype replicate): per LOCO chromosome a kinship operator without
    that chromosome, the randomized low-rank null model (rank 256 + 10 oversampling, two power
    iterations, REML) and scanLowRank of that chromosome's SNPs.  genotypes/s = n m / total."""
    reader, bounds, X, Y, grm_ok = ctx["reader"], ctx["bounds"], ctx["X"], ctx["Y"], ctx["grm_ok"]
    n, m = ctx["n"], ctx["m"]
    y = np.ascontiguousarray(Y[:, 0])
    csa.synchronize()
    t0 = time.perf_counter()
    t_null, ranks, h2, tested = 0.0, [], [], 0
    for c in range(22):
        lo, hi = bounds[c], bounds[c + 1]
        idx = np.arange(m)
        loco = idx[grm_ok & ((idx < lo) | (idx >= hi))].astype(np.int32)
        t1 = time.perf_counter()
        op = csa.PlinkKinshipOperator(reader, None, loco)
        lr = csa.LowRankLmm(op, y, X, rank=rank, oversample=oversample)
        t_null += time.perf_counter() - t1
        out = lr.scan(reader, None, np.arange(lo, hi, dtype=np.int32))
        d = lr.diag
        ranks.append(int(d["effective_rank"]))
        h2.append(d["sigma2_g"] / (d["sigma2_g"] + d["sigma2_e"]))
        tested += int(np.sum(out["passes"]))
        lr.close()
        op.close()
    csa.synchronize()
    total = time.perf_counter() - t0
    res = dict(workload=f"n={n} x m={m}, 22 LOCO chromosomes, 1 trait, low-rank rank {rank} + {oversample}",
               total_s=total, null_model_s=t_null, genotypes_per_s=n * m / total,
               effective_rank=[min(ranks), max(ranks)], h2_range=[float(np.min(h2)), float(np.max(h2))],
               n_tested=tested)
    log(f"low-rank leg: {total:.1f} s for 22 chromosomes (null models {t_null:.1f} s), "
        f"{res['genotypes_per_s']:.3e} genotypes/s, h2 {res['h2_range'][0]:.3f}-{res['h2_range'][1]:.3f}")
    return res


def mv_leg(csa, n=50_000, m=300_000, d=4, seed=20261207, n_check=32):
    """Multi-trait LMM workload: d = 4 correlated traits, Vg ~ Wishart(4, 0.3 I),
    Ve ~ Wishart(4, 0.7 I), 3 records per individual averaged, 500 causal SNPs; standardised GRM
    (alg 0, FP16-split MFMA), MvEigenVce prepare (QR + dense eigendecomposition: rocSOLVER dsyevd
    up to n = 32766, the 64-bit tridiagonal divide-and-conquer path beyond), AI-REML, and the
    multivariate Wald scan of all m SNPs.  The scan is checked against an FP64 recomputation
    (W_proj^T g in FP64 on the host) for n_check SNPs."""
    rng = np.random.default_rng(seed)
    out = dict(workload=f"n={n} x m={m}, {d} traits (3 records each, averaged)")
    t0 = time.perf_counter()
    reader = csa.PlinkReader.synthetic(seed, n, m, 0.05, 0.5, 0.01)
    A = rng.normal(0, np.sqrt(0.3), (4, d))
    B = rng.normal(0, np.sqrt(0.7), (4, d))
    Vg, Ve = A.T @ A, B.T @ B
    causal = np.sort(rng.choice(m, 500, replace=False)).astype(np.int32)
    dZ = csa.DeviceArray(n, causal.size)
    csa.markers_device(reader, None, causal, dZ, min_maf=0.0)
    Z = dZ.download()
    dZ.close()
    Z /= np.maximum(Z.std(axis=0), 1e-12)
    g = Z @ (rng.standard_normal((causal.size, d)) @ np.linalg.cholesky(Vg).T) / np.sqrt(causal.size)
    e = np.mean([rng.standard_normal((n, d)) @ np.linalg.cholesky(Ve).T for _ in range(3)], axis=0)
    X = np.column_stack([np.ones(n), rng.standard_normal((n, 2))]).astype(np.float32)
    Y = X[:, 1:].astype(np.float64) @ rng.normal(0, 0.2, (2, d)) + g + e
    out["setup_s"] = time.perf_counter() - t0
    csa.synchronize()
    t0 = time.perf_counter()
    raw = csa.DeviceArray(n, n)
    S, _, nv = csa.grm_build_raw(reader, raw, min_maf=0.01, alg=0)
    G = csa.DeviceArray(n, n)
    csa.grm_loco_subset_scaled(raw, None, n, None, nv, 1.0, S, 0.0, G)
    raw.close()
    csa.synchronize()
    out["grm_s"] = time.perf_counter() - t0
    log(f"mv: standardised GRM {out['grm_s']:.1f} s")
    t0 = time.perf_counter()
    mv = csa.MvLmm(Y, X)
    mv.prepare(G)
    G.close()
    out["prepare_s"] = time.perf_counter() - t0
    log(f"mv: prepare (QR, A^T G A, eigendecomposition, W) {out['prepare_s']:.1f} s")
    t0 = time.perf_counter()
    conv, em_it, ai_it = mv.runNullModel("aireml")
    out["null_s"] = time.perf_counter() - t0
    Vg_hat, Ve_hat, _, _, ll = mv.variance()
    out.update(em_iterations=em_it, ai_iterations=ai_it, converged=conv,
               vg_rel_err_vs_truth=float(np.linalg.norm(Vg_hat - Vg) / np.linalg.norm(Vg)),
               ve_rel_err_vs_truth=float(np.linalg.norm(Ve_hat - Ve / 3) / np.linalg.norm(Ve / 3)))   # mean of 3 records
    log(f"mv: AI-REML {out['null_s']:.2f} s (EM {em_it}, AI {ai_it})")
    csa.synchronize()
    t0 = time.perf_counter()
    res = mv.scan(reader)
    dt = time.perf_counter() - t0
    out.update(scan_s=dt, scan_genotypes_per_s=n * m / dt, n_tested=int(np.sum(res["status"] >= 0)),
               rotation_fp16_tflops_issued=3 * 2.0 * m * n * (n - 3) / dt / 1e12)
    log(f"mv: scan of {m} SNPs {dt:.2f} s ({out['scan_genotypes_per_s']:.3e} genotypes/s)")
    # FP64 recomputation for n_check SNPs
    snps = np.sort(rng.choice(m, n_check, replace=False)).astype(np.int32)
    dG = csa.DeviceArray(n, n_check)
    csa.markers_device(reader, None, snps, dG, min_maf=0.0)
    Gm = dG.download()
    dG.close()
    D, Uty, W = mv.eigen(with_w=True)
    Sr = W.T @ Gm
    del W
    Sinv = np.linalg.inv(D[:, None, None] * Vg_hat + Ve_hat)
    H = np.einsum("is,iab->sab", Sr * Sr, Sinv)
    r = Sr.T @ np.einsum("iab,ib->ia", Sinv, Uty)
    beta = np.linalg.solve(H, r[:, :, None])[:, :, 0]
    wald = np.einsum("sa,sab,sb->s", beta, H, beta)
    se = np.sqrt(np.diagonal(np.linalg.inv(H), axis1=1, axis2=2))
    from scipy.special import chdtrc
    lp = -np.log10(chdtrc(d, wald))
    lp_gpu = -np.log10(res["p"][snps])
    out.update(check_snps=n_check,
               max_abs_dlog10p_vs_fp64=float(np.max(np.abs(lp_gpu - lp))),
               max_dbeta_over_se_vs_fp64=float(np.max(np.abs(res["beta"][snps] - beta) / se)),
               max_rel_dwald_vs_fp64=float(np.max(np.abs(res["wald"][snps] - wald) / (wald + 1.0))))
    log(f"mv: vs FP64 on {n_check} SNPs: |dlog10P| {out['max_abs_dlog10p_vs_fp64']:.2e}, "
        f"|dbeta|/SE {out['max_dbeta_over_se_vs_fp64']:.2e}")
    mv.close()
    reader.close()
    return out


def serving_leg(csa, n=100_000, m=1_000_000, n_requests=200, rate=5.0, seed=20261206, budget_gib=64.0):
    """Poisson stream of scan requests (rate req/s) over a resident n x m matrix with an HBM budget
    for the requests' buffers; each request: a random cohort (log-uniform 10k..n), 1-64 traits,
    0-10 covariates (+ intercept), h2 ~ U(0.1, 0.8), 16 calibration SNPs, a scan of all m SNPs with
    the exact GLS stage at its default threshold.  Latency = arrival to the client holding the full
    results (freq, n_obs, pass, beta, SE, P, log10 P of every SNP and trait) on the host; p50 / p99
    over all requests and over the second half of the arrivals (steady state), the service rate
    (completions per second while requests were arriving or queued) beside the arrival rate."""
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    reader = csa.PlinkReader.synthetic(seed, n, m, 0.05, 0.5, 0.005)
    srv = csa.ScanServer(reader, max_columns=512, hbm_budget_bytes=int(budget_gib * 2**30))
    setup = time.perf_counter() - t0
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    reqs = []
    for _ in range(n_requests):
        nc = int(np.exp(rng.uniform(np.log(10_000), np.log(n))))
        keep = np.sort(rng.choice(n, nc, replace=False)).astype(np.int32)
        T = int(rng.integers(1, 65))
        c = int(rng.integers(0, 11)) + 1
        # column-major, the layout the server's submit takes (no conversion copy at submission)
        X = np.asfortranarray(np.column_stack([np.ones(nc)] + [rng.standard_normal(nc) for _ in range(c - 1)]),
                              dtype=np.float32)
        Y = np.asfortranarray(rng.standard_normal((nc, T)))
        h2 = float(rng.uniform(0.1, 0.8))
        cal = np.sort(rng.choice(m, 16, replace=False)).astype(np.int32)
        reqs.append(dict(Y=Y, X=X, sigma2_g=h2, sigma2_e=1.0 - h2, cal_snps=cal, keep=keep))
    ids, submitted, done, done_dev, fetch = {}, 0, {}, {}, {}
    start = time.perf_counter()

    def fetch_one(rid, i):
        # the client thread holding the results: the copy to host runs beside the server's steps
        res = srv.result(rid)
        t = time.perf_counter() - start - arrivals[i]
        assert res["log10p"].shape == (m, reqs[i]["Y"].shape[1])
        ne = int(res["stats"]["n_exact"])
        del res
        srv.release(rid)
        return t, ne

    with ThreadPoolExecutor(max_workers=1) as pool:
        while len(done) < n_requests:
            now = time.perf_counter() - start
            while submitted < n_requests and arrivals[submitted] <= now:
                ids[srv.submit(**reqs[submitted])] = submitted
                submitted += 1
            if srv.step() == (0, 0, 0):
                wait = arrivals[submitted] - (time.perf_counter() - start) if submitted < n_requests else 0.01
                time.sleep(max(0.0, min(0.01, wait)))
            for rid, i in list(ids.items()):
                if rid in fetch:
                    if rid not in done and fetch[rid].done():
                        done[rid] = fetch[rid].result()
                    continue
                st, info = srv.status(rid)
                if st == 2:
                    done_dev[rid] = time.perf_counter() - start - arrivals[i]
                    fetch[rid] = pool.submit(fetch_one, rid, i)
                elif st == -1:
                    raise RuntimeError(f"request {i} failed")
            if len(fetch) == n_requests and submitted == n_requests:
                for rid, f in fetch.items():
                    if rid not in done:
                        done[rid] = f.result()
    n_exact = sum(ne for _, ne in done.values())
    genos = float(sum(reqs[i]["keep"].size * m * reqs[i]["Y"].shape[1] for i in ids.values()))
    wall = time.perf_counter() - start
    order = sorted(ids.items(), key=lambda kv: kv[1])
    lat = np.array([done[rid][0] for rid, _ in order])
    lat_dev = np.array([done_dev[rid] for rid, _ in order])
    half = lat[n_requests // 2:]
    srv.close()
    reader.close()
    return dict(workload=f"n={n} x m={m} resident, {n_requests} requests, Poisson {rate} req/s, "
                         f"HBM budget {budget_gib:g} GiB",
                p50_latency_s=float(np.percentile(lat, 50)), p99_latency_s=float(np.percentile(lat, 99)),
                p50_latency_second_half_s=float(np.percentile(half, 50)),
                p99_latency_second_half_s=float(np.percentile(half, 99)),
                p50_device_done_s=float(np.percentile(lat_dev, 50)),
                arrival_rate_per_s=float(n_requests / arrivals[-1]), service_rate_per_s=float(n_requests / wall),
                wall_s=wall, trait_genotypes_per_s=genos / wall, n_exact=n_exact, setup_s=setup)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=200_000)
    ap.add_argument("--m", type=int, default=600_000)
    ap.add_argument("--k", type=int, default=16)
    ap.add_argument("--prec", choices=["int8_l3", "int8_l2"], default="int8_l3")
    ap.add_argument("--missing-mode", type=int, choices=[0, 1, 2], default=0,
                    help="K.V missing-genotype form: 0 auto, 1 dense O products, 2 sparse missing lists")
    ap.add_argument("--kv-only", action="store_true", help="only the headline K.V applies (for profiling)")
    ap.add_argument("--cpu-slice-snps", type=int, default=160_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-loco", action="store_true")
    ap.add_argument("--loco-exact-log10p", type=float, default=0.0,
                    help="exact GLS stage level of the LOCO leg (-log10 of the trait-adjusted P; 0 = the "
                         "library default 5)")
    ap.add_argument("--no-c2", action="store_true", help="skip the n=50k x m=500k legs (K.V, GRMs, scan)")
    ap.add_argument("--miss", type=float, default=0.005,
                    help="missing-genotype rate of the n=50k x m=500k genotypes (the sparse K.V form runs at "
                         "0.15 %% or less)")
    ap.add_argument("--no-grm", action="store_true")
    ap.add_argument("--no-scan", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serving-requests", type=int, default=200)
    ap.add_argument("--mv", action="store_true", help="also run the multi-trait LMM leg")
    ap.add_argument("--lowrank", action="store_true",
                    help="also run the low-rank matrix-free GWAS leg (the reference path, headline genotypes)")
    ap.add_argument("--mv-n", type=int, default=50_000)
    ap.add_argument("--mv-m", type=int, default=300_000)
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 FETCH_SIZE / WRITE_SIZE child runs (roofline.traffic then comes "
                         "from the committed profile)")
    ap.add_argument("--grm-only", action="store_true",
                    help="only the n=50k x m=500k GRM legs (exact and standardised; for counter runs)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 2 GiB device copy 3 times (HBM counter calibration)")
    args = ap.parse_args()

    if args.copy_probe:
        import cosmicscanner_amd as csa
        nbytes = 2 << 30
        src = csa.DeviceArray(nbytes // 8)
        dst = csa.DeviceArray(nbytes // 8)
        for _ in range(3):
            csa._check(csa.lib().cs_dev_copy(dst.ptr, src.ptr, nbytes))
        print(json.dumps({"copy_probe_bytes": nbytes, "launches": 3}))
        return 0

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                                   "(multi-GPU sharding is design-only, see DESIGN.md)"}))
        return 2

    import cosmicscanner_amd as csa

    if args.grm_only:
        reader2 = csa.PlinkReader.synthetic(20261204, 50_000, 500_000, 0.01, 0.5, 0.005)
        grm = grm_leg(csa, reader2, 50_000, 500_000)
        grm["standardised"] = grm_std_leg(csa, reader2, 50_000, 500_000)
        reader2.close()
        print(json.dumps({"grm": grm}), flush=True)
        return 0

    N, M, k = args.n, args.m, args.k
    seed = 20261205
    t0 = time.perf_counter()
    ctx = c3_setup(csa, N, M, 16, seed)
    log(f"resident genotypes {ctx['reader'].device_bytes() / 2**30:.2f} GiB, setup {time.perf_counter() - t0:.1f} s")
    kv = kv_leg(csa, ctx["reader"], N, M, k, args.steps, args.warmup, args.prec, args.missing_mode,
                extras=not args.kv_only)
    value, ms_step = kv["genotypes_per_s"], kv["ms_per_step"]
    if not args.kv_only and not args.no_traffic:
        tr = measured_traffic(args)
        key = "k_geno2s" if "k_geno2s" in kv["roofline"]["kernel"] else "k_geno1s"
        if tr is not None:
            kv["roofline"]["traffic"] = round(tr[key], -6)
            kv["roofline"]["traffic_source"] = (f"rocprofv3 FETCH_SIZE x{tr['fetch_scale']} + WRITE_SIZE "
                                               f"x{tr['write_scale']} in this run (plain-copy calibration)")
            kv["roofline"]["traffic_other_product"] = round(tr["k_geno1s" if key == "k_geno2s" else "k_geno2s"], -6)
            log(f"measured HBM traffic per launch: k_geno1s {tr['k_geno1s'] / 1e9:.2f} GB, "
                f"k_geno2s {tr['k_geno2s'] / 1e9:.2f} GB")
        elif kv["roofline"]["traffic"] is not None:
            kv["roofline"]["traffic_source"] = "committed rocprofv3 summary in profiles/"
    if not args.kv_only:
        # context for the roofline: the same kernel against a measured plain copy (read + write)
        cr = copy_rate(csa)
        kv["roofline"]["copy_gbs_measured"] = round(cr, 1)
        kv["roofline"]["frac_of_copy"] = round(kv["roofline"]["achieved"] / cr, 4)
        log(f"plain device copy: {cr:.0f} GB/s (read + write); dominant K.V kernel at "
            f"{kv['roofline']['frac_of_copy']:.3f} of it")

    fp32 = None
    if not args.no_fp32 and not args.kv_only:
        fp32 = fp32_leg(csa, ctx["reader"], N, M, k)

    locores = None
    if not args.no_loco and not args.kv_only:
        locores = loco_leg(csa, ctx, exact_log10p=args.loco_exact_log10p)

    lowres = None
    if args.lowrank and not args.kv_only:
        lowres = lowrank_leg(csa, ctx)
    ctx["reader"].close()
    del ctx

    kv_c2 = grm = scan = None
    if not args.no_c2 and not args.kv_only:
        N2, M2 = 50_000, 500_000
        reader2 = csa.PlinkReader.synthetic(20261204, N2, M2, 0.01, 0.5, args.miss)
        kv_c2 = kv_leg(csa, reader2, N2, M2, 16, args.steps, args.warmup, args.prec, args.missing_mode, extras=True)
        if not args.no_grm:
            grm = grm_leg(csa, reader2, N2, M2)
            log(f"GRM: {grm['ms']:.1f} ms, {grm['tops_full_square']:.1f} TOPS (2 n^2 m), "
                f"int8 products {grm['ms_products']:.1f} ms in {grm['launches']} launches")
            grm["standardised"] = grm_std_leg(csa, reader2, N2, M2)
            gs = grm["standardised"]
            log(f"standardised GRM: {gs['ms']:.1f} ms, {gs['tops_full_square']:.1f} TOPS (2 n^2 m), "
                f"fp16 products {gs['ms_products']:.1f} ms = {gs['roofline']['achieved']} TFLOP/s issued")
        if not args.no_scan:
            scan = scan_leg(csa, reader2, N2, M2)
            log(f"projected LMM scan: {scan['total_ms']:.1f} ms end to end ({scan['pcg_iterations']} PCG "
                f"iterations), scan pass {scan['scan_pass_ms']:.1f} ms = {scan['scan_genotypes_per_s']:.3e} genotypes/s")
        reader2.close()

    serving = None
    if not args.no_serving and not args.kv_only:
        serving = serving_leg(csa, n_requests=args.serving_requests)
        log(f"serving: p50 {serving['p50_latency_s']:.2f} s, p99 {serving['p99_latency_s']:.2f} s "
            f"(second half {serving['p50_latency_second_half_s']:.2f} / {serving['p99_latency_second_half_s']:.2f} s), "
            f"service {serving['service_rate_per_s']:.2f} req/s, {serving['trait_genotypes_per_s']:.3e} trait-genotypes/s")

    mvres = None
    if args.mv and not args.kv_only:
        mvres = mv_leg(csa, n=args.mv_n, m=args.mv_m)

    cpu = None
    if not args.no_cpu and not args.kv_only:
        cpu = cpu_baseline(N, args.cpu_slice_snps, k, seed)
        log(f"cpu baseline: {cpu['value']:.3e} genotypes/s on {cpu['cores']} threads")

    out = {
        "metric": "genotypes/s in K.V (matrix-free kinship apply, k=16) on one MI355X",
        "value": value,
        "unit": "genotypes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8 MFMA (exact genotypes, 3x8-bit fixed-point RHS limbs), int32 acc, fp64 epilogue"
        if args.prec == "int8_l3" else "int8 MFMA (2x8-bit limbs), int32 acc, fp64 epilogue",
        "data": "synthetic HWE genotypes generated on device (p~U(0.01,0.5), 0.5% missing), random RHS",
        "config": {"workload": f"LOCO matrix-free LMM genotypes n={N:,} x m={M:,}: K.V apply, k={k} (all SNPs)",
                   "n": N, "m": M, "k": k, "precision": args.prec},
        "phases_ms": kv["phases_ms"],
        "host_buffer_genotypes_per_s": kv.get("host_buffer_genotypes_per_s"),
        "small_k_apply_ms": kv.get("small_k_apply_ms"),
        "fp32_gpu_genotypes_per_s": fp32["genotypes_per_s"] if fp32 else None,
        "speedup_vs_fp32_gpu": (value / fp32["genotypes_per_s"]) if fp32 else None,
        "fp32_gpu_tflops": fp32["fp32_tflops"] if fp32 else None,
        "roofline": kv["roofline"],
        "loco_scan": locores,
        "lowrank_scan": lowres,
        "kv_n50k_m500k": kv_c2,
        "grm": grm,
        "lmm_scan": scan,
        "serving": serving,
        "multivariate": mvres,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
