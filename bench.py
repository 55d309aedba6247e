"""Benchmark: GPU-resident SP2 purification on ELLPACK-R (thresholded X^2
SpGEMM) on one MI355X.

Workload (the 32^3 SP2 configuration as written): 32^3 cubic lattice, N =
32768, one orbital per site, onsite +-1 checkerboard, nearest-neighbour
hopping -0.5 + U[-0.05, 0.05] disorder per bond, threshold 1e-5, nocc = N/2,
Gershgorin bounds, stop at ||X^2 - X||_F <= 1e-8 or 60 iterations.  The
density matrix keeps ~2440 (at most ~2530) entries per row above 1e-5, so
the ELLPACK-R width is M = 3072; at the configured M = 256 the reference
exits with "Number of non-zeroes per row > M" (and so does this library;
tests/test_gpu_configs.py).  At this threshold ||X^2 - X||_F stalls near
4e-3 .. 1e-2 (the truncation floor): no solve reaches 1e-8, and every solve
runs its 60 iterations; the JSON line says so (converged, idempotency).

A step is one SP2 iteration in the solve's steady regime: a solve first
runs SETTLE iterations untimed (the fill reaches ~2440 entries per row),
then each step continues it by one iteration (bml_hip_sp2_ellpack with
resume): tr X and sum |X_ij|^2, the branch, the fused window product (X^2's
upper triangle, X <- alpha X + beta X^2 thresholded, ||X - X^2||_F), the
mirror.  The timed steps run the mixed-precision iteration (FP32 storage
and products, FP64 window sums, traces, norm and branch); the same number
of FP64 steps is timed afterwards.  `value` is the executed X^2 GFLOP/s of
the timed steps (2 flops per partial product formed; the upper triangle is
formed, so about half of the full product's), whole-iteration time.  Full
solves (mixed: FP32 iterations, the last 4 in FP64; and FP64) are timed once
each beside it.  Inputs are synthetic (seeded generator, bml_amd.generators).

Prints one JSON line (rank 0) with the roofline of the dominant kernel
(HIP-event durations inside the timed region) and the reference's CPU
bml_multiply_x2 on a row block of the same steady iterate (oracle/_ref,
bml's own C code) as cpu_baseline.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}
L, THR, TOL, MAXIT, REFINE = 32, 1e-5, 1e-8, 60, 4
HOP, DISORDER = -0.5, 0.05
M = 3072      # holds the density matrix's ~2530 entries per row at 1e-5
SETTLE = 24   # iterations before the timed steps (the fill is steady by then)
# the batched and single-system legs below keep the earlier lattice
# parameters where the configured ones do not fit (see their docstrings)
HOP_SMALL, DISORDER_SMALL = -0.1, 0.01


def bytes_alg(prof, p, s):
    """Streamed-Gustavson bytes of the precision-p row-product launches in a
    profile: A rows (s = element + index bytes per entry) and nnz, the
    entries of B streamed (one referenced B row per A entry; the window
    engine's row blocks stream each row referenced by any of a block's rows
    once for the block, which the profile counts) plus one nnz read per A
    entry, the C rows and nnz."""
    streamed = prof.streamed[p] or prof.products[p]
    return (s * prof.in_nnz[p] + 4 * prof.rows[p] + s * streamed
            + 4 * prof.in_nnz[p] + s * prof.out_nnz[p] + 4 * prof.rows[p])


def roofline(prof, p, s, name):
    n = prof.launches[p]
    if not n:
        return None
    avg_ms = prof.kernel_ms[p] / n
    per_launch = bytes_alg(prof, p, s) / n
    ach = per_launch / (avg_ms * 1e-3) / 1e9
    return {"kernel": name, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
            "avg_launch_ms": round(avg_ms, 4), "launches": n,
            "bytes_alg_per_launch": int(per_launch),
            "products_per_launch": int(prof.products[p] / n),
            "streamed_per_launch": int((prof.streamed[p] or prof.products[p]) / n),
            "touched_per_row": round(prof.touched[p] / max(prof.rows[p], 1), 1),
            "table_overflow_rows_per_launch": round(prof.ovf_rows[p] / n, 1),
            "gflops": round(2.0 * prof.products[p] / (prof.kernel_ms[p] * 1e-3) / 1e9, 2)}


def ell_max_diff(A, B):
    """max |A - B| over the union of the two sparsity patterns."""
    ra, ca, va = A.to_coo()
    rb, cb, vb = B.to_coo()
    ka, kb = ra * A.N + ca, rb * B.N + cb
    keys = np.union1d(ka, kb)
    xa = np.zeros(len(keys), va.dtype)
    xb = np.zeros(len(keys), vb.dtype)
    xa[np.searchsorted(keys, ka)] = va
    xb[np.searchsorted(keys, kb)] = vb
    return float(np.abs(xa - xb).max(initial=0.0))


def sp2_steps(hip, G, dtype, steps, warmup):
    """SETTLE untimed iterations of a solve in precision dtype, warmup more,
    then `steps` timed ones, one bml_hip_sp2_ellpack(resume) call each; the
    profiling window covers the timed steps.  Returns (seconds, profile,
    history of the timed steps, X)."""
    H = G.lattice_hamiltonian(L, L, L, M=16, hopping=HOP, disorder=DISORDER, dtype=dtype)
    dH = hip.from_ell(H)
    X = hip.zero_matrix(H.precision, H.N, M)
    hip.sp2(dH, X, H.N / 2, threshold=THR, tol=TOL, maxit=SETTLE)
    for _ in range(warmup):
        hip.sp2(dH, X, H.N / 2, threshold=THR, tol=TOL, maxit=1, resume=True)
    hip.synchronize()
    hip.profile_begin()
    t0 = time.perf_counter()
    hist = []
    for _ in range(steps):
        r, h = hip.sp2(dH, X, H.N / 2, threshold=THR, tol=TOL, maxit=1, resume=True)
        hist.append(h[0])
    hip.synchronize()
    t1 = time.perf_counter()
    prof = hip.profile_end()
    dH.free()
    return t1 - t0, prof, np.array(hist), X


def full_solve(hip, dH, N, mixed):
    """One whole solve (60 iterations unless it converges), timed."""
    X = hip.zero_matrix(dH.precision, N, M)
    hip.synchronize()
    t0 = time.perf_counter()
    r, hist = hip.sp2(dH, X, N / 2, threshold=THR, tol=TOL, maxit=MAXIT, mixed=mixed,
                      refine_iters=REFINE, mixed_switch=False)
    hip.synchronize()
    dt = time.perf_counter() - t0
    return dt, r, hist, X


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(Xd, target_s=12.0):
    """The reference's bml_multiply_x2 (bml's own OpenMP C code, built from
    its sources in oracle/_ref) on the host cores, on a row block of the
    same FP64 steady iterate: bml_multiply_x2 runs over the matrix's row
    domain (bml_multiply_ellpack_typed.c:155-157), set to rows [0, R), with
    R sized from a 256-row run to about target_s seconds.  GFLOP/s = 2 x the
    block's partial products / time (the full product's, which the
    reference forms)."""
    drv = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    if not os.path.exists(drv):
        return None
    threads = len(os.sched_getaffinity(0))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
               OMP_PLACES="cores", OMP_STACKSIZE="64M")
    A = Xd.get()
    A = A.with_m(max(1, int(A.nnz.max())))
    N = A.N

    def run(f, rows):
        out = subprocess.run([drv, "bench_x2_rows", f, "4096", THR.hex(), str(rows), "1"],
                             capture_output=True, text=True, env=env, timeout=600)
        if out.returncode != 0:
            return None
        tok = out.stdout.split()
        return float(tok[tok.index("median_ms") + 1]), int(tok[tok.index("products") + 1])

    with tempfile.TemporaryDirectory() as tmp:
        f = os.path.join(tmp, "x.ell")
        A.write(f)
        del A
        cal = run(f, 256)
        if cal is None:
            return None
        rows = int(min(N, max(256, 256 * target_s * 1e3 / max(cal[0], 1e-3))))
        got = run(f, rows)
    if got is None:
        return None
    ms, prods = got
    return {"value": round(2.0 * prods / (ms * 1e-3) / 1e9, 3), "unit": "GFLOP/s",
            "cores": threads, "kind": "reference",
            "cpu_model": cpu_model(),
            "sample": f"bml_multiply_x2 (FP64, threshold {THR}) over rows [0, {rows}) of the "
                      f"{L}^3 steady iterate (~{int(Xd.row_stats()[1] / N)} entries per row): "
                      f"{prods:.3e} partial products in {ms / 1e3:.2f} s on {threads} threads "
                      f"(OMP_PROC_BIND=close, OMP_PLACES=cores)"}


def copy_calibration(hip):
    """Three k_copy_rows launches of known HBM bytes for the FETCH_SIZE /
    WRITE_SIZE counters: an FP64 ELLPACK matrix with N = 32768 full rows of
    M = 1024 entries (403 MB, more than the 32 MB of L2 plus the 256 MB
    Infinity Cache, so every copy streams from and to HBM) copied into one
    with wider rows.  Each copy reads nnz and the N*M (value, index) entries
    and writes the same."""
    from bml_amd.ell import Ell
    n, m = 32768, 1024
    A = Ell(np.ones((n, m)), np.tile(np.arange(m, dtype=np.int32), (n, 1)),
            np.full(n, m, dtype=np.int32))
    dA = hip.from_ell(A)
    del A
    W = hip.zero_matrix(dA.precision, n, m + 64)
    hip.synchronize()
    for _ in range(3):
        hip.copy(dA, W)
    hip.synchronize()
    W.free()
    dA.free()
    return {"kernel": "k_copy_rows<double>", "launches": 3, "bytes_each_way": 4 * n + 12 * n * m}


def lib_digest():
    import hashlib
    h = hashlib.sha256()
    with open(os.path.join(ROOT, "bml_amd", "libbml_hip.so"), "rb") as f:
        h.update(f.read())
    return h.hexdigest()[:16]


def kernel_digest():
    """Digest of what the window kernel is compiled from: its source file,
    the device header and the build flags (host-side changes leave it)."""
    import hashlib
    h = hashlib.sha256()
    for f in ("hip/ellpack.hip", "hip/devtypes.h", "Makefile"):
        with open(os.path.join(ROOT, "bml_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def profiled_traffic():
    """HBM bytes per launch of the timed steps' window kernel, measured with
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on this same command and scaled
    by the copy calibration (summary committed under profiles/).  Taken only
    when the summary was measured on this very kernel source and build flags
    (their digest), so a changed kernel reports null rather than a stale
    figure."""
    f = os.path.join(ROOT, "profiles", "sp2_lattice32_window_traffic.md")
    out = {}
    if os.path.exists(f):
        lib = None
        for line in open(f):
            if line.startswith("kernel_source_sha256_16"):
                lib = line.split()[1]
            if line.startswith("traffic_bytes_per_launch"):
                for kv in line.split()[1:]:
                    k, v = kv.split("=")
                    out[int(k[1:])] = int(v)
        if lib != kernel_digest():
            return {}
    return out


def ellblock_sp2(hip, G):
    """Config 3: SP2 on ELLBLOCK, 16x16x8 sites x 32 orbitals (N = 65536,
    2048 block rows of 32), FP64 MFMA vs mixed (FP32 MFMA or bf16-input
    MFMA until the single-precision iterate stalls, then FP64).  Onsite
    shift +-6 and coupling 0.02 instead of +-2 and 0.1: with the configured
    values the random 32x32 onsite blocks close the gap (0.003 on a 5x5x4
    cut) and the density matrix fills the whole lattice; these keep a gap of
    1.9 and 25 blocks per block row above 1e-5.  MB = 160 holds the X^2
    blocks touched before thresholding (the reference's overflow test)."""
    H = G.block_lattice_hamiltonian(16, 16, 8, b=32, MB=8, onsite_shift=6.0, coupling=0.02)
    MB = 160
    dH = hip.from_blk(H, MB)
    out = {"workload": "SP2 on ELLBLOCK, 16x16x8 sites x 32 orbitals (N=65536, 32x32 blocks), "
                       "onsite shift +-6, coupling 0.02, MB=160, block threshold 1e-5, "
                       "stop at ||X^2-X||_F<=1e-8 or 60 iterations; H is exactly symmetric, so "
                       "each X^2 forms about half of its block pairs and mirrors the rest "
                       "(block_pairs_per_launch and the MFMA rates count the pairs formed)",
           "N": H.N}
    runs = [("fp64", hip.BLOCK_ENGINE_MFMA, False), ("mixed_fp32", hip.BLOCK_ENGINE_MFMA, True),
            ("mixed_bf16", hip.BLOCK_ENGINE_MFMA_BF16, True)]
    for name, engine, mixed in runs:
        hip.set_block_engine(engine)
        X = hip.block_matrix(H.precision, H.bsize, MB)
        hip.eb_sp2(dH, X, H.N / 2, threshold=THR, tol=TOL, maxit=MAXIT, mixed=mixed)  # warm-up
        hip.synchronize()
        reps = 3
        hip.profile_begin()
        t0 = time.perf_counter()
        for _ in range(reps):
            r, _h = hip.eb_sp2(dH, X, H.N / 2, threshold=THR, tol=TOL, maxit=MAXIT, mixed=mixed)
        hip.synchronize()
        dt = (time.perf_counter() - t0) / reps
        prof = hip.profile_end()
        # trace_error: |tr X - nocc| of the returned density matrix (one
        # trace after the timed solves); the result's trace_x is that of the
        # last iteration's input
        entry = {"solves_per_s": round(1.0 / dt, 3), "iterations": r.iterations,
                 "single_iterations": r.single_iterations, "converged": bool(r.converged),
                 "idempotency": r.idempotency, "trace_error": abs(hip.eb_trace(X) - H.N / 2),
                 "trace_error_last_input": abs(r.trace_x - H.N / 2)}
        for p, key in ((2, "fp64"), (1, "bf16" if engine == hip.BLOCK_ENGINE_MFMA_BF16 else "fp32")):
            n = prof.launches[p]
            if not n:
                continue
            flops = 2.0 * 32 ** 3 * prof.products[p]
            tf = flops / (prof.kernel_ms[p] * 1e-3) / 1e12
            entry[f"x2_{key}"] = {"kernel": "k_eb_numeric", "launches": n,
                                  "avg_launch_ms": round(prof.kernel_ms[p] / n, 4),
                                  "block_pairs_per_launch": int(prof.products[p] / n),
                                  "tflops": round(tf, 2), "mfma_peak_tflops": MFMA_PEAK_TFLOPS[key],
                                  "mfma_frac": round(tf / MFMA_PEAK_TFLOPS[key], 4)}
        out[name] = entry
        X.free()
    hip.set_block_engine(hip.BLOCK_ENGINE_AUTO)
    out["mixed_fp32_over_fp64"] = round(out["mixed_fp32"]["solves_per_s"] / out["fp64"]["solves_per_s"], 3)
    return out


def batched_sp2(hip, G, plain_sample=32):
    """Config 4: 512 independent SP2 solves, N log-uniform in [256, 8192]
    (near-cubic lattices), batched into one block-diagonal matrix: FP64 and
    mixed (FP32, last 4 iterations FP64), against the plain FP64 path (one
    bml_hip_sp2_ellpack solve after another, timed on the first
    plain_sample systems and scaled to 512) on the same row engine and on
    the exact engine.  Hopping -0.1 and M = 256 as in
    config 2 (at -0.5 the rows overflow M); nocc = floor(N/2)."""
    from bml_amd import generators as Gm
    sizes = Gm.batch_sizes(512)
    Hs = []
    for n in sizes:
        Lx, Ly, Lz = Gm.near_cubic(int(n))
        Hs.append(Gm.lattice_hamiltonian(Lx, Ly, Lz, M=16, hopping=HOP_SMALL, disorder=0.0))
    Mb = 256
    dHs = [hip.from_ell(H) for H in Hs]
    nocc = [float(H.N // 2) for H in Hs]
    rows = int(sum(H.N for H in Hs))
    out = {"workload": f"512 SP2 solves, N log-uniform in [256, 8192] on near-cubic lattices "
                       f"(sum N = {rows}), hopping {HOP_SMALL}, onsite +-1, M={Mb}, threshold {THR}, "
                       f"stop at ||X^2-X||_F<={TOL} or {MAXIT} iterations", "systems": len(Hs),
           "rows": rows}
    # mixed_refine6: the last 6 iterations in FP64 instead of 4 (on the hash
    # engine a single-precision iteration costs what a double one does, as
    # both sum in double, so the longer refinement is free)
    for name, mixed, refine in (("fp64", False, 4), ("mixed", True, 4), ("mixed_refine6", True, 6)):
        Xs = [hip.zero_matrix(H.precision, H.N, Mb) for H in Hs]
        hip.sp2_batch(dHs, Xs, nocc, threshold=THR, tol=TOL, maxit=MAXIT, mixed=mixed,
                      refine_iters=refine)  # warm-up
        hip.synchronize()
        t0 = time.perf_counter()
        res = hip.sp2_batch(dHs, Xs, nocc, threshold=THR, tol=TOL, maxit=MAXIT, mixed=mixed,
                            refine_iters=refine)
        hip.synchronize()
        dt = time.perf_counter() - t0
        out[name] = {"solves_per_s": round(len(Hs) / dt, 2), "seconds": round(dt, 3),
                     "converged": int(sum(r.converged for r in res)),
                     "max_trace_error": max(abs(hip.trace(X) - c) for X, c in zip(Xs, nocc)),
                     "max_trace_error_last_input": max(abs(r.trace_x - c)
                                                       for r, c in zip(res, nocc))}
        for X in Xs:
            X.free()
    # plain FP64 path: one solve after another, on the same row engine and
    # on the exact engine (the library's default)
    sub = list(range(plain_sample))
    share = sum(Hs[k].N for k in sub) / rows
    engine = hip.get_row_engine()
    for key, eng in (("plain_fp64", engine), ("plain_fp64_exact_engine", hip.ROW_ENGINE_EXACT)):
        hip.set_row_engine(eng)
        Xp = [hip.zero_matrix(Hs[k].precision, Hs[k].N, Mb) for k in sub]
        hip.sp2(dHs[0], Xp[0], nocc[0], threshold=THR, tol=TOL, maxit=MAXIT)
        hip.synchronize()
        t0 = time.perf_counter()
        for k in sub:
            hip.sp2(dHs[k], Xp[k], nocc[k], threshold=THR, tol=TOL, maxit=MAXIT)
        hip.synchronize()
        dt = time.perf_counter() - t0
        out[key] = {"solves_per_s": round(len(sub) / dt, 2),
                    "sample": f"first {len(sub)} systems ({100 * share:.1f}% of the rows)"}
        for X in Xp:
            X.free()
    hip.set_row_engine(engine)
    for name in ("mixed", "mixed_refine6", "fp64"):
        for key in ("plain_fp64", "plain_fp64_exact_engine"):
            out[f"{name}_batched_over_{key}"] = round(out[name]["solves_per_s"] /
                                                      out[key]["solves_per_s"], 2)
    return out


def x2_config1(hip, G, cpu=True):
    """Config 1: one bml_multiply_x2 (threshold 1e-5) and one bml_multiply
    (alpha 1, beta 0, A != B: the AB + add path, C with M = 1024) of the
    random banded N = 4096 matrix (~32 nnz/row within |i-j| <= 256), X^2
    into M2 = 256;
    with the reference's CPU x2 on the host beside it."""
    A = G.random_banded()
    dA = hip.from_ell(A)
    dB = hip.copy_new(dA)
    X2 = hip.zero_matrix(A.precision, A.N, 256)
    # the AB path checks the pre-threshold row against C's M, as the
    # reference does: C gets M = 1024 as in the reference probe
    Cm = hip.zero_matrix(A.precision, A.N, 1024)
    hip.multiply_x2(dA, X2, THR)
    hip.multiply(dA, dB, Cm, 1.0, 0.0, THR)
    reps = 20
    hip.synchronize()
    hip.profile_begin()
    t0 = time.perf_counter()
    for _ in range(reps):
        hip.multiply_x2(dA, X2, THR)
    hip.synchronize()
    dt = (time.perf_counter() - t0) / reps
    prof = hip.profile_end()
    t0 = time.perf_counter()
    for _ in range(reps):
        hip.multiply(dA, dB, Cm, 1.0, 0.0, THR)
    hip.synchronize()
    dtm = (time.perf_counter() - t0) / reps
    n = prof.launches[2]
    out = {"workload": "random banded N=4096, ~32 nnz/row, |i-j|<=256, X^2 threshold 1e-5 into M2=256",
           "x2_ms": round(dt * 1e3, 4), "x2_kernel_ms": round(prof.kernel_ms[2] / n, 4),
           "x2_gflops": round(2.0 * prof.products[2] / n / dt / 1e9, 2),
           "multiply_AB_plus_add_ms": round(dtm * 1e3, 4),
           "bytes_alg_gbs": round(bytes_alg(prof, 2, 12) / n / (prof.kernel_ms[2] / n * 1e-3) / 1e9, 1)}
    drv = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    if cpu and os.path.exists(drv):
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
                   OMP_PLACES="cores", OMP_STACKSIZE="64M")
        with tempfile.TemporaryDirectory() as tmp:
            f = os.path.join(tmp, "a.ell")
            A.write(f)
            r = subprocess.run([drv, "bench_x2", f, "256", THR.hex(), "20"], capture_output=True,
                               text=True, env=env, timeout=300)
        if r.returncode == 0:
            ms = float(r.stdout.split()[1])
            out["cpu_reference_x2_ms"] = ms
            out["cpu_reference_threads"] = threads
            out["gpu_over_cpu_x2"] = round(ms / (dt * 1e3), 1)
    return out


def x2_config5(hip, G):
    """Config 5: thresholded X^2 of the double complex Hermitian 64^3
    lattice (Peierls phases, N = 262144), X = the third SP2 iterate
    (~280 nnz/row), M = 512, threshold 1e-6."""
    H = G.peierls_hamiltonian(64, M=16)
    dH = hip.from_ell(H)
    X = hip.zero_matrix(H.precision, H.N, 512)
    hip.sp2(dH, X, H.N / 2, threshold=1e-6, tol=1e-8, maxit=3)
    X2 = hip.zero_matrix(H.precision, H.N, 512)
    hip.multiply_x2(X, X2, 1e-6)
    hip.synchronize()
    reps = 3
    hip.profile_begin()
    t0 = time.perf_counter()
    for _ in range(reps):
        hip.multiply_x2(X, X2, 1e-6)
    hip.synchronize()
    dt = (time.perf_counter() - t0) / reps
    prof = hip.profile_end()
    eng = "unord" if hip.get_row_engine() == hip.ROW_ENGINE_UNORDERED else "mw"
    roof = roofline(prof, 4, 20, f"k_spgemm_{eng}<cf64> (double complex X^2)")
    mx, tot = X.row_stats()
    m2, tot2 = X2.row_stats()
    return {"workload": "double complex Hermitian 64^3 lattice (Peierls phases), N=262144, "
                        "X = SP2 iterate 3, M=512, threshold 1e-6",
            "x_nnz_per_row": round(tot / H.N, 1), "x2_nnz_per_row": round(tot2 / H.N, 1),
            "x2_max_nnz": m2, "x2_ms": round(dt * 1e3, 2),
            "x2_gflops": round(8.0 * prof.products[4] / reps / dt / 1e9, 1),
            "roofline": roof}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--only-headline", action="store_true",
                    help="only the headline steps (no full solves, other configurations, CPU)")
    ap.add_argument("--only-ellblock", action="store_true",
                    help="only the ELLBLOCK SP2 measurement (16x16x8 sites x 32 orbitals)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    from bml_amd import generators as G
    from bml_amd import hip

    hip.init(0)
    if args.only_ellblock:
        print(json.dumps({"ellblock_sp2": ellblock_sp2(hip, G)}), flush=True)
        return
    # the unordered setting: the window engine for these wide rows (the
    # exact engine's LDS tables cannot hold them)
    hip.set_row_engine(hip.ROW_ENGINE_UNORDERED)
    N = L ** 3

    t_mix, p_mix, h_mix, X32 = sp2_steps(hip, G, np.float32, args.steps, args.warmup)
    X32.free()
    t_dbl, p_dbl, h_dbl, X64 = sp2_steps(hip, G, np.float64, args.steps, args.warmup)

    # dominant kernel of the timed steps: the fused window product (+ mirror)
    roof = roofline(p_mix, 1, 8, "k_spgemm_winb<float, 8, 10, 4, true, true> + "
                                 "k_sym_mirror<float> (fused FP32 SP2 step, 4-row blocks)")
    roof64 = roofline(p_dbl, 2, 12, "k_spgemm_winb<double, 8, 10, 4, true, true> + "
                                    "k_sym_mirror<double> (fused FP64 SP2 step, 4-row blocks)")
    traffic = profiled_traffic()
    for r, p in ((roof, 1), (roof64, 2)):
        if r is not None:
            r["traffic"] = traffic.get(p)
    gflops_mix = 2.0 * p_mix.products[1] / t_mix / 1e9
    gflops_dbl = 2.0 * p_dbl.products[2] / t_dbl / 1e9

    solves = None
    if not args.only_headline:
        H = G.lattice_hamiltonian(L, L, L, M=16, hopping=HOP, disorder=DISORDER)
        dH = hip.from_ell(H)
        solves = {}
        Xs = {}
        for name, mixed in (("mixed", True), ("fp64", False)):
            dt, r, hist, Xf = full_solve(hip, dH, N, mixed)
            solves[name] = {"solves_per_s": round(1.0 / dt, 4), "seconds": round(dt, 3),
                            "iterations": r.iterations, "single_iterations": r.single_iterations,
                            "converged": bool(r.converged), "idempotency": r.idempotency,
                            "trace_error": abs(hip.trace(Xf) - N / 2),
                            "trace_error_last_input": abs(r.trace_x - N / 2)}
            Xs[name] = Xf
        solves["mixed_over_fp64"] = round(solves["mixed"]["solves_per_s"] /
                                          solves["fp64"]["solves_per_s"], 3)
        solves["max_abs_diff_mixed_vs_fp64"] = ell_max_diff(Xs["mixed"].get(), Xs["fp64"].get())
        for Xf in Xs.values():
            Xf.free()
        dH.free()
    calib = copy_calibration(hip)

    line = {
        "metric": "X^2 SpGEMM GFLOP/s + SP2 solves/s, mixed vs FP64 at ||X^2-X||_F <= 1e-8",
        "value": round(gflops_mix, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * t_mix / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic",
        "config": {
            "workload": f"SP2 purification, ELLPACK-R, {L}^3 cubic lattice N={N}, onsite +-1 "
                        f"checkerboard, hopping {HOP} +- {DISORDER} disorder, threshold {THR}, "
                        f"nocc N/2, Gershgorin, M={M} (the density matrix keeps ~2440-2530 "
                        f"entries per row; the configured M=256 overflows as in the reference); "
                        f"a step is one SP2 iteration in the steady regime (after {SETTLE} "
                        f"untimed iterations): traces, branch, fused X^2 upper triangle + "
                        f"update + norm, mirror; mixed = FP32 storage and products, FP64 sums, "
                        f"traces, norm and branch; value = executed X^2 GFLOP/s (2 x partial "
                        f"products formed) over whole-iteration time",
            "N": N, "M": M, "threshold": THR, "settle_iterations": SETTLE,
        },
        "fp64_gflops": round(gflops_dbl, 2),
        "fp64_ms_per_step": round(1000.0 * t_dbl / args.steps, 3),
        "mixed_over_fp64_per_step": round(t_dbl / t_mix, 3),
        "steady_idempotency": {"mixed": float(h_mix[-1, 2]), "fp64": float(h_dbl[-1, 2])},
        "steady_trace_error": {"mixed": float(abs(h_mix[-1, 0] - N / 2)),
                               "fp64": float(abs(h_dbl[-1, 0] - N / 2))},
        "full_solves": solves,
        "roofline": roof,
        "roofline_fp64": roof64,
        "row_engine": "unordered setting (window engine for these rows)",
        "copy_calibration": calib,
        "library_sha256_16": lib_digest(),
        "kernel_source_sha256_16": kernel_digest(),
        "x2_config1": None if args.only_headline else x2_config1(hip, G, cpu=not args.no_cpu),
        "x2_config5": None if args.only_headline else x2_config5(hip, G),
        "ellblock_sp2": None if args.only_headline else ellblock_sp2(hip, G),
        "batched_sp2": None if args.only_headline else batched_sp2(hip, G),
        "cpu_baseline": None if (args.no_cpu or args.only_headline) else cpu_baseline(X64),
    }
    X64.free()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
