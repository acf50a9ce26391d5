"""Benchmark of the H2-matrix kernel action on one MI355X.

Workload (the configuration the headline metric is quoted on):
  n = 1,000,000 points uniform in [0,1]^3 (numpy default_rng(0)), Matern-3/2
  kernel with lengthscale 0.5, tensor Chebyshev degree 4, eta = 0.8,
  32 right-hand sides x with i.i.d. N(0,1) entries.

One step = one solve of (K + 100 I) a = x for the 32 resident columns to a
1e-8 relative FP64 residual: GCR whose matvecs use FP32 kernel entries with
BF16 three-term matrix-core contractions, the residual replaced by the FP64
action's each time it has dropped by 1e-3, and vectors, inner products and
updates in FP64 (the precision form with FP64 refinement).
value = n * nrhs * (iterations FMCA's plain FP64 CG needs for the same
tolerance) / seconds per refined solve: the FP64-equivalent matvec
throughput of the refined solver.  Beside it, each once: FMCA's FP64 CG on
the same system, GCR with FP64 matvecs, CG with reliable updates on the same
BF16 matvecs; single matvecs of the three precision forms (1 and 32 rhs, per
phase), the host-buffer boundary, the differentiation side results (n = 2e5
2-D and n = 8e6 3-D: forward, backward, ratio, memory, and a refined solve
at n = 8e6) and the FP64 CPU oracle on the host cores.

Prints ONE JSON line.  Usage:
  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
                  [--no-side] [--workload NAME] [--calibrate-copy-mib M]
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

METRIC = "H2 matvec Mpoint*RHS/s, n=1e6 3D, FP32+refine vs FP64 at 1e-8 rel err"
PEAK_FP32_TFLOPS = 157.3   # MI355X dense FP32 (vector and matrix)
# FP64 residual replacement of the GCR solves on BF16-split matvecs: once the
# iterated residual has dropped by this factor, then the check at the end
# (the recurrence runs to half of tol first): two FP64 actions per solve, with
# the drift of the second cycle (about 1e-4 of its starting residual) well
# below tol
GCR_REFINE = 3e-5
PEAK_FP64_TFLOPS = 78.6
PEAK_BF16_TFLOPS = 2500.0  # MI355X dense BF16 matrix
PEAK_HBM_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--n", type=int, default=1_000_000)
    p.add_argument("--nrhs", type=int, default=32)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-solve", action="store_true",
                   help="time the single matvecs only and print them (for profiler runs)")
    p.add_argument("--no-side", action="store_true",
                   help="skip the differentiation side results (n = 2e5 2-D and n = 8e6 3-D)")
    p.add_argument("--ridge", type=float, default=100.0)
    p.add_argument("--solve-precision", default="fp32_bf16", choices=["fp32", "fp32_bf16"],
                   help="matvecs of the refined solve: FP32 entries with FP32 (fp32) or "
                        "BF16 three-term (fp32_bf16) matrix-core contractions")
    p.add_argument("--solve-method", default="gcr", choices=["gcr", "cg"],
                   help="method of the refined solve (FP64 residual replacement in both)")
    p.add_argument("--refine-factor", type=float, default=None,
                   help="residual drop between FP64 replacements (default 1e-3 for gcr, "
                        "0.1 for cg)")
    p.add_argument("--workload", default="matvec",
                   choices=["matvec", "small-dense", "clustered-cg", "autograd", "large",
                            "total-degree", "plan", "assembled", "clustered-spectrum", "samplets"],
                   help="matvec: the headline line (default); the others print a side line "
                        "for the other configurations")
    p.add_argument("--spectrum-cases", default="4:0.8,5:0.8,6:0.8,7:0.8,8:0.8,9:0.8",
                   help="clustered-spectrum: degree:eta pairs")
    p.add_argument("--spectrum-steps", type=int, default=8,
                   help="clustered-spectrum: block Krylov steps (block size 16)")
    p.add_argument("--calibrate-copy-mib", type=int, default=0,
                   help="also run one elementwise copy of this size (reads and writes "
                        "M MiB) as the reference kernel for HBM traffic counters")
    p.add_argument("--cpu-seconds", type=float, default=12.0,
                   help="target CPU time of the sampled coupling/near-field loop")
    return p.parse_args()


def timed_steps(H, x, out, precision, steps, warmup, profile):
    import torch
    for _ in range(warmup):
        H.action_device(x, out=out, precision=precision)
    torch.cuda.synchronize()
    if profile:
        H.phase_times(reset=True)
        H.set_profiling(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        H.action_device(x, out=out, precision=precision)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    phases = None
    if profile:
        H.set_profiling(False)
        phases = H.phase_times(reset=True)
    return (t1 - t0) / steps, phases


def mfma_tiles(nrhs):
    """rhs tiles of 16 columns per workgroup pass of the MFMA kernels."""
    return 1 if nrhs <= 16 else (2 if nrhs <= 32 else 4)


def traffic_from_profiles(kernel_name):
    """HBM bytes per launch of `kernel_name` from rocprofv3 --pmc FETCH_SIZE and
    WRITE_SIZE runs, each scaled by what the counter reads for a copy kernel of
    known size (profiles/traffic.json, written from those runs); None if the
    kernel was not profiled."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        k = d.get("kernels", {}).get(kernel_name)
        return None if k is None else k["hbm_bytes_per_launch"]
    except (OSError, ValueError, KeyError):
        return None


def host_threads():
    """Threads for the CPU baseline: every core this process may run on
    (its affinity mask), unless OMP_NUM_THREADS caps it, as it does on hosts
    that give each GPU a fixed share of the cores.  Returns (threads, note)."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = min(affinity, int(omp)) if omp and omp.isdigit() else affinity
    note = (f"{threads} threads: nproc reports {os.cpu_count()}, the process may run on "
            f"{affinity} cores, OMP_NUM_THREADS={omp or 'unset'} (the host's CPU share for "
            f"this GPU)")
    return threads, note


def cpu_baseline(P, x, info, budget_seconds):
    """FMCA's OpenMP CPU H2 action as restated by the FP64 oracle, timed on a
    bounded sample of the column groups and scaled by kernel entries."""
    import oracle
    nthreads, cores_note = host_threads()
    A = oracle.OracleH2(P, P, "MATERN32", 0.5, 4, 0.8)
    total = info["near_entries"] + info["far_entries"]
    # calibration pass, then a pass sized to the time budget
    done, times = A.action_sample(x, 2e8, nthreads)
    rate = done / max(times[1], 1e-6)
    budget = min(total, max(2e8, rate * budget_seconds))
    done, times = A.action_sample(x, budget, nthreads)
    t_full = times[0] + times[2] + times[1] * total / done
    nrhs = x.shape[1]
    return {
        "value": P.shape[1] * nrhs / t_full / 1e6,
        "unit": "Mpoint*RHS/s",
        "cores": nthreads,
        "cores_note": cores_note,
        "kind": "port",
        "sample": (f"oracle (FP64 C++/OpenMP restatement of FMCA's action, same tree/blocks) "
                   f"on the same n={P.shape[1]} points and {nrhs} rhs: forward pass "
                   f"{times[0]:.2f}s and backward pass {times[2]:.2f}s timed in full, the "
                   f"coupling/near-field loop timed on {done / total * 100:.1f}% of the "
                   f"{total:.3e} kernel entries ({times[1]:.1f}s) and scaled by entries; "
                   f"setup {A.setup_seconds:.1f}s not counted"),
        "seconds_per_matvec_extrapolated": t_full,
    }


def _sync_time(fn, reps=1):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        out = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps, out


def side_small_dense(args):
    """n = 1e4 uniform 3-D, EXPONENTIAL l = 0.5, deg 3, eta 0.8, one rhs
    U[-1,1]: H2 error of the FP64 and FP32 actions against dense K x (FP64,
    formed on the GPU with torch), GPU times, and the CPU oracle's full action."""
    import torch
    import fmca_amd
    import oracle
    n = 10000
    rng = np.random.default_rng(0)
    P = np.asfortranarray(rng.random((3, n)))
    xh = np.asfortranarray(rng.uniform(-1, 1, (n, 1)))
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("EXPONENTIAL", 0.5), P, P, 3, 0.8)
    x = torch.from_numpy(xh).cuda()
    Pt = torch.from_numpy(np.ascontiguousarray(P.T)).cuda()
    yd = torch.exp(-torch.cdist(Pt, Pt) / 0.5) @ x
    res = {"workload": "small-dense: n=1e4 uniform 3D, EXPONENTIAL l=0.5, deg 3, eta 0.8, 1 rhs"}
    for name, prec in (("fp64", fmca_amd.FP64), ("fp32", fmca_amd.FP32)):
        H.action_device(x, precision=prec)
        t, y = _sync_time(lambda: H.action_device(x, precision=prec), 10)
        # the same action captured once in a graph and replayed (no launch overhead)
        yg = torch.empty_like(x)
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            H.action_device(x, out=yg, precision=prec)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            H.action_device(x, out=yg, precision=prec)
        g.replay()
        tg, _ = _sync_time(g.replay, 20)
        res[name] = {"ms": t * 1e3, "graph_replay_ms": tg * 1e3,
                     "rel_err_vs_dense": float(torch.linalg.norm(y - yd) / torch.linalg.norm(yd)),
                     "graph_equals_eager": bool(torch.equal(yg, y))}
    A = oracle.OracleH2(P, P, "EXPONENTIAL", 0.5, 3, 0.8)
    nthreads, _ = host_threads()
    t0 = time.perf_counter()
    yo = A.action(xh, nthreads)
    res["cpu_oracle"] = {"ms": (time.perf_counter() - t0) * 1e3, "threads": nthreads,
                         "fp64_gpu_vs_oracle": float(np.linalg.norm(
                             H.action(xh) - yo) / np.linalg.norm(yo))}
    res["info"] = H.info()
    return res


def side_clustered_cg(args):
    """n = 1e6 points from a mixture of 16 isotropic Gaussians (sigma 0.02,
    centres uniform in [0,1]^3), GAUSSIAN l = 0.1, deg 4, eta 0.8, 64 rhs
    y = sin(2 pi x1) cos(2 pi x2) + 0.01 N(0,1): matvec times per phase in the
    three precisions, their differences from FP64, CG on (K + 1e-4 I) a = y
    (FP64, to the first p.(K~ + lambda I)p <= 0: the H2 matrix plus this ridge
    is indefinite) and CG to 1e-8 at a ridge where it is definite
    (--ridge, default 10 here): FP32 and FP32-BF16 matvecs with FP64
    reliable updates against plain FP64 CG."""
    import torch
    import fmca_amd
    from fmca_amd import workloads as W
    n, nrhs = 1_000_000, 64
    ridge = args.ridge if args.ridge != 100.0 else 10.0
    P = W.gaussian_mixture(n, 3, 16, 0.02, 0)
    t0 = time.perf_counter()
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("GAUSSIAN", 0.1), P, P, 4, 0.8)
    plan = time.perf_counter() - t0
    y = torch.from_numpy(W.smooth_rhs(P, nrhs)).cuda()
    out = torch.empty((nrhs, n), dtype=torch.float64, device="cuda").t()
    res = {"workload": "clustered-cg: n=1e6 16-Gaussian mixture (sigma 0.02), GAUSSIAN l=0.1, "
                       "deg 4, eta 0.8, 64 rhs", "plan_s": plan, "info": H.info()}
    ys = {}
    for name, prec in (("fp32", fmca_amd.FP32), ("fp32_bf16", fmca_amd.FP32_BF16),
                       ("fp64", fmca_amd.FP64)):
        t, ph = timed_steps(H, y, out, prec, 2, 1, True)
        ys[name] = out.clone()
        res[name] = {"ms": t * 1e3, "mpts": n * nrhs / t / 1e6,
                     "phases_ms": {k: v[0] / max(v[1], 1) for k, v in ph.items()}}
    for name in ("fp32", "fp32_bf16"):
        res[name]["vs_fp64_rel_err"] = float(torch.linalg.norm(ys[name] - ys["fp64"])
                                             / torch.linalg.norm(ys["fp64"]))
    del ys, out
    try:
        _, info = H.cg(y, x0=torch.zeros_like(y), ridge=1e-4, tol=1e-8,
                       precision=fmca_amd.FP64, max_iter=500)
    except (RuntimeError, ValueError) as e:
        info = {"error": str(e)}
    res["cg_fp64_ridge_1e-4"] = info
    sol = {}
    for name, prec in (("fp32_bf16", fmca_amd.FP32_BF16), ("fp64", fmca_amd.FP64)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        try:
            a, info = H.cg(y, x0=torch.zeros_like(y), ridge=ridge, tol=1e-8, precision=prec,
                           max_iter=3000)
            sol[name] = a
        except (RuntimeError, ValueError) as e:
            info = {"error": str(e)}
        torch.cuda.synchronize()
        info["wall_s"] = time.perf_counter() - t0
        res[f"cg_{name}_ridge_{ridge:g}"] = info
    if len(sol) == 2:
        res["solution_rel_diff"] = float(torch.linalg.norm(sol["fp32_bf16"] - sol["fp64"])
                                         / torch.linalg.norm(sol["fp64"]))
    return res


def side_clustered_spectrum(args):
    """Is K~ + 1e-4 I positive definite for the clustered configuration
    (n = 1e6, 16-Gaussian mixture sigma 0.02, GAUSSIAN l = 0.1), and at which
    interpolation degree?  K~ is the H2 matrix, K the exact kernel matrix
    (an eta = 0 actor: every block dense).  For each (degree, eta):
    block Krylov space of E = K~ - K (block 16, --spectrum-steps blocks,
    full reorthogonalisation, FP64 actions), then Rayleigh-Ritz of E and of
    K~ on that space.
      * min Ritz value of K~ + ridge I over the space: a Rayleigh quotient, so
        a negative one proves K~ + ridge I indefinite (the vector is explicit);
      * min Ritz value of E: estimates lambda_min(E); by Weyl,
        lambda_min(K~ + ridge I) >= ridge + lambda_min(E) since K is PSD;
      * ||K~||_2 from 12 block power steps on K~, giving cond estimates.
    One JSON line per case, printed as it finishes."""
    import torch
    import fmca_amd
    from fmca_amd import workloads as W
    n, b, steps, ridge = args.n, 16, args.spectrum_steps, 1e-4
    cases = [(int(c.split(":")[0]), float(c.split(":")[1]))
             for c in args.spectrum_cases.split(",")]
    P = W.gaussian_mixture(n, 3, 16, 0.02, 0)
    K = fmca_amd.CovarianceKernel("GAUSSIAN", 0.1)
    t0 = time.perf_counter()
    Hd = fmca_amd.H2kernelActor(K, P, P, 4, 0.0)  # the exact K: no admissible blocks
    dinfo = Hd.info()
    assert dinfo["far_blocks"] == 0
    print(json.dumps({"dense_actor": {"plan_s": time.perf_counter() - t0,
                                      "near_entries": dinfo["near_entries"]}}), flush=True)
    gen = torch.Generator(device="cuda")
    out = []
    for deg, eta in cases:
        t0 = time.perf_counter()
        H = fmca_amd.H2kernelActor(K, P, P, deg, eta)
        info = H.info()
        gen.manual_seed(1234)
        Q = torch.linalg.qr(torch.randn(n, b, dtype=torch.float64, device="cuda",
                                        generator=gen))[0]
        Qs, KtQ, KQ = [], [], []
        t_h2 = t_dense = 0.0
        for _ in range(steps):
            torch.cuda.synchronize()
            ta = time.perf_counter()
            A1 = H.action_device(Q)
            torch.cuda.synchronize()
            tb = time.perf_counter()
            A2 = Hd.action_device(Q)
            torch.cuda.synchronize()
            t_h2 += tb - ta
            t_dense += time.perf_counter() - tb
            Qs.append(Q)
            KtQ.append(A1)
            KQ.append(A2)
            Wm = A1 - A2
            Qa = torch.cat(Qs, dim=1)
            for _ in range(2):
                Wm = Wm - Qa @ (Qa.t() @ Wm)
            Q = torch.linalg.qr(Wm)[0]
        Qa = torch.cat(Qs, dim=1)
        Kt = torch.cat(KtQ, dim=1)
        Ke = torch.cat(KQ, dim=1)
        Tk = Qa.t() @ Kt
        Te = Tk - Qa.t() @ Ke
        Tk, Te = 0.5 * (Tk + Tk.t()), 0.5 * (Te + Te.t())
        th_k, S_k = torch.linalg.eigh(Tk)
        th_e, S_e = torch.linalg.eigh(Te)
        # residual of the lowest Ritz pair of E, and the Rayleigh quotients of
        # K~ + ridge I and of K at that Ritz vector
        s = S_e[:, :1]
        u = Qa @ s
        r_e = (Kt @ s - Ke @ s) - th_e[0] * u
        rq_k = float((u.t() @ (Kt @ s)).item()) + ridge
        rq_exact = float((u.t() @ (Ke @ s)).item())
        res_e = float(torch.linalg.norm(r_e).item())
        del Qs, KtQ, KQ, Qa, Kt, Ke, u, r_e
        # ||K~||_2 by block power steps
        gen.manual_seed(99)
        X = torch.linalg.qr(torch.randn(n, b, dtype=torch.float64, device="cuda",
                                        generator=gen))[0]
        for _ in range(12):
            X = torch.linalg.qr(H.action_device(X))[0]
        lam_max = float(torch.linalg.eigvalsh(X.t() @ H.action_device(X))[-1].item())
        res = {"degree": deg, "eta": eta, "q": info["q"], "far_blocks": info["far_blocks"],
               "near_entries": info["near_entries"], "far_entries": info["far_entries"],
               "krylov_dim": b * steps,
               "min_ritz_Kh_plus_ridge": float(th_k[0].item()) + ridge,
               "indefinite_proven": bool(float(th_k[0].item()) + ridge < 0),
               "min_ritz_E": float(th_e[0].item()), "max_ritz_E": float(th_e[-1].item()),
               "ritz_residual_E": res_e,
               "rayleigh_Kh_plus_ridge_at_min_E_vector": rq_k,
               "rayleigh_K_at_min_E_vector": rq_exact,
               "weyl_lower_bound_estimate": ridge + float(th_e[0].item()),
               "lambda_max_Kh": lam_max,
               "cond_estimate": lam_max / max(float(th_k[0].item()) + ridge, ridge),
               "h2_action_s": t_h2 / steps, "dense_action_s": t_dense / steps,
               "case_s": time.perf_counter() - t0}
        print(json.dumps({"spectrum": res}), flush=True)
        out.append(res)
        del H
        torch.cuda.empty_cache()
    return {"workload": f"clustered-spectrum: n={n} 16-Gaussian mixture (sigma 0.02), "
                        f"GAUSSIAN l=0.1, ridge {ridge:g}", "cases": out}


def side_samplets(args):
    """Samplet compression (FMCA's tests/test_H2SampletCompressor.cpp setting:
    n = 1e5 uniform 2-D, EXPONENTIAL l = 1, dtilde 2..5, multipole degree
    2 (dtilde - 1)) with CompareCluster at eta 0.8, and the same at dtilde 3
    in 3-D: tree build, compression phases, triplets, a-posteriori triplets
    at 1e-3 and the compression error of FMCA's test (ten columns of K)."""
    import fmca_amd
    import oracle
    import scipy.sparse as sp
    n = 100_000
    out = []
    for dim, dtilde in ((2, 2), (2, 3), (2, 4), (2, 5), (3, 3)):
        P = np.asfortranarray(np.random.default_rng(11).random((dim, n)))
        t0 = time.perf_counter()
        T = fmca_amd.SampletTree(P, dtilde)
        t_tree = time.perf_counter() - t0
        t0 = time.perf_counter()
        comp = fmca_amd.SampletKernelCompressor(T, fmca_amd.CovarianceKernel("EXPONENTIAL", 1.0),
                                                P, 0.8, 0.0)
        t_comp = time.perf_counter() - t0
        st = comp.statistics()
        r, c, v = comp.triplets()
        ra, ca, va = comp.aposteriori_triplets_fast(1e-3)
        perm = T.indices()
        idx = np.random.default_rng(12).choice(n, 10, replace=False)
        X = np.zeros((n, 10))
        X[idx, np.arange(10)] = 1.0
        Y1 = oracle.dense_matrix(P[:, perm], P[:, perm[idx]], "EXPONENTIAL", 1.0).T

        def err_of(rr, cc, vv):
            S = sp.csr_matrix((vv, (rr, cc)), shape=(n, n))
            S = S + sp.triu(S, 1).T
            Y2 = fmca_amd.inverseSampletTransform(T, S @ fmca_amd.sampletTransform(T, X))
            return float(np.linalg.norm(Y2 - Y1) / np.linalg.norm(Y1))
        res = {"dim": dim, "dtilde": dtilde, "tree_s": t_tree, "compress_s": t_comp,
               "nnz_per_row": len(v) / n, "nnz_per_row_aposteriori_1e-3": len(va) / n,
               "compression_error": err_of(r, c, v),
               "compression_error_aposteriori_1e-3": err_of(ra, ca, va),
               **{k: st[k] for k in ("pattern_entries", "dense_jobs", "far_jobs", "job_elements",
                                     "plan_s", "upload_s", "gpu_s", "download_s", "combine_s",
                                     "triplets_s")}}
        print(json.dumps({"samplets": res}), flush=True)
        out.append(res)
        del comp, r, c, v
    return {"workload": "samplets: n=1e5 uniform, EXPONENTIAL l=1, eta 0.8, CompareCluster",
            "cases": out}


def side_total_degree(args):
    """n = 1e6 uniform 3-D, MATERN32 l = 0.5, deg 4, eta 0.8 with
    TotalDegreeInterpolator (q = 35, leaves of 35-69 points), 32 rhs N(0,1):
    plan time, FP32 / FP64 action times per phase, 1-rhs times, and the H2
    error on 200 sampled rows against exact K x, beside the tensor-product
    form (q = 125) of the same degree."""
    import torch
    import fmca_amd
    from fmca_amd import workloads as W
    n, nrhs = 1_000_000, 32
    P = W.uniform(n, 3, 0)
    x = torch.from_numpy(W.normal_rhs(n, nrhs, 1)).cuda()
    out = torch.empty((nrhs, n), dtype=torch.float64, device="cuda").t()
    rows = np.sort(np.random.default_rng(2).choice(n, 200, replace=False))
    Pt = torch.from_numpy(np.ascontiguousarray(P.T)).cuda()
    r = torch.cdist(Pt[torch.from_numpy(rows).cuda()], Pt) * (np.sqrt(3.0) / 0.5)
    exact = ((1 + r) * torch.exp(-r)) @ x
    del r
    res = {"workload": "total-degree: n=1e6 uniform 3D, MATERN32 l=0.5, deg 4, eta 0.8, "
                       "TotalDegreeInterpolator (q=35), 32 rhs"}
    for interp in ("total_degree", "tensor_product"):
        t0 = time.perf_counter()
        H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("MATERN32", 0.5), P, P, 4, 0.8,
                                   interpolator=interp)
        entry = {"plan_s": time.perf_counter() - t0, "info": H.info()}
        ys = {}
        for name, prec in (("fp32", fmca_amd.FP32), ("fp64", fmca_amd.FP64)):
            t, ph = timed_steps(H, x, out, prec, 3, 1, True)
            ys[name] = out.clone()
            entry[name] = {"ms": t * 1e3, "mpts": n * nrhs / t / 1e6,
                           "phases_ms": {k: v[0] / max(v[1], 1) for k, v in ph.items()}}
        entry["fp32_vs_fp64_rel_err"] = float(torch.linalg.norm(ys["fp32"] - ys["fp64"])
                                              / torch.linalg.norm(ys["fp64"]))
        sel = ys["fp64"][torch.from_numpy(rows).cuda()]
        entry["h2_rel_err_200_rows"] = float(torch.linalg.norm(sel - exact)
                                             / torch.linalg.norm(exact))
        x1 = x[:, :1].contiguous()
        for name, prec in (("fp32", fmca_amd.FP32), ("fp64", fmca_amd.FP64)):
            H.action_device(x1, precision=prec)
            t, _ = _sync_time(lambda: H.action_device(x1, precision=prec), 3)
            entry[name]["rhs1_ms"] = t * 1e3
        res[interp] = entry
        del ys, H
    return res


def side_autograd(args):
    """n = 2e5 uniform 2-D, MATERN32 l = 0.3, deg 5, eta 0.8, 8 rhs N(0,1):
    L = 0.5 ||K(l) alpha - y||^2, forward and backward (d alpha, d l, d points)
    times of the differentiable H2 matvec, plain FP64 form and FP32 form."""
    import torch
    from fmca_amd import FP32, FP64
    from fmca_amd.autograd import H2Spec, h2_matvec
    n, nrhs = 200_000, 8
    rng = np.random.default_rng(0)
    pts = torch.from_numpy(rng.random((2, n))).cuda().requires_grad_(True)
    alpha = torch.from_numpy(rng.standard_normal((n, nrhs))).cuda().requires_grad_(True)
    yt = torch.from_numpy(rng.standard_normal((n, nrhs))).cuda()
    l = torch.tensor(0.3, dtype=torch.float64, requires_grad=True)
    res = {"workload": "autograd: n=2e5 uniform 2D, MATERN32 l=0.3, deg 5, eta 0.8, 8 rhs"}
    for name, prec in (("fp64", FP64), ("fp32", FP32)):
        spec = H2Spec("MATERN32", deg=5, eta=0.8, precision=prec)
        h2_matvec(alpha, l, pts, spec)  # plan
        tf, tb = [], []
        for _ in range(max(3, args.steps)):
            for g in (alpha, l, pts):
                g.grad = None
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            loss = 0.5 * (h2_matvec(alpha, l, pts, spec) - yt).pow(2).sum()
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            loss.backward()
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            tf.append(t1 - t0)
            tb.append(t2 - t1)
        f, b = min(tf), min(tb)
        res[name] = {"forward_ms": f * 1e3, "backward_ms": b * 1e3, "backward_over_forward": b / f}
    return res


def side_assembled(args):
    """The assembled H2 matrix (computeH2Matrix + operator*) against the
    matrix-free action, uniform 3-D, MATERN32 l = 0.5, deg 4 (q = 125), eta
    0.8: FP32 at n = 1e6 (32 and 1 rhs) and FP64 at n = 5e5 (32 rhs; the FP64
    blocks of n = 1e6 would not fit next to the rest).  Per configuration:
    assembly seconds, bytes held, both action times, the streamed bandwidth
    (stored bytes / stored action time) and the relative difference."""
    import torch
    import fmca_amd
    res = {"workload": "assembled: uniform 3D, MATERN32 l=0.5, deg 4 (q=125), eta 0.8"}
    for n, prec, nrhs_list in ((1_000_000, fmca_amd.FP32, (32, 1)),
                               (500_000, fmca_amd.FP64, (32,))):
        P = np.asfortranarray(np.random.default_rng(0).random((3, n)))
        H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("MATERN32", 0.5), P, P, 4, 0.8)
        name = "fp32" if prec == fmca_amd.FP32 else "fp64"
        xs = {}
        free = {}
        for nrhs in nrhs_list:
            x = torch.from_numpy(np.asfortranarray(
                np.random.default_rng(1).standard_normal((n, nrhs)))).cuda()
            x = x.t().contiguous().t() if nrhs > 1 else x
            y = torch.empty(nrhs, n, dtype=torch.float64, device="cuda").t()
            H.action_device(x, out=y, precision=prec)  # sizes the workspaces
            free[nrhs] = _sync_time(lambda: H.action_device(x, out=y, precision=prec),
                                    reps=max(3, args.steps))[0]
            xs[nrhs] = (x, y.clone())
        t0 = time.perf_counter()
        nbytes = H.assemble(prec)
        torch.cuda.synchronize()
        t_asm = time.perf_counter() - t0
        for nrhs in nrhs_list:
            x, yf = xs[nrhs]
            y = torch.empty_like(yf.t()).t()
            H.action_device(x, out=y, precision=prec)
            ts = _sync_time(lambda: H.action_device(x, out=y, precision=prec),
                            reps=max(3, args.steps))[0]
            diff = (torch.linalg.norm(y - yf) / torch.linalg.norm(yf)).item()
            res[f"{name}_n{n}_rhs{nrhs}"] = {
                "matrix_free_ms": free[nrhs] * 1e3, "assembled_ms": ts * 1e3,
                "speedup": free[nrhs] / ts, "assemble_s": t_asm, "bytes": nbytes,
                "streamed_TBps": nbytes / ts / 1e12, "rel_diff": diff,
                "mpts": n * nrhs / ts / 1e6}
        H.release_assembled()
        del H
    return res


def side_plan(args):
    """Planning (cluster tree + block pattern) of the n = 8e6 and n = 1e6
    uniform 3-D workloads, degree 4, eta 0.8, on the host and on the GPU: the
    create call's seconds and the plan report.  The GPU plan is timed twice
    (the first call also loads the planner's kernels)."""
    import torch
    import fmca_amd
    torch.zeros(1, device="cuda")
    res = {"workload": "plan: uniform 3D, deg 4 (q=125), eta 0.8; trees + block pattern"}
    for n in (8_000_000, 1_000_000):
        P = np.asfortranarray(np.random.default_rng(0).random((3, n)))
        K = fmca_amd.CovarianceKernel("EXPONENTIAL", 0.25)
        runs = {}
        for tag, planner in (("device_first", "device"), ("device", "device"), ("host", "host")):
            t0 = time.perf_counter()
            H = fmca_amd.H2kernelActor(K, P, P, 4, 0.8, planner=planner)
            runs[tag] = {"create_s": time.perf_counter() - t0, **H.plan_report()}
            info = H.info()
            del H
        runs["clusters"] = info["row_clusters"]
        runs["blocks"] = info["far_blocks"] + info["near_blocks"]
        res[f"n={n}"] = runs
    return res


def side_large(args):
    """n = 8e6 uniform 3-D, EXPONENTIAL l = 0.25, deg 4, eta 0.8, 16 rhs:
    plan time, FP32 forward, backward (d alpha, d l) with moments recomputed,
    their ratio, and the device-memory high-water mark."""
    import torch
    import fmca_amd
    from fmca_amd.autograd import H2Spec, h2_matvec
    n, nrhs = 8_000_000, 16
    rng = np.random.default_rng(0)
    P = rng.random((3, n))
    fmca_amd.device_memory(reset_peak=True)
    torch.cuda.reset_peak_memory_stats()
    pts = torch.from_numpy(P).cuda()
    alpha = torch.from_numpy(rng.standard_normal((n, nrhs))).cuda().requires_grad_(True)
    l = torch.tensor(0.25, dtype=torch.float64, requires_grad=True)
    K = fmca_amd.CovarianceKernel("EXPONENTIAL", 0.25)
    Pf = np.asfortranarray(P)
    plan = {}
    for planner in ("host", "device"):
        t0 = time.perf_counter()
        Hp = fmca_amd.H2kernelActor(K, Pf, Pf, 4, 0.8, planner=planner)
        plan[planner] = {"create_s": time.perf_counter() - t0, **Hp.plan_report()}
        del Hp
    spec = H2Spec("EXPONENTIAL", deg=4, eta=0.8, precision=fmca_amd.FP32, planner="device")
    t0 = time.perf_counter()
    y = h2_matvec(alpha, l, pts, spec)
    torch.cuda.synchronize()
    first = time.perf_counter() - t0
    res = {"workload": "large: n=8e6 uniform 3D, EXPONENTIAL l=0.25, deg 4, eta 0.8, 16 rhs, "
                       "FP32 form for forward and backward; FP32-BF16 + FP64-refined solve",
           "first_forward_incl_plan_s": first, "plan": plan,
           "info": spec._actor.info()}
    tf, tb = [], []
    for _ in range(max(2, args.steps // 2)):
        alpha.grad = None
        l.grad = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        y = h2_matvec(alpha, l, pts, spec)
        loss = 0.5 * y.pow(2).sum()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        tf.append(t1 - t0)
        tb.append(time.perf_counter() - t1)
    f, b = min(tf), min(tb)
    res.update({"forward_ms": f * 1e3, "mpts": n * nrhs / f / 1e6, "backward_ms": b * 1e3,
                "backward_over_forward": b / f,
                "device_memory": {"library_peak_bytes": fmca_amd.device_memory()[1],
                                  "torch_peak_bytes": int(torch.cuda.max_memory_allocated())}})
    # FP32 bulk + FP64 refinement: solve (K + 1e-4 n I) a = alpha to a 1e-8
    # relative FP64 residual, GCR on the BF16-split FP32 matvecs with FP64
    # residual replacement, on the same actor (its plan and resident points)
    H = spec._actor
    b_rhs = alpha.detach()
    ridge = 1e-4 * n
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    a, info = H.cg(b_rhs, x0=torch.zeros_like(b_rhs), ridge=ridge, tol=1e-8,
                   precision=fmca_amd.FP32_BF16, method="gcr", refine_factor=GCR_REFINE,
                   max_iter=2000)
    torch.cuda.synchronize()
    ts = time.perf_counter() - t0
    r = b_rhs - H.action_device(a) - ridge * a
    resid = (torch.linalg.norm(r, dim=0) / torch.linalg.norm(b_rhs, dim=0)).max().item()
    del r
    res["refined_solve"] = {
        "system": f"(K + {ridge:g} I) a = alpha, 16 rhs, tol 1e-8, from a = 0",
        "method": "gcr", "precision": "fp32_bf16", "refine_factor": GCR_REFINE, "seconds": ts,
        "iterations": info["iterations"], "matvecs_fp64": info["matvecs_fp64"],
        "rel_residual_fp64_rechecked": resid, "converged": info["converged"],
        "device_memory": {"library_peak_bytes": fmca_amd.device_memory()[1],
                          "torch_peak_bytes": int(torch.cuda.max_memory_allocated())}}
    return res


def refined_solves(H, x, ridge, steps, warmup, precision, method, factor):
    """The headline step: one solve of (K + ridge I) a = x for all columns
    to a 1e-8 relative FP64 residual (fmca_h2_cg_device), FP32-class matvecs
    driving the iteration and FP64 residual replacement, from a = 0.
    Profiling events (pooled, created before the timed region) stay on
    through the timed solves, so that the leaf-output launch time of the
    roofline is measured on the launches of the timed region itself."""
    import torch
    import fmca_amd
    infos = []

    def one():
        a, info = H.cg(x, x0=torch.zeros_like(x), ridge=ridge, tol=1e-8,
                       precision=precision, refine_factor=factor, method=method,
                       max_iter=2000)
        return a, info
    H.set_profiling(True)
    for _ in range(warmup):
        one()
    torch.cuda.synchronize()
    H.phase_times(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        a, info = one()
        infos.append(info)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / steps
    H.set_profiling(False)
    ph32 = H.phase_times(reset=True, precision=precision)
    H.phase_times(reset=True)
    return t, a, infos, ph32


def once_solve(H, x, ridge, precision, method="cg", factor=0.1):
    """One solve of the same system from a = 0, timed once (FMCA's plain FP64
    CG by default)."""
    import torch
    H.cg(x[:, :1].contiguous(), x0=torch.zeros_like(x[:, :1].contiguous()), ridge=ridge,
         tol=1e-8, precision=precision, method=method, max_iter=3)  # warms the path
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    a, info = H.cg(x, x0=torch.zeros_like(x), ridge=ridge, tol=1e-8, precision=precision,
                   refine_factor=factor, method=method, max_iter=2000)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, a, info


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              f"the action is not sharded across GPUs.", file=sys.stderr)
        sys.exit(2)
    import torch
    import fmca_amd

    side = {"small-dense": side_small_dense, "clustered-cg": side_clustered_cg,
            "autograd": side_autograd, "large": side_large,
            "total-degree": side_total_degree, "plan": side_plan,
            "assembled": side_assembled,
            "clustered-spectrum": side_clustered_spectrum,
            "samplets": side_samplets}.get(args.workload)
    if side is not None:
        print(json.dumps(side(args)), flush=True)
        return

    n, nrhs = args.n, args.nrhs
    rng = np.random.default_rng(0)
    P = np.asfortranarray(rng.random((3, n)))
    K = fmca_amd.CovarianceKernel("MATERN32", 0.5)
    torch.zeros(1, device="cuda")  # HIP runtime up before the planning clock starts
    t0 = time.perf_counter()
    Hh = fmca_amd.H2kernelActor(K, P, P, 4, 0.8, planner="host")
    setup_plan_host = time.perf_counter() - t0
    del Hh
    t0 = time.perf_counter()
    H = fmca_amd.H2kernelActor(K, P, P, 4, 0.8, planner="device")
    setup_plan = time.perf_counter() - t0
    plan_report = H.plan_report()
    t0 = time.perf_counter()
    H.upload()
    torch.cuda.synchronize()
    setup_upload = time.perf_counter() - t0
    info = H.info()
    xh = np.asfortranarray(rng.standard_normal((n, nrhs)))
    x = torch.from_numpy(xh).cuda()  # column-major (F-ordered) device copy
    x = x.t().contiguous().t() if x.stride(0) != 1 else x
    out32 = torch.empty((nrhs, n), dtype=torch.float64, device="cuda").t()
    out64 = torch.empty((nrhs, n), dtype=torch.float64, device="cuda").t()

    # single matvecs of both precisions (per-phase times, FP32 vs FP64 difference)
    mv_steps = 5
    t32, ph32m = timed_steps(H, x, out32, fmca_amd.FP32, mv_steps, 1, True)
    t64, ph64m = timed_steps(H, x, out64, fmca_amd.FP64, 3, 1, True)
    err = float(torch.linalg.norm(out32 - out64) / torch.linalg.norm(out64))
    tbf, phbfm = timed_steps(H, x, out32, fmca_amd.FP32_BF16, mv_steps, 1, True)
    err_bf = float(torch.linalg.norm(out32 - out64) / torch.linalg.norm(out64))
    del out32, out64
    x1 = x[:, :1].contiguous()
    o1 = torch.empty((n, 1), dtype=torch.float64, device="cuda")
    t32_1, _ = timed_steps(H, x1, o1, fmca_amd.FP32, mv_steps, 1, False)
    solve_prec = fmca_amd.FP32_BF16 if args.solve_precision == "fp32_bf16" else fmca_amd.FP32
    t64_1, _ = timed_steps(H, x1, o1, fmca_amd.FP64, 3, 1, False)
    del o1, x1
    # the C-ABI host-buffer boundary (fmca_h2_action): H2D + action + D2H
    H.action(xh, precision=fmca_amd.FP32)
    t0 = time.perf_counter()
    H.action(xh, precision=fmca_amd.FP32)
    t_host = time.perf_counter() - t0

    if args.no_solve:
        if args.calibrate_copy_mib > 0:
            nel = args.calibrate_copy_mib * (1 << 20) // 8
            src = torch.ones(nel, dtype=torch.float64, device="cuda")
            dst = torch.empty_like(src)
            torch.neg(src, out=dst)
            torch.cuda.synchronize()
            del src, dst
        print(json.dumps({"workload": "single matvecs (profiler run)", "n": n, "nrhs": nrhs,
                          "fp32_ms": t32 * 1e3, "fp32_bf16_ms": tbf * 1e3, "fp64_ms": t64 * 1e3,
                          "fp32_rhs1_ms": t32_1 * 1e3, "fp64_rhs1_ms": t64_1 * 1e3}), flush=True)
        return
    # the headline: refined solves (timed steps); then, once each, FMCA's plain
    # FP64 CG (whose iteration count the value credits), the same method in
    # FP64, and the other method with the same matvecs
    factor = args.refine_factor or (GCR_REFINE if args.solve_method == "gcr" else 0.1)
    t_ref, a_ref, infos, ph_solve = refined_solves(H, x, args.ridge, args.steps, args.warmup,
                                                   solve_prec, args.solve_method, factor)
    t_64, a_64, info64 = once_solve(H, x, args.ridge, fmca_amd.FP64)
    sol_diff = float(torch.linalg.norm(a_ref - a_64) / torch.linalg.norm(a_64))
    del a_ref
    other = "cg" if args.solve_method == "gcr" else "gcr"
    t_64m, a_64m, info64m = (once_solve(H, x, args.ridge, fmca_amd.FP64, "gcr", 0.1)
                             if args.solve_method == "gcr" else (t_64, None, info64))
    del a_64m
    t_oth, a_oth, info_oth = once_solve(H, x, args.ridge, solve_prec, other,
                                        0.1 if other == "cg" else 1e-3)
    oth_diff = float(torch.linalg.norm(a_oth - a_64) / torch.linalg.norm(a_64))
    del a_oth, a_64
    iters64 = info64["iterations"]
    value = n * nrhs * iters64 / t_ref / 1e6
    value_fp64 = n * nrhs * iters64 / t_64 / 1e6

    if args.calibrate_copy_mib > 0:
        nel = args.calibrate_copy_mib * (1 << 20) // 8
        src = torch.ones(nel, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.neg(src, out=dst)
        torch.cuda.synchronize()
        del src, dst

    E_near, E_far, q = info["near_entries"], info["far_entries"], int(info["q"])

    # Dominant kernel of the timed step: the FP32 leaf output (near field +
    # V^T l_t), one launch per FP32 matvec of the solves, timed with HIP
    # events on the launch stream over the timed solves.  Algorithmic flops
    # per launch: 2 E_near nrhs (near entries) + 2 q n nrhs (far basis).
    def avg(ph, k):
        return ph[k][0] / max(ph[k][1], 1)
    dom = max(("leaf_output", "coupling"), key=lambda k: avg(ph_solve, k))
    t_dom = avg(ph_solve, dom) / 1e3
    nt = mfma_tiles(nrhs)
    bf = args.solve_precision == "fp32_bf16"
    if dom == "leaf_output":
        flops = 2.0 * E_near * nrhs + 2.0 * q * n * nrhs
        bf16_flops = 3 * 2.0 * E_near * nrhs  # the three BF16 products of the near field
        entries = E_near
        kname = (f"k_leaf_output_bf16<3, 1, {nt}>" if bf
                 else f"k_leaf_output_mfma<float, 3, 1, {nt}, false>")
    else:
        flops = 2.0 * E_far * nrhs
        bf16_flops = 3 * 2.0 * E_far * nrhs
        entries = E_far
        kname = (f"k_coupling_bf16<3, 1, {nt}>" if bf
                 else f"k_coupling_mfma<float, 3, 1, {nt}, false>")
    achieved = flops / t_dom / 1e12
    roofline = {
        # the products are FP32-accurate (FP32 entries; BF16 three-term split or
        # FP32 matrix cores): the roofline is the FP32 dense matrix peak
        "bound": "mfma",
        "kernel": kname,
        "achieved": achieved,
        "peak": PEAK_FP32_TFLOPS,
        "unit": "TFLOP/s",
        "frac": achieved / PEAK_FP32_TFLOPS,
        "bf16_matrix_core_frac": (bf16_flops / t_dom / 1e12 / PEAK_BF16_TFLOPS) if bf else None,
        "traffic": traffic_from_profiles(kname),
        "traffic_source": ("profiles/traffic.json: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE "
                           "runs of this workload's FP32 matvec, scaled by the counters' reading "
                           "for a copy kernel of known size; not measured in this run"),
        "avg_launch_ms": t_dom * 1e3,
        "launches_timed": int(ph_solve[dom][1]),
        "algorithmic_flops_per_launch": flops,
        "kernel_entries_per_launch": entries,
        "kernel_entries_per_s": entries / t_dom,
    }
    res = {
        "metric": METRIC,
        "value": value,
        "unit": "Mpoint*RHS/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_ref * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic: uniform points in [0,1]^3 and N(0,1) rhs, numpy default_rng(0)",
        "config": {
            "workload": (f"solve of (K + {args.ridge:g} I) a = x to a 1e-8 relative FP64 "
                         f"residual, n={n} uniform 3D points, MATERN32 l=0.5, tensor Chebyshev "
                         f"deg 4 (q={q}), eta=0.8, {nrhs} rhs N(0,1); "
                         f"{args.solve_method.upper()} on {args.solve_precision.upper()} "
                         f"matvecs + FP64 residual replacement (factor {factor:g}); value = "
                         f"n * nrhs * (iterations of FMCA's plain FP64 CG to the same "
                         f"tolerance) / seconds per refined solve"),
            "n": n, "nrhs": nrhs, "kernel": "MATERN32", "lengthscale": 0.5, "degree": 4,
            "eta": 0.8, "ridge": args.ridge, "tol": 1e-8,
            "far_blocks": int(info["far_blocks"]),
            "near_blocks": int(info["near_blocks"]), "near_entries": E_near,
            "far_entries": E_far,
        },
        "roofline": roofline,
        "solve": {
            "fp32_refined": {**infos[-1], "method": args.solve_method,
                             "precision": args.solve_precision, "refine_factor": factor,
                             "seconds_mean": t_ref,
                             "iterations_all_steps": [i["iterations"] for i in infos],
                             "rel_residual_max_all_steps": max(i["rel_residual"] for i in infos)},
            "fp64": {**info64, "method": "cg", "wall_s": t_64, "value": value_fp64},
            "fp64_same_method": {**info64m, "method": args.solve_method, "wall_s": t_64m},
            "refined_other_method": {**info_oth, "method": other, "wall_s": t_oth,
                                     "solution_rel_diff": oth_diff},
            "solution_rel_diff": sol_diff,
            "speedup_fp32_refined_vs_fp64": t_64 / t_ref,
            "speedup_fp32_refined_vs_fp64_same_method": t_64m / t_ref,
        },
        "matvec": {
            "fp32": {"ms": t32 * 1e3, "mpts": n * nrhs / t32 / 1e6,
                     "phases_ms": {k: v[0] / max(v[1], 1) for k, v in ph32m.items()}},
            "fp64": {"ms": t64 * 1e3, "mpts": n * nrhs / t64 / 1e6,
                     "phases_ms": {k: v[0] / max(v[1], 1) for k, v in ph64m.items()}},
            "fp32_bf16": {"ms": tbf * 1e3, "mpts": n * nrhs / tbf / 1e6,
                          "phases_ms": {k: v[0] / max(v[1], 1) for k, v in phbfm.items()},
                          "vs_fp64_rel_err": err_bf},
            "fp32_speedup": t64 / t32,
            "fp32_bf16_speedup": t64 / tbf,
            "fp32_vs_fp64_rel_err": err,
            "rhs1": {"fp32_mpts": n / t32_1 / 1e6, "fp64_mpts": n / t64_1 / 1e6,
                     "fp32_ms": t32_1 * 1e3, "fp64_ms": t64_1 * 1e3},
            "host_boundary": {"fp32_ms": t_host * 1e3, "mpts": n * nrhs / t_host / 1e6,
                              "note": "fmca_h2_action on host arrays: H2D copy + action + D2H"},
        },
        "setup_s": {"plan": setup_plan, "plan_host": setup_plan_host, "upload": setup_upload,
                    "plan_report": plan_report},
        "device_memory": {"library_peak_bytes": fmca_amd.device_memory()[1],
                          "torch_peak_bytes": int(torch.cuda.max_memory_allocated())},
    }
    del H, x
    torch.cuda.empty_cache()
    if not args.no_side:
        res["autograd_2d"] = side_autograd(args)
        res["large_8e6"] = side_large(args)
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(P, xh, info, args.cpu_seconds)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
