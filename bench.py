"""Benchmark: time-to-solution for 32 BdG modes at 128^3 on one MI355X.

One step = one full bdg_solve of the 128^3 dipolar cigar (real psi0, nev = 32,
block 32, tol 1e-8) with psi0 and all fields resident on the device, under
three policies: plain FP64 (the reference's algorithm), mixed (preconditioner
in FP32), and mixed + learned (the learned preconditioner, trained on the GPU
beforehand on pairs from a related problem; its training time is reported
separately).  `value` is the mixed + learned time-to-solution; the FP64 and
mixed times, iteration counts and the eigenvalue error against FP64 are
reported beside it.  The roofline of the dominant kernel is measured live
with HIP events on the context stream (bdgx_bench_pass); the CPU baseline is
the FP64 C oracle (one vector at a time, as the reference) on a bounded
sample, scaled by the column applies of the FP64 device solve.

`--blas` times the Gram and block-combine kernels at the solver's shapes
(FP64 MFMA utilisation).  `--probe` only launches the FP64 FFT pass kernels and a copy kernel of known
size a few times (for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE traffic runs,
whose summary in profiles/ supplies `roofline.traffic`).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
# residual tolerance of the extra mixed + learned solve that stops where the
# mixed path's eigenvalue target (relative omega error <= 1e-7) is reached
ACC_TOL = 1e-6


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def rel_err_modes(w, ref, zero):
    nz = ~zero
    e_rel = float(np.max(np.abs(w[nz] - ref[nz]) / np.abs(ref[nz]))) if nz.any() else 0.0
    e_abs0 = float(np.max(np.abs(w[zero] - ref[zero]))) if zero.any() else 0.0
    return e_rel, e_abs0


def max_cluster_angle(w, A, B, skip, rtol=1e-7):
    """Largest principal angle (radians) between span(A[:, J]) and span(B[:, J])
    over the clusters J of equal omega (degenerate modes compared as
    subspaces), without the zero modes and the cluster holding the last mode
    (its partners may lie beyond the computed modes)."""
    from scipy.linalg import subspace_angles

    n, worst, j0 = len(w), 0.0, 0
    while j0 < n:
        j1 = j0 + 1
        while j1 < n and abs(w[j1] - w[j0]) <= rtol * max(1.0, abs(w[j0])):
            j1 += 1
        J = [j for j in range(j0, j1) if not skip[j]]
        if J and j1 < n:
            worst = max(worst, float(np.max(subspace_angles(A[:, J], B[:, J]))))
        j0 = j1
    return worst


def solve_timed(cfg, psi, mu, precision, steps, warmup, theta=None, tol=None):
    from bdg_lobpcg_amd.workloads import make_problem

    times, last = [], None
    for i in range(warmup + steps):
        p = make_problem(cfg, psi, mu, precision=precision, tol=tol)
        p.set_verbosity(1)
        if theta is not None:
            p.set_learned(theta)
        t0 = time.perf_counter()
        rc = p.solve()
        t = time.perf_counter() - t0
        if rc != 0:
            raise RuntimeError(f"bdg_solve failed with {rc}")
        st = p.stats()
        w = p.eigenvalues(cfg["nev"])
        tag = "mixed+learned" if theta is not None else ("mixed" if precision else "fp64")
        if tol is not None:
            tag += f" (tol {tol:g})"
        log(f"{tag} solve {i} ({'warmup' if i < warmup else 'timed'}): {t:.3f} s,"
            f" iters {st['iterations']}, column applies {st['column_applies']}"
            f" (fp32 {st['column_applies_fp32']}), learned iterations {st['learned_iterations']},"
            f" converged {p.converged()}/{cfg['nev']}")
        if i >= warmup:
            times.append(t)
        final = i == warmup + steps - 1
        last = dict(w=w, zero=p.zero_modes(cfg["nev"]), res=p.residuals(cfg["nev"]), stats=st,
                    conv=p.converged(), arrays=p.arrays() if final else None,
                    v=p.modes(cfg["nev"])[1] if final else None)
        p.close()
    return float(np.mean(times)), last


PASSES = ["xfwd", "yfwd", "z", "yinv", "xinv", "plane_fwd", "plane_inv"]
def kernel_name(pname, cfg, variants):
    """Name of the FP64 kernel that runs pass `pname` under the tuned launch
    variants (plane: 0/1 the one-plane kernel with 4N/8N threads, 2/3 the
    persistent kernel with 4N/8N threads; z: 256 or 128 threads at Nz = 128)."""
    nx, nz = cfg["N"][0], cfg["N"][2]
    v = variants.get(pname, 0)
    if pname in ("plane_fwd", "plane_inv"):
        d = "fwd" if pname == "plane_fwd" else "inv"
        k = f"k_plane_{d}_persist" if v >= 2 else f"k_plane_{d}_real"
        return f"{k}<double, {nx}, {(8 if v & 1 else 4) * nx}>"
    if pname == "z":
        t = (256 if nz == 128 else 128) >> v
        return f"k_zpass_fixed<double, {nz}, 8, {t}>"
    return pname


KERNEL_ROLE = {"plane_fwd": "x+y forward, whole z-planes",
               "plane_inv": "y+x inverse + combine",
               "z": "z forward, k-space multiply, z inverse"}


def roofline(cfg):
    """Per-launch time and algorithmic bytes of every FFT pass kind (32
    spectrum batches, 10 launches between HIP events on the context stream)."""
    import bdg_lobpcg_amd as B

    d = B.DeviceContext(cfg["N"], cfg["L"], cfg["cplx"])
    out = {}
    for prec, name in ((B.FP32, "fp32"), (B.FP64, "fp64")):
        out[f"{name}_variants"] = d.autotune(prec, cfg["m"])
        for pid, pname in enumerate(PASSES):
            try:
                ms, by = d.bench_pass(pid, prec, cfg["m"], iters=10)
            except B.BdgError:
                continue
            out[f"{name}_{pname}"] = dict(ms=ms, bytes=by, GBs=by / (ms * 1e-3) / 1e9)
    d.close()
    return out


def probe(cfg):
    """Launch the three FP64 FFT pass kernels (32 spectrum batches) and a
    1 GiB device copy (k_convert<double, double>) three times each."""
    import bdg_lobpcg_amd as B

    d = B.DeviceContext(cfg["N"], cfg["L"], cfg["cplx"])
    d.autotune(B.FP64, cfg["m"])
    info = {}
    for name in ("plane_fwd", "z", "plane_inv"):
        _, by = d.bench_pass(PASSES.index(name), B.FP64, cfg["m"], iters=3)
        info[name] = by
    n = (1 << 30) // 8
    src, dst = B.DeviceBuffer(8 * n), B.DeviceBuffer(8 * n)
    for _ in range(3):
        B.check(B.lib().bdgx_convert(d.h, B.FP64, src.ptr, n, B.FP64, dst.ptr, n, n, 1), "copy")
    d.sync()
    print(json.dumps({"probe": "fp64 plane_fwd, z, plane_inv x3 each (+1 warm-up), copy 1 GiB x3",
                      "algorithmic_bytes": info, "copy_bytes_read": 8 * n,
                      "copy_bytes_written": 8 * n}), flush=True)
    d.close()


GRAM_SHAPES = [(96, 32, False), (96, 96, False), (96, 96, True), (36, 2, False)]
COMBINE_SHAPES = [(96, 64), (96, 32), (36, 34)]
FP64_MFMA_PEAK_TFS = 78.6


def blas(cfg, iters=10):
    """Gram V^T Y and combine V C at the solver's shapes for this grid (S rows,
    basis [X|P|W] of 3m columns), FP64, HIP events on the context stream:
    achieved TF/s against the FP64 MFMA peak and GB/s against HBM."""
    import bdg_lobpcg_amd as B

    d = B.DeviceContext(cfg["N"], cfg["L"], cfg["cplx"])
    variants = d.autotune(B.FP64, cfg["m"])
    S = int(np.prod(cfg["N"]))
    rng = np.random.default_rng(1)
    kmax = 96
    A = np.asfortranarray(rng.standard_normal((S, kmax)))
    Bm = np.asfortranarray(rng.standard_normal((S, kmax)))
    dA, dB = d.upload(A), d.upload(Bm)
    dG = B.DeviceBuffer(8 * kmax * kmax)
    out = []
    for p, q, same in GRAM_SHAPES:
        dY = dA if same else dB

        def run():
            B.check(B.lib().bdgx_gram(d.h, B.FP64, dA.ptr, S, p, dY.ptr, S, q, S, 0, dG.ptr), "gram")

        run()
        G = d.download(dG, (p, q), np.float64)
        ref = A[:, :p].T @ (A if same else Bm)[:, :q]
        err = float(np.abs(G - ref).max() / np.abs(ref).max())
        d.timer_begin()
        for _ in range(iters):
            run()
        ms = d.timer_end() / iters
        by = S * (p + (0 if same else q)) * 8
        fl = 2.0 * S * p * q
        out.append(dict(kernel="gram", shape=f"{p}x{q}" + (" (A^T A)" if same else ""), us=ms * 1e3,
                        GBs=by / ms / 1e6, TFs=fl / ms / 1e9,
                        mfma_frac=fl / ms / 1e9 / FP64_MFMA_PEAK_TFS, rel_err=err))
    for k, m in COMBINE_SHAPES:
        Cm = np.asfortranarray(rng.standard_normal((k, m)))
        dC = d.upload(Cm)
        dY = B.DeviceBuffer(S * m * 8)

        def run():
            B.check(B.lib().bdgx_combine(d.h, B.FP64, dA.ptr, S, k, dC.ptr, dY.ptr, S, m, S),
                    "combine")

        run()
        Y = d.download(dY, (S, m), np.float64)
        rows = slice(0, S, 997)
        ref = A[rows, :k] @ Cm
        err = float(np.abs(Y[rows] - ref).max() / np.abs(ref).max())
        d.timer_begin()
        for _ in range(iters):
            run()
        ms = d.timer_end() / iters
        by = S * (k + m) * 8
        fl = 2.0 * S * k * m
        out.append(dict(kernel="combine", shape=f"{k}->{m}",
                        variant=variants["combine"], rows_per_lane=2 if variants["combine"] in (1, 2) else 4,
                        persistent=variants["combine"] >= 2, us=ms * 1e3,
                        GBs=by / ms / 1e6,
                        TFs=fl / ms / 1e9, mfma_frac=fl / ms / 1e9 / FP64_MFMA_PEAK_TFS,
                        rel_err=err))
    # the update of a block whose other columns are frozen: 36 basis columns
    # read once, the 4 changing ones written through a column map
    import ctypes
    k, cols = 36, [30, 31, 32, 33]
    Cm = np.asfortranarray(rng.standard_normal((k, len(cols))))
    dC = d.upload(Cm)
    dY = B.DeviceBuffer(S * 34 * 8)
    cidx = (ctypes.c_int * len(cols))(*cols)

    def run():
        B.check(B.lib().bdgx_combine_cols(d.h, B.FP64, dA.ptr, S, k, dC.ptr, dY.ptr, S, cidx,
                                          len(cols), S), "combine_cols")

    run()
    Y = d.download(dY, (S, 34), np.float64)
    rows = slice(0, S, 997)
    ref = A[rows, :k] @ Cm
    err = float(np.abs(Y[rows][:, cols] - ref).max() / np.abs(ref).max())
    d.timer_begin()
    for _ in range(iters):
        run()
    ms = d.timer_end() / iters
    by = S * (k + len(cols)) * 8
    fl = 2.0 * S * k * len(cols)
    out.append(dict(kernel="combine_cols", shape=f"{k}->{len(cols)} of 34", us=ms * 1e3,
                    GBs=by / ms / 1e6, TFs=fl / ms / 1e9,
                    mfma_frac=fl / ms / 1e9 / FP64_MFMA_PEAK_TFS, rel_err=err))
    d.close()
    for r in out:
        print(json.dumps(r), flush=True)


def batch(nprob=16, precision=1, nthreads=0):
    """The many-problems workload (workloads.batch_configs): psi0 per problem by
    device ITP, then all solves one after another and again concurrently with
    bdg_solve_batch (one worker thread and HIP stream per problem).  Prints one
    JSON line with both wall times and checks that the concurrent eigenvalues
    equal the sequential ones bit for bit."""
    import bdg_lobpcg_amd as B
    from bdg_lobpcg_amd.workloads import batch_configs, ground_state, make_problem

    cfgs = batch_configs(nprob)
    states = []
    t0 = time.perf_counter()
    for c in cfgs:
        psi, mu, _ = ground_state(c)
        states.append((psi, mu))
    t_itp = time.perf_counter() - t0
    log(f"batch: {nprob} ground states in {t_itp:.1f} s")

    def build():
        return [make_problem(c, psi, mu, precision=precision) for c, (psi, mu) in zip(cfgs, states)]

    probs = build()
    t0 = time.perf_counter()
    rc_seq = [p.solve() for p in probs]
    t_seq = time.perf_counter() - t0
    w_seq = [p.eigenvalues(c["nev"]) for p, c in zip(probs, cfgs)]
    its = [p.stats()["iterations"] for p in probs]
    conv = [p.converged() for p in probs]
    for p in probs:
        p.close()
    probs = build()
    t0 = time.perf_counter()
    rc_bat = B.Problem.solve_batch(probs, nthreads)
    t_bat = time.perf_counter() - t0
    if any(rc_seq) or any(rc_bat):
        log(f"batch: solve codes sequential {rc_seq}, concurrent {rc_bat}: "
            + "; ".join(sorted({B.error_string(r) for r in rc_seq + rc_bat if r})))
        sys.exit(1)
    same = all(np.array_equal(p.eigenvalues(c["nev"]), w) for p, c, w in zip(probs, cfgs, w_seq))
    for p in probs:
        p.close()
    log(f"batch: sequential {t_seq:.2f} s, concurrent {t_bat:.2f} s ({t_seq / t_bat:.2f}x), "
        f"identical eigenvalues {same}")
    print(json.dumps({
        "workload": f"{nprob} problems, grids {{48^3, 64^3, 80^3, 96^3, 64^2x128}}, nev 16, "
                    f"{'mixed' if precision else 'fp64'}",
        "problems": [dict(grid=list(c["N"]), g=round(c["g"], 1), eps_dd=round(c["edd"], 3),
                          aspect=round(c["omega"][2], 3), iterations=i, converged=cv, rc=r)
                     for c, i, cv, r in zip(cfgs, its, conv, rc_seq)],
        "psi0_seconds": t_itp, "sequential_seconds": t_seq, "concurrent_seconds": t_bat,
        "speedup": t_seq / t_bat, "concurrent_rc": rc_bat, "identical_eigenvalues": same,
    }), flush=True)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE, each scaled by the same counter on a
    1 GiB copy), or None if the summary does not list it."""
    path = os.path.join(HERE, "profiles", "traffic_fft_passes_128c.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("| `") and kernel in line:
                    return float(line.split("|")[-2])
    except OSError:
        pass
    return None


def cpu_baseline(cfg, psi, arrays, column_applies):
    """FP64 C oracle: one column of (M then K) and (T_K then T_M) timed on the
    host cores, scaled by the FP64 device solve's column applies."""
    from oracle import oracle as O

    c = O.Ctx(cfg["N"], cfg["L"], cfg["cplx"])
    c.set_fields(wf=psi)
    c.set_dipolar(cfg["g"] * cfg["edd"], (0.0, 0.0, 1.0), cfg["R"])
    c.set_fields(ltK=arrays["ltK"], ltM=arrays["ltM"], pK=arrays["pK"], pM=arrays["pM"],
                 mu=arrays["mu"])
    rng = np.random.default_rng(0)
    ncol = 2
    V = rng.standard_normal((c.size, ncol))
    t0 = time.perf_counter()
    for j in range(ncol):
        c.apply("K", c.apply("M", V[:, j]))
        c.apply("precondM", c.apply("precondK", V[:, j]))
    t = (time.perf_counter() - t0) / ncol
    return dict(value=t * column_applies, unit="s", cores=O.num_threads(), kind="port",
                sample=f"{ncol} columns of (M then K) + (precondK then precondM) at "
                       f"{'x'.join(map(str, cfg['N']))} FP64 ({t:.3f} s/column), scaled by "
                       f"{column_applies} column applies of the FP64 device solve; "
                       f"operator applies only (Gram/Rayleigh-Ritz excluded)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="cfg1_128c_d")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--probe", action="store_true")
    ap.add_argument("--no-graphs", action="store_true",
                    help="solve with eager launches instead of graph replays of the iteration chain")
    ap.add_argument("--batch", type=int, default=0, metavar="N",
                    help="solve the N-problem batch workload sequentially and concurrently")
    ap.add_argument("--blas", action="store_true",
                    help="time the Gram and combine kernels at the solver's shapes and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: the solver runs on a single GPU "
              "(multi-GPU sharding is a design note only); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    import bdg_lobpcg_amd as B  # noqa: F401  (fails loudly without the native library)
    from bdg_lobpcg_amd import workloads

    workloads.GRAPHS = not args.no_graphs
    from bdg_lobpcg_amd.workloads import CONFIGS, ground_state, train_learned

    cfg = CONFIGS[args.config]
    if args.probe:
        probe(cfg)
        return
    if args.blas:
        blas(cfg)
        return
    if args.batch:
        batch(args.batch)
        return
    log(f"config {args.config}: N={cfg['N']} nev={cfg['nev']} ITP {cfg['itp_steps']} steps")
    psi, mu, t_itp = ground_state(cfg)
    log(f"psi0 by device ITP: {t_itp:.2f} s, mu = {mu:.10f}")
    t64, r64 = solve_timed(cfg, psi, mu, 0, args.steps, args.warmup)
    tmx, rmx = solve_timed(cfg, psi, mu, 1, args.steps, args.warmup)
    theta, t_train, tinfo = train_learned(cfg)
    log(f"learned preconditioner trained in {t_train:.1f} s: {tinfo}")
    tlr, rlr = solve_timed(cfg, psi, mu, 1, args.steps, args.warmup, theta=theta)
    # the same policy run only to the residual that the eigenvalue accuracy
    # target of the mixed path needs (relative omega error <= 1e-7 against
    # FP64; the error is quadratic in the residual), measured, not assumed
    tacc, racc = solve_timed(cfg, psi, mu, 1, args.steps, args.warmup, theta=theta,
                             tol=ACC_TOL)
    zero = r64["zero"] | rlr["zero"] | rmx["zero"]
    e_rel, e_abs0 = rel_err_modes(rlr["w"], r64["w"], zero)
    e_acc, _ = rel_err_modes(racc["w"], r64["w"], zero | racc["zero"])
    # eigenvectors: principal angles against the FP64 modes (v half)
    ang = {name: max_cluster_angle(r64["w"], r["v"], r64["v"], zero | r["zero"])
           for name, r in (("mixed", rmx), ("mixed_learned", rlr), ("mixed_learned_at_omega_target", racc))}
    log("max subspace angle vs FP64 modes: " + ", ".join(f"{k} {v:.2e}" for k, v in ang.items()))
    e_rel_mx, _ = rel_err_modes(rmx["w"], r64["w"], zero)
    log(f"FP64 {t64:.3f} s, mixed {tmx:.3f} s ({t64 / tmx:.2f}x), mixed+learned {tlr:.3f} s "
        f"({t64 / tlr:.2f}x); max rel err vs FP64: mixed {e_rel_mx:.2e}, mixed+learned {e_rel:.2e}")
    log(f"mixed+learned to residual {ACC_TOL:g}: {tacc:.3f} s ({t64 / tacc:.2f}x), "
        f"iters {racc['stats']['iterations']}, max rel omega err vs FP64 {e_acc:.2e}")
    rf = roofline(cfg)
    variants = {k: rf.pop(k) for k in [k for k in rf if k.endswith("_variants")]}
    log("passes: " + ", ".join(f"{k} {v['GBs']:.0f} GB/s" for k, v in rf.items()))
    log(f"autotuned launch variants: {variants}")
    # dominant kernel: the FP64 pass with the longest launch (each operator
    # apply runs every pass once per channel, so it takes the largest share)
    cand = [k for k in ("fp64_plane_fwd", "fp64_z", "fp64_plane_inv") if k in rf] or \
        [k for k in rf if k.startswith("fp64_")]
    dom_key = max(cand, key=lambda k: rf[k]["ms"])
    dom = rf[dom_key]
    dom_pass = dom_key.split("_", 1)[1]
    dom_kernel = kernel_name(dom_pass, cfg, variants.get("fp64_variants", {}))
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(cfg, psi, r64["arrays"], r64["stats"]["column_applies"])
        log(f"cpu baseline {cpu['value']:.1f} s ({cpu['cores']} threads)")
    line = {
        "metric": "time-to-solution, 32 BdG modes at 128^3: mixed+learned vs plain FP64 on GPU",
        "value": tlr,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": tlr * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic (psi0 from device imaginary-time evolution of a Gaussian)",
        "config": {"workload": f"{args.config}: {'x'.join(map(str, cfg['N']))} dipolar cigar, "
                               f"real psi0, nev {cfg['nev']}, block {cfg['m']}, tol {cfg['tol']}, "
                               f"g {cfg['g']}, eps_dd {cfg['edd']}", "grid": list(cfg["N"]),
                   "nev": cfg["nev"], "parallelism": "single GPU"},
        "fp64_seconds": t64,
        "mixed_seconds": tmx,
        "mixed_learned_seconds": tlr,
        "speedup_vs_fp64": t64 / tlr,
        "speedup_mixed_only": t64 / tmx,
        "mixed_learned_vs_fp64_max_rel_err": e_rel,
        "mixed_vs_fp64_max_rel_err": e_rel_mx,
        "max_subspace_angle_vs_fp64": ang,
        "goldstone_abs_diff": e_abs0,
        "iterations": {"fp64": r64["stats"]["iterations"], "mixed": rmx["stats"]["iterations"],
                       "mixed_learned": rlr["stats"]["iterations"]},
        "column_applies": {"fp64": r64["stats"]["column_applies"],
                           "mixed": rmx["stats"]["column_applies"],
                           "mixed_learned": rlr["stats"]["column_applies"]},
        # column applies (one column through M then K, or T_K then T_M) per
        # second of time-to-solution, the whole solve included
        "column_applies_per_second": {
            "fp64": r64["stats"]["column_applies"] / t64,
            "mixed": rmx["stats"]["column_applies"] / tmx,
            "mixed_learned": rlr["stats"]["column_applies"] / tlr},
        "rayleigh_ritz_seconds": {"fp64": r64["stats"]["seconds_rr"],
                                  "mixed": rmx["stats"]["seconds_rr"],
                                  "mixed_learned": rlr["stats"]["seconds_rr"]},
        "learned": dict(tinfo, training_seconds=t_train,
                        learned_iterations=rlr["stats"]["learned_iterations"],
                        fallback_at=rlr["stats"]["learned_fallback_at"]),
        "converged": [r64["conv"], rmx["conv"], rlr["conv"]],
        "mixed_learned_at_omega_target": {
            "tol": ACC_TOL, "omega_rel_err_target": 1e-7, "seconds": tacc,
            "speedup_vs_fp64": t64 / tacc, "iterations": racc["stats"]["iterations"],
            "max_rel_err_vs_fp64": e_acc, "meets_target": bool(e_acc <= 1e-7),
            "max_residual": float(np.max(racc["res"][~(zero | racc["zero"])])),
            "converged": racc["conv"]},
        "omega_fp64": [float(x) for x in r64["w"]],
        "psi0_seconds": t_itp,
        "roofline": {"bound": "hbm", "achieved": dom["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom["GBs"] / HBM_PEAK_GBS,
                     "traffic": measured_traffic(dom_kernel),
                     "kernel": f"{dom_kernel} ({KERNEL_ROLE.get(dom_pass, dom_pass)})",
                     "bytes_per_launch": dom["bytes"], "ms_per_launch": dom["ms"]},
        "passes": {k: round(v["GBs"], 1) for k, v in rf.items()},
        "launch_variants": variants,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
