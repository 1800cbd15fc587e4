#!/usr/bin/env python
"""LRes SCF wall time at 1024 electrons on one MI355X: mixed precision vs FP64.

Workload: 512 H2 (1024 electrons) on an 8x8x8 simple-cubic lattice, a = 6.0
bohr, bond 1.4 bohr, random orientations and 0.15 bohr jitter from seed 0, grid
10 points per cell (80^3), default LCAO basis (rcut 4.5, s 1.2/0.4/0.15, p 0.5),
local pseudopotential rc 0.6, closed-shell LDA, window radius 8 bohr (512
windows of 84 functions), M = 32 poles, SCF from the uniform density to
|dn|/N < 1e-6 (the reference scf_lcao defaults).  Synthetic geometry: no data.

One step = one full SCF (cold start) with the mixed-precision window step;
the same SCF with the FP64 window step is timed alongside (same iteration
count) and reported as fp64_scf_s.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "LRes SCF wall time @1024 e⁻: FP32+refine vs FP64 GPU at ΔE ≤2e-7 Ha/e⁻"
PEAK_FP64 = 78.6e12      # MI355X dense FP64 (vector and matrix), FLOP/s
PEAK_FP32 = 157.3e12     # MI355X FP32, FLOP/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=8, help="lattice cells per axis")
    ap.add_argument("--a", type=float, default=6.0,
                    help="lattice constant in bohr (grid round(a / 0.6) points per cell; 5.1 is the 0.85x compressed lattice)")
    ap.add_argument("--rw", type=float, default=8.0)
    ap.add_argument("--M", type=int, default=32)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--random", type=int, default=0,
                    help="instead of the lattice: this many H2 at random centres (min separation 3.0 bohr) in a "
                         "periodic cube at the density of the a = 6.0 lattice (uneven windows)")
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every SCF iteration's mixing/potential/H(v) step eagerly instead of "
                         "replaying it from a captured hipGraph")
    ap.add_argument("--no-forces", action="store_true",
                    help="skip the adjoint and the total-energy forces at the converged SCF (after the timed region)")
    ap.add_argument("--forces-fd", type=int, default=0, metavar="K",
                    help="also check the forces of K seeded random atoms against central differences of the "
                         "fixed-N band energy (steps 1e-4 / 5e-5 bohr, Richardson)")
    ap.add_argument("--profile-scf", action="store_true",
                    help="profiling mode: one warm-up SCF per precision, then one mixed and one fp64 SCF, each "
                         "preceded and followed by 0.2 s of GPU idle time (the SCFs' borders in a kernel trace); "
                         "prints the host wall of each and no JSON line")
    ap.add_argument("--profile-pass", type=int, default=0, metavar="K",
                    help="profiling mode: K warm window passes per precision at the workload's guess potential "
                         "and a 256 MiB device copy (the byte-counter calibration), no SCF, no JSON line")
    return ap.parse_args()


def run_scf(C, cr, grid, vion, precision, tol, graph=True):
    import torch
    from lresdft_amd.scf import scf_lcao
    C.set_precision(precision)
    C.mu = None
    C.history = []
    C.pass_log = []
    p0 = C.passes
    torch.cuda.synchronize()
    t = time.perf_counter()
    r = scf_lcao(cr, C, tol=tol, grid=grid, vion=vion, graph=graph)
    torch.cuda.synchronize()
    r["wall"] = time.perf_counter() - t
    r["passes"] = C.passes - p0
    r["history"] = list(C.history)
    r["pass_log"] = list(C.pass_log)
    return r


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pole_kernels_traffic_1024e_rw8_m32.txt")


def measured_traffic(tag):
    """HBM bytes per launch of the pole kernel (FETCH_SIZE + WRITE_SIZE, scaled by a
    calibration copy) from the committed rocprofv3 summary of this workload."""
    try:
        for line in open(TRAFFIC_FILE):
            line = line.strip()
            if line.startswith(f"traffic_bytes_per_launch_{tag} ="):
                return float(line.split("=")[1])
    except OSError:
        pass
    return None


def kernel_name(bucket, fp32):
    T, G, kind = bucket[5:8] if fp32 else bucket[2:5]
    if kind == 2 and fp32 and G == 0:
        return f"k_pole_mfma<float, tiles {T}, window-fused: 4 waves per window, pole sums and reduction inside>"
    if kind == 2:
        return f"k_pole_mfma<{'float' if fp32 else 'double'}, tiles {T}, waves {G}>"
    if T == 0:
        return f"k_pole_big<{'float' if fp32 else 'double'}, 1024 threads, panel 16>"
    return f"k_pole_gj<T {T}, {'float' if fp32 else 'double'}, grid {G}x{G}>"


def roofline(stats, flops, peak, kernel, traffic):
    passes, ms = stats
    if passes == 0 or ms <= 0:
        return None
    t = ms * 1e-3 / passes
    achieved = flops / t
    return {"bound": "mfma", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
            "frac": achieved / peak, "traffic": traffic, "kernel": kernel, "launch_ms": ms / passes,
            "flops_per_launch": flops,
            "traffic_note": ("HBM bytes per launch, rocprofv3 FETCH_SIZE + WRITE_SIZE scaled by a 256 MiB copy "
                             "(profiles/pole_kernels_traffic_1024e_rw8_m32.txt)") if traffic else None}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n, a, rw, M, tol, vion_host, sample_windows=64):
    """The oracle's LRes SCF timed on the host: the reference's eigen-route algorithm
    (lrft_windows, dsygvd / dsygvx per window, src/lrft_cpp.cpp:129-233) in the C++/OpenMP
    restatement oracle/lres_cpu.cpp with scipy's LAPACK, the reference's cold/warm/mu
    orchestration (LResEigen, lcao3d.py:400-456: truncation at c + 1.5 R, float32 Y^2
    re-density) and SCF driver (scf_lcao, lcao3d.py:237-284: scipy H(v), numpy FFT
    Hartree, LDA, DIIS).  Two full SCFs from the uniform density are timed, on one
    thread and on every core of the process's CPU set (at most 16, the host share of
    one GPU); the ion potential is formed outside the timed region, as in the GPU
    timing.  The pole route (zgesv per pole and window, window_resolvent_vs_eig.py:58-65)
    is timed per pass at the guess potential: all cores on every window, one thread on
    `sample_windows` windows (scaled to all)."""
    from oracle import cpu_native
    from oracle import lres_oracle as orc
    pts = int(round(a / 0.6))
    cr = orc.h2_crystal((n, n, n), a=a, pts_per_cell=pts, rc=0.6)
    b = orc.LCAOBasis(cr, **orc.BASIS)
    cells = orc.cell_centres((n, n, n), a)
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    out = {}
    for threads in (1, cores):
        L = orc.LResEigen(b, cells, R_w=rw, M=M, native_threads=threads, fast_windows=True)
        t = time.perf_counter()
        r = orc.scf_lcao(cr, L, tol=tol, vion=vion_host)
        wall = time.perf_counter() - t
        out[threads] = dict(scf_s=wall, iterations=r["iterations"], passes=L.passes, energy=r["energy"],
                            converged=bool(r["converged"]))
    # pole route, one pass at the guess potential
    d0 = np.full(cr.shape, cr.n_elec / (cr.ng * cr.dV))
    v = vion_host + orc.hartree(cr, d0) + orc.lda_xc(d0)[1]
    H = b.H(v)
    L = orc.LResEigen(b, cells, R_w=rw, M=M, native_threads=cores, fast_windows=True)
    _, lam, _, _, _ = cpu_native.eig_pass(L._tab, H, 0.0, -1.0, M, None, False, cores)
    e_lo = lam.min() - 0.3
    mu = e_lo + 0.65
    t = time.perf_counter()
    cpu_native.pole_pass(L._tab, H, mu, e_lo, M, False, cores)
    t_pole_all = time.perf_counter() - t
    k = min(sample_windows, L._tab.nwin)
    sub = cpu_native.Tables(b, L.windows[:k], L.own[:k])
    t = time.perf_counter()
    cpu_native.pole_pass(sub, H, mu, e_lo, M, False, 1)
    t_pole_1 = (time.perf_counter() - t) * L._tab.nwin / k
    one, allc = out[1], out[cores]
    return {"value": allc["scf_s"], "unit": "s", "cores": cores, "kind": "port",
            "value_1thread": one["scf_s"], "cpu_model": cpu_model(),
            "sample": (f"full LRes SCF (tol {tol}) of the same {n}^3 crystal on the host, eigen route "
                       f"(oracle/lres_cpu.cpp, dsygvd/dsygvx per window from scipy's LAPACK, OpenMP over windows) "
                       f"with the reference's orchestration and SCF driver; value = {cores} threads, value_1thread "
                       f"= 1 thread; ion potential precomputed as in the GPU timing"),
            "scf_1thread": one, "scf_all_cores": allc,
            "pole_route_pass_s": {"all_cores": t_pole_all,
                                  "one_thread_scaled": t_pole_1,
                                  "note": f"one pole-route pass (M {M}, zgesv per pole and window); one thread on "
                                          f"{k} of {L._tab.nwin} windows, scaled"}}


def forces_fd(C, v, F, cells, args, K, seed=11):
    """Central differences (steps 1e-4 and 5e-5 bohr, Richardson) of the fixed-N band
    energy along one random unit direction for K random atoms, v frozen, e_lo fixed."""
    from lresdft_amd import system
    from lresdft_amd.lcao3d import LRFTLCAOCpp, csr64
    cr, N, mu, e_lo = C.b.cr, C.b.cr.n_elec, C.mu, C.e_lo

    def e_band(pos):
        crp = system.Crystal(pos, cr.Z, cr.cell, cr.shape, cr.rc)
        bp = system.LCAOBasis(crp, **system.DEFAULT_BASIS)
        Cp = LRFTLCAOCpp(bp, cells, R_w=args.rw, M=args.M, device_h=False)
        Cp.plan.set_h(csr64(bp.H(v)))
        m, _ = Cp.plan.mu_solve(N, e_lo, e_lo + 1e-6, mu + 1.0, xtol=1e-14, ftol=1e-13 * N)
        return 2.0 * Cp.plan.run(m, e_lo, want_density=False)[1]

    rng = np.random.default_rng(seed)
    out, worst = [], 0.0
    for A in rng.choice(len(cr.pos), size=K, replace=False):
        u = rng.standard_normal(3)
        u /= np.linalg.norm(u)

        def f(h):
            p = cr.pos.copy()
            p[A] += h * u
            return e_band(p)

        h = 1e-4
        d1 = (f(h) - f(-h)) / (2 * h)
        d2 = (f(h / 2) - f(-h / 2)) / h
        fd = -(4 * d2 - d1) / 3
        an = float(F[A] @ u)
        worst = max(worst, abs(an - fd) / max(abs(an), 1e-3))
        out.append({"atom": int(A), "F.u": an, "-dE/dh": fd})
    return {"atoms": out, "max_rel_err": worst}


def adjoint_timing(C, grid, vion, res, cells, args):
    """The adjoint kernels (fp64) and the total-energy forces at the converged input
    density of an SCF result, timed after the benchmark's timed region."""
    import torch
    from lresdft_amd.forces import total_forces
    C.set_precision("fp64")
    v = grid.potential(res["dens_in"], vion)
    C.plan.set_potential(v)
    C.plan.adjoint(C.mu, C.e_lo)            # first call allocates the workspaces
    torch.cuda.synchronize()
    t = time.perf_counter()
    C.plan.adjoint(C.mu, C.e_lo)
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t
    total_forces(C, grid, vion, res["dens_in"])   # first call builds the per-atom boxes
    torch.cuda.synchronize()
    t = time.perf_counter()
    F, parts = total_forces(C, grid, vion, res["dens_in"])
    torch.cuda.synchronize()
    t_f = time.perf_counter() - t
    fd = forces_fd(C, v.cpu().numpy(), parts["band"], cells, args, args.forces_fd) if args.forces_fd else None
    return {"adjoint_ms": 1e3 * t_adj, "forces_s": t_f, "max_abs_force": float(abs(F).max()),
            "max_abs_band_force": float(abs(parts["band"]).max()), "fd_check": fd,
            "note": ("adjoint = d(2 sum E_B)/d(H_W, S_W) on the GPU (fp64, all windows); forces = -dE/dR of the "
                     "Harris-Foulkes SCF energy at the converged input density: fixed-N band-energy forces with the "
                     "potential frozen (beta from two passes, adjoint, contraction with the basis derivatives in "
                     "per-atom boxes, all on the GPU) plus the ion-potential response (FFT); fd_check: central "
                     "differences of the band energy along random directions")}


def profile_passes(C, cr, grid, vion, K):
    """Warm passes at the guess potential (uniform density), both precisions, each with
    its H(v) assembly and its density, then a 256 MiB device-to-device copy whose
    FETCH_SIZE / WRITE_SIZE calibrate the byte counters."""
    import torch
    d0 = torch.full(cr.shape, cr.n_elec / (cr.ng * cr.dV), dtype=torch.float64, device="cuda")
    v = grid.potential(d0, vion)
    C.plan.set_potential(v)
    lmin = C.plan.lambda_min()
    e_lo, mu = lmin - C.margin, lmin + 0.35
    for prec in ("mixed", "fp64"):
        C.set_precision(prec)
        for _ in range(K + 1):          # the first pass also runs the autotuner
            C.plan.set_potential(v)
            C.plan.run(mu, e_lo, want_density=False)
            C.plan.density_last()
    src = torch.ones(32 * 2 ** 20, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"profile passes done: {K} per precision, buckets {C.plan.buckets()}", flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the window step shards by "
              "windows across GPUs, which this build does not launch", file=sys.stderr)
        sys.exit(2)
    from lresdft_amd import _native
    _native.lib()
    import torch
    from lresdft_amd import system
    from lresdft_amd.lcao3d import LRFTLCAOCpp
    from lresdft_amd.scf import GridOps

    n = args.n
    if args.random:
        box = 6.0 * args.random ** (1.0 / 3.0)
        cr, cells = system.random_h2_box(args.random, box, dmin=3.0, seed=args.seed, dx=0.6, rc=0.6)
        workload = (f"{args.random} H2 at random centres (d_min 3.0, seed {args.seed}) in a {box:.2f} bohr cube, "
                    f"{cr.n_elec} e-, R_w {args.rw}, M {args.M}, SCF tol {args.tol}")
    else:
        pts = int(round(args.a / 0.6))
        cr = system.h2_crystal((n, n, n), a=args.a, pts_per_cell=pts, rc=0.6)
        cells = system.cell_centres((n, n, n), args.a)
        workload = (f"h2_crystal {n}x{n}x{n} a={args.a} ({pts} grid points per cell) jitter 0.15 seed 0, "
                    f"{cr.n_elec} e-, R_w {args.rw}, M {args.M}, SCF tol {args.tol}")
    # per-geometry set-up, timed beside the SCFs (not part of their wall): basis on the host,
    # grid operators and the ion potential on the device, windows and plan (device tables);
    # the device is initialised first, so that its start-up is not counted
    torch.zeros(1, device="cuda")
    torch.cuda.synchronize()
    t_s = time.perf_counter()
    b = system.LCAOBasis(cr, **system.DEFAULT_BASIS)
    t_basis = time.perf_counter() - t_s
    t_s = time.perf_counter()
    grid = GridOps(cr)
    vion = grid.ion_potential()
    torch.cuda.synchronize()
    t_ion = time.perf_counter() - t_s
    t_s = time.perf_counter()
    C = LRFTLCAOCpp(b, cells, R_w=args.rw, M=args.M)
    torch.cuda.synchronize()
    t_plan = time.perf_counter() - t_s
    flops = C.plan.flops_per_pass()

    if args.profile_pass:
        profile_passes(C, cr, grid, vion, args.profile_pass)
        return
    if args.profile_scf:
        for prec in ("fp64", "mixed"):
            run_scf(C, cr, grid, vion, prec, args.tol, not args.no_graph)
        for prec in ("mixed", "fp64"):
            torch.cuda.synchronize()
            time.sleep(0.2)
            r = run_scf(C, cr, grid, vion, prec, args.tol, not args.no_graph)
            time.sleep(0.2)
            print(f"profile SCF {prec}: wall {1e3 * r['wall']:.3f} ms, {r['iterations']} iterations, "
                  f"{r['passes']} passes, passes per iteration {r['pass_log']}, history {r['history']}, "
                  f"residuals {['%.1e' % e for e in r['residuals']]}", flush=True)
        return

    for _ in range(args.warmup):
        run_scf(C, cr, grid, vion, "fp64", args.tol, not args.no_graph)
        run_scf(C, cr, grid, vion, "mixed", args.tol, not args.no_graph)

    C.plan.stats(reset=True)
    fp64 = [run_scf(C, cr, grid, vion, "fp64", args.tol, not args.no_graph) for _ in range(args.steps)]
    st64 = C.plan.stats(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    mixed = [run_scf(C, cr, grid, vion, "mixed", args.tol, not args.no_graph) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t0
    st32 = C.plan.stats(reset=True)

    t_mixed = t_total / args.steps
    bk = C.plan.buckets()
    default_workload = (not args.random and args.n == 8 and args.rw == 8.0 and args.M == 32 and len(bk) == 1)
    t_fp64 = sum(r["wall"] for r in fp64) / args.steps
    dE = (mixed[-1]["energy"] - fp64[-1]["energy"]) / cr.n_elec
    line = {
        "metric": METRIC, "value": t_mixed, "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * t_mixed, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64+f64", "data": "synthetic",
        "config": {"workload": workload,
                   "windows": int(C.plan.nwin), "n_W": int(C.plan.n.max()), "n_W_min": int(C.plan.n.min()),
                   "n_W_mean": float(C.plan.n.mean()), "grid": list(cr.shape),
                   "buckets": C.plan.buckets()},
        "fp64_scf_s": t_fp64, "speedup_vs_fp64": t_fp64 / t_mixed, "dE_per_e_vs_fp64": dE,
        "energy_fp64": fp64[-1]["energy"], "energy_mixed": mixed[-1]["energy"],
        "iterations": {"fp64": fp64[-1]["iterations"], "mixed": mixed[-1]["iterations"]},
        "passes_per_scf": {"fp64": fp64[-1]["passes"], "mixed": mixed[-1]["passes"]},
        "history": {"fp64": fp64[-1]["history"], "mixed": mixed[-1]["history"]},
        "passes_per_iteration": {"fp64": fp64[-1]["pass_log"], "mixed": mixed[-1]["pass_log"]},
        "graph": {"mixing_potential_H_step_replayed": not args.no_graph,
                  "replays_per_scf": {"fp64": fp64[-1]["graph_replays"], "mixed": mixed[-1]["graph_replays"]}},
        "roofline": roofline(st32["fp32"], flops, PEAK_FP32, kernel_name(bk[0], True),
                             measured_traffic("fp32_fused" if bk[0][6] == 0 and bk[0][7] == 2 else "fp32")
                             if default_workload else None),
        "roofline_fp64": roofline(st64["fp64"], flops, PEAK_FP64, kernel_name(bk[0], False),
                                  measured_traffic("fp64") if default_workload else None),
        "cpu_baseline": None,
        "setup_s": {"basis_host": t_basis, "grid_ops_and_ion_potential": t_ion, "windows_and_plan": t_plan,
                    "note": ("per geometry, outside the SCF wall (value); grid_ops_and_ion_potential includes the "
                             "FFT library's first-use set-up (runtime kernel compilation, once per process); the "
                             "first pass of each precision also runs the kernel autotuner, which the warm-up SCFs "
                             "absorb")},
    }
    if not args.no_forces:
        line["adjoint"] = adjoint_timing(C, grid, vion, mixed[-1], cells, args)
    if not args.no_cpu_baseline and not args.random:
        line["cpu_baseline"] = cpu_baseline(n, args.a, args.rw, args.M, args.tol,
                                            vion.cpu().numpy().reshape(cr.shape))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
