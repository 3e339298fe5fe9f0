#!/usr/bin/env python3
"""Benchmark: seconds per BSE build + lowest-16 exciton solve on one MI355X.

Workload (default "c4"): seeded 18-orbital triangular model (3 atoms x 6 orbitals,
a = 3.19 A), 96 x 96 Monkhorst-Pack mesh, nv = nc = 2 (BSE dimension 36864), Keldysh
r0 = 40 A, eps_m = eps_s = 1, lattice-sum cutoff 25 cells, exchange on, block
Davidson (block 32) for the lowest 16 states.

A step is: lattice Fourier sums + BSE assembly + Davidson solve, with the band
stacks already resident in HBM (the host band problem runs once, before timing).
Two forms are timed: the FP64 plain path (FP64 assembly, FP64 products) and the
mixed path (FP32 assembly and FP32 products, then FP64 refinement with matrix-free
FP64 products). ``value`` is the mixed path's seconds per step; the FP64 time, the
speed-up and the per-eigenvalue deviation from the FP64 path are reported beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

CONFIGS = {
    # name: (natoms, orbitals/atom, n, nbands, r0, cutoff, nstates, block, matrix_free)
    "c2": (3, 6, 48, 2, 40.0, 25.0, 16, 32, False),
    "c3": (3, 6, 36, 2, 40.0, 25.0, 8, 32, False),
    "c4": (3, 6, 96, 2, 40.0, 25.0, 16, 32, False),
    "c5": (4, 9, 120, 3, 40.0, 25.0, 16, 32, True),
}

PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
TOL = 1e-5  # FP64 residual tolerance (eV) of both forms


def measured_traffic(config, kind):
    """HBM bytes per dense-product launch from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or None if that summary is absent."""
    path = ROOT / "profiles" / f"dense_product_{config}_traffic.md"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            vals = dict(kv.split("=") for kv in line.split(":", 1)[1].split())
            v = float(vals[kind]) if kind in vals else float("nan")
            return v if np.isfinite(v) else None
    return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    """The host CPU's model name (first processor in /proc/cpuinfo) and its logical CPU count."""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{name} ({os.cpu_count()} logical CPUs on the host)"


def cpu_baseline(model_path, cfg, n_products_fp64, nstates, block):
    """Time the CPU oracle (the reference algorithm restated in C++, OpenMP) on a
    bounded sample of the same workload and scale it to one full step."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O

    natoms, orb, n, nb, r0, cutoff, _, _, _ = cfg
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    ns = 32  # sample mesh: 32 x 32 of the same model
    s = O.System(str(model_path))
    s.brillouin_zone_mesh(ns)
    ex = O.Exciton(s, ns, nb, 0, (1.0, 1.0, r0))
    ex.cutoff = cutoff
    ex.exchange = True
    t0 = time.perf_counter()
    ex.initialize()           # bands (numpy) + lattice sums at the sample mesh (C++ oracle)
    t_ft_sample = time.perf_counter() - t0
    t0 = time.perf_counter()
    H = ex.bse()              # per-element assembly loop (C++ oracle, OpenMP)
    t_bse_sample = time.perf_counter() - t0
    X = np.random.default_rng(0).normal(size=(ex.N, block)) + 0j
    t0 = time.perf_counter()
    for _ in range(3):
        H @ X
    t_mv_sample = (time.perf_counter() - t0) / 3
    N_full = (n ** 2) * nb * nb
    scale2 = (N_full / ex.N) ** 2
    # lattice sums: cost ~ mesh points x cells; cells at the full mesh from the oracle itself
    ncells_full = len(s.truncate_supercell(n, s.a * cutoff))
    t_ft = t_ft_sample * (n ** 2 / ns ** 2) * (ncells_full / len(ex.cells))
    t_bse = t_bse_sample * scale2
    t_solve = t_mv_sample * scale2 * (n_products_fp64 / block)
    total = t_ft + t_bse + t_solve
    return {
        "value": round(total, 3), "unit": "s per build+solve (extrapolated)", "cores": threads, "kind": "port",
        "cpu_model": cpu_model(),
        "sample": (f"oracle at a {ns}x{ns} mesh of the same model (N={ex.N}): lattice sums {t_ft_sample:.2f}s, "
                   f"per-element assembly {t_bse_sample:.2f}s, one {block}-vector FP64 product {t_mv_sample*1e3:.1f}ms; "
                   f"scaled to N={N_full} (assembly and products x(N/Ns)^2, sums x mesh x cells) with "
                   f"{n_products_fp64} product vectors as in the GPU FP64 solve"),
        "parts_s": {"fourier": round(t_ft, 3), "assembly": round(t_bse, 3), "solve": round(t_solve, 3)},
    }


def extras(ex, config, nstates, mf):
    """Workload-specific figures measured through the ExcitonTB facade after the timed steps:
    c3 - Hellmann-Feynman gradients of the lowest 8 energies w.r.t. r0, eps_s and every
         hopping amplitude (FP64 adjoint kernels);
    c5 - matrix-free mixed solve, absorption spectrum along x (0.05 eV Lorentzians, 2000
         points over 0-6 eV) and dE/d(r0, eps_s) of the lowest exciton."""
    ex.setPrecision("fp32" if config == "c5" else "fp64")
    ex.setMatrixFree(bool(mf))
    ex.setDavidsonTolerance(TOL)
    ex.BShamiltonian()
    t0 = time.perf_counter()
    res = ex.diagonalize("davidson", nstates)
    t_solve = time.perf_counter() - t0
    out = {"solve_s": round(t_solve, 4), "energies_eV": [round(float(e), 6) for e in np.asarray(res.eigval)]}
    if config == "c3":
        t0 = time.perf_counter()
        g = np.asarray(ex.hfGradients(None, nstates, ["r0", "eps_s"]))
        t_hf = time.perf_counter() - t0
        t0 = time.perf_counter()
        G = np.asarray(ex.hoppingGradients(None, nstates))
        t_hop = time.perf_counter() - t0
        out.update({"hf_gradients_s": round(t_hf, 4), "hopping_gradients_s": round(t_hop, 4),
                    "n_hopping_entries": int(G[0].size), "dE_dr0": [float(f"{x:.6g}") for x in g[:, 0]],
                    "dE_deps_s": [float(f"{x:.6g}") for x in g[:, 1]]})
    else:
        t0 = time.perf_counter()
        spec = np.asarray(res.absorptionSpectrum(0.0, 6.0, 2000, 0.05))
        t_abs = time.perf_counter() - t0
        t0 = time.perf_counter()
        g = np.asarray(ex.hfGradients(None, 1, ["r0", "eps_s"]))
        t_hf = time.perf_counter() - t0
        i = int(np.argmax(spec[:, 1]))
        out.update({"absorption_s": round(t_abs, 4), "absorption_peak_eV": float(spec[i, 0]),
                    "hf_gradients_s": round(t_hf, 4), "dE0_dr0": float(g[0, 0]), "dE0_deps_s": float(g[0, 1])})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c4", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the FP64 plain path")
    ap.add_argument("--max-basis", type=int, default=0, help="Davidson basis size before restart (0: library default)")
    ap.add_argument("--block", type=int, default=0, help="Davidson block size (0: the workload's)")
    ap.add_argument("--time-tiles", action="store_true",
                    help="matrix-free workloads: also time one 32-vector product of the recomputed-tiles "
                         "operator (FP32 tiles) against the convolution operator on the same vectors")
    args = ap.parse_args()
    # exactly one JSON line on stdout: everything else (including what the native library
    # prints) goes to stderr
    sys.stdout.flush()
    json_fd = os.dup(1)
    os.dup2(2, 1)
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path runs on a single "
              "MI355X; the design notes describe how it would shard)", file=sys.stderr)
        sys.exit(2)

    import xatu_amd as X
    from xatu_amd import capi, synthetic

    cfg = CONFIGS[args.config]
    natoms, orb, n, nb, r0, cutoff, nstates, block, mf = cfg
    if args.block > 0:
        block = args.block
    tmp = Path(tempfile.mkdtemp(prefix="xatu_bench_"))
    model = synthetic.write_model(tmp / "model.model", natoms=natoms, orb_per_atom=orb)

    log(f"config {args.config}: natoms={natoms} orb={orb} mesh {n}x{n} nv=nc={nb}")
    t0 = time.perf_counter()
    ex = X.ExcitonTB(X.SystemConfiguration(str(model)), n, nb, 0, [1.0, 1.0, r0])
    ex.setCutoff(cutoff)
    ex.setExchange(True)
    ex.brillouinZoneMesh(n)
    ex.initializeHamiltonian()  # host bands + device lattice sums (untimed set-up)
    t_setup = time.perf_counter() - t0
    N = int(ex.excitonbasisdim)
    log(f"N={N}; host bands + set-up {t_setup:.2f}s")

    s = ex.system
    ctx = capi.Context(0)
    tau = np.asarray(s.motif)[:, :3]
    ctx.set_lattice(np.asarray(s.bravaisLattice), tau, s.orbitalsPerAtom(), s.meshPoints)
    ctx.set_bands(nb, nb, ex.eigvalKStack, ex.eigvecKStack, np.asarray(s.kpointOffsets),
                  ex.eigvalKQStack, ex.eigvecKQStack)
    inter = dict(eps_m=1.0, eps_s=1.0, r0=r0, cutoff=cutoff, ncell=n, total_cells=int(ex.totalCells),
                 regularization=float(ex.regularization), exchange=True)

    def step(prec, arith):
        ctx.fourier_sums(**inter)
        ctx.build(prec if not mf else capi.XATU_MATRIX_FREE)
        t_build = ctx.timings()
        w, v, rep = ctx.solve(method=capi.XATU_SOLVE_DAVIDSON, nstates=nstates, precision=prec,
                              matrix_free=mf, block=block, max_basis=args.max_basis, tol=TOL, want_vectors=False,
                              fp32_arith=arith)
        return w, rep, t_build

    results = {}
    skip_plain = args.no_plain or mf  # the FP64 form of a matrix-free workload is not a product path worth timing
    # mixed: products with the FP32 matrix on the FP16 matrix cores with split operands (the
    # default); mixed_fp32mfma: the same solve with the products on the FP32 matrix cores
    forms = (("fp64", capi.XATU_FP64, 0), ("mixed_fp32mfma", capi.XATU_FP32, capi.XATU_ARITH_FP32),
             ("mixed", capi.XATU_FP32, capi.XATU_ARITH_SPLIT16))
    for name, prec, arith in forms:
        if name != "mixed" and skip_plain:
            continue
        for i in range(args.warmup):
            log(f"{name} warmup {i}")
            step(prec, arith)
        times, reps = [], []
        for i in range(args.steps):
            t0 = time.perf_counter()
            w, rep, tb = step(prec, arith)
            dt = time.perf_counter() - t0
            times.append(dt)
            reps.append(rep)
            log(f"{name} step {i}: {dt:.3f}s (fourier {tb.ms_fourier:.2f}ms, build {tb.ms_build:.2f}ms, "
                f"solve {rep.ms_total:.1f}ms [products {rep.ms_products:.1f}, RR {rep.ms_rayleigh_ritz:.1f}, "
                f"proj {rep.ms_projections:.1f}], iters {rep.iterations}, products {rep.products}, "
                f"refine {rep.refine_iterations}, resid {rep.max_residual:.2e})")
        tm = ctx.timings()
        results[name] = dict(times=times, w=w, rep=reps[-1], tim=tm, build=tb)
        if not mf:
            ctx.free_bse(prec)

    mixed = results["mixed"]
    ms_step = float(np.median(mixed["times"])) * 1e3
    out = {
        "metric": "s per BSE build+lowest-16 solve (mixed form: FP32 matrix, split-FP16 MFMA products with FP32 "
                  "accumulation, FP64 refinement; every eigenvalue within 0.1 meV of FP64)",
        "value": round(ms_step / 1e3, 4), "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_step, 2), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (split-f16 MFMA, f32 accumulate) + f64", "data": "synthetic (seeded tight-binding model)",
        "config": {"workload": f"{args.config}: seeded {natoms * orb}-orbital triangular model, {n}x{n} k-mesh, "
                               f"nv=nc={nb}, BSE dim {N}, Keldysh r0={r0}, cutoff {cutoff} cells, exchange on, "
                               f"Davidson block {block}, lowest {nstates}" + (", matrix-free" if mf else ""),
                   "bse_dim": N},
    }
    if "fp64" in results:
        p = results["fp64"]
        ms64 = float(np.median(p["times"])) * 1e3
        dev = (mixed["w"] - p["w"]) * 1e3
        out["fp64_plain_s"] = round(ms64 / 1e3, 4)
        out["speedup_mixed_vs_fp64"] = round(ms64 / ms_step, 3)
        out["eig_dev_meV_max"] = float(np.abs(dev).max())
        out["eig_dev_meV"] = [float(f"{x:.3g}") for x in dev]
    if "mixed_fp32mfma" in results:
        q = results["mixed_fp32mfma"]
        out["mixed_fp32mfma_s"] = round(float(np.median(q["times"])), 4)
        if "fp64" in results:
            out["mixed_fp32mfma_eig_dev_meV_max"] = float(np.abs((q["w"] - results["fp64"]["w"]) * 1e3).max())
    rep = mixed["rep"]
    tim = mixed["tim"]
    # throughput figures of the headline step: converged eigenpairs per second of the whole
    # step; vector products H x per second (all of the solve's products over their kernel
    # time, FP32 phase and FP64 refinement); lattice Fourier sums as algorithmic bytes
    # (8 natoms^2 nR of V_ab(R) + 16 natoms^2 nMesh of sums + 24 nR of cell vectors) per
    # second of their kernels
    out["eigenpairs_per_s"] = round(nstates / (ms_step / 1e3), 2)
    if rep.ms_products > 0:
        out["matvecs_per_s"] = round(rep.products / (rep.ms_products / 1e3), 1)
    n_r = len(ctx.get_cells())
    ft_bytes = 8.0 * natoms ** 2 * n_r + 16.0 * natoms ** 2 * n * n + 24.0 * n_r
    if mixed["build"].ms_fourier > 0:
        out["fourier_GBs"] = round(ft_bytes / (mixed["build"].ms_fourier / 1e3) / 1e9, 2)
        out["fourier_bytes"] = ft_bytes
    if "fp64" in results and "mixed_fp32mfma" in results:
        out["speedup_fp32mfma_vs_fp64"] = round(float(np.median(results["fp64"]["times"])) /
                                                float(np.median(results["mixed_fp32mfma"]["times"])), 3)
    out["mixed_parts_ms"] = {"fourier": round(mixed["build"].ms_fourier, 3),
                             "build": round(mixed["build"].ms_build, 3), "solve": round(rep.ms_total, 2),
                             "products": round(rep.ms_products, 2), "rayleigh_ritz": round(rep.ms_rayleigh_ritz, 2),
                             "projections": round(rep.ms_projections, 2),
                             "iterations": rep.iterations, "refine_iterations": rep.refine_iterations}
    # Roofline of the dominant kernel of the mixed step: the product with the FP32 matrix
    # (staging, product and reduction kernels, bracketed by HIP events on their stream).
    # With N even it runs on the FP16 matrix cores with split operands and reads only the
    # block-lower half of the Hermitian matrix; its algorithmic bytes are the N(N+1)/2
    # distinct complex entries (8 bytes each). Otherwise the full matrix on the FP32 matrix
    # cores.
    if tim.n_matvec_kernel > 0:
        nvec_avg = tim.n_matvec_vectors / tim.n_matvec_kernel
        flops = 8.0 * N * N * nvec_avg
        t = tim.ms_matvec_kernel / 1e3
        common = {"avg_ms": round(tim.ms_matvec_kernel, 3), "launches": tim.n_matvec_kernel,
                  "avg_vectors": round(nvec_avg, 2), "complex_TFLOPs_4M_equiv": round(flops / t / 1e12, 2)}
        if N % 2 == 0:
            algo = 4.0 * N * (N + 1)
            gbs = algo / t / 1e9
            out["roofline"] = {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                               "frac": round(gbs / PEAK_HBM_GBS, 4),
                               "traffic": measured_traffic(args.config, "split16"),
                               "kernel": "Hermitian split-FP16 product of the FP32 matrix H X (block-lower half of "
                                         "N x N read once, by N x 32)",
                               "algorithmic_bytes": algo, "full_matrix_equivalent_GBs": round(8.0 * N * N / t / 1e9, 1),
                               **common}
        else:
            ach = flops / t / 1e12
            out["roofline"] = {"bound": "mfma", "achieved": round(ach, 2), "peak": PEAK_FP32_TFLOPS,
                               "unit": "TFLOP/s", "frac": round(ach / PEAK_FP32_TFLOPS, 4),
                               "traffic": measured_traffic(args.config, "fp32"),
                               "kernel": "FP32 dense product H X (N x N by N x 32)",
                               "hbm_GBs": round(8.0 * N * N / t / 1e9, 1), "algorithmic_bytes": 8.0 * N * N,
                               **common}
    if args.time_tiles and mf:
        X = np.asfortranarray(np.random.default_rng(1).normal(size=(N, 32)) + 0j)
        forms_mf = (("conv", capi.XATU_MATRIX_FREE_CONV), ("tiles", capi.XATU_MATRIX_FREE_TILES))
        got = {}
        for name, form in forms_mf:
            log(f"one 32-vector product, matrix-free form {name}")
            t0 = time.perf_counter()
            got[name] = ctx.matvec(X, precision=capi.XATU_FP32, matrix_free=form)
            got[name + "_s"] = time.perf_counter() - t0
        out["matrix_free_product_s"] = {"convolution": round(got["conv_s"], 4), "tiles": round(got["tiles_s"], 3),
                                        "vectors": 32, "rel_diff": float(np.abs(got["conv"] - got["tiles"]).max() /
                                                                         np.abs(got["tiles"]).max())}
    if args.config in ("c3", "c5"):
        ctx.close()
        out["extras"] = extras(ex, args.config, nstates, mf)
        ctx = None
    if not args.no_cpu:
        n_products = results["fp64"]["rep"].products if "fp64" in results else rep.products
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(model, cfg, n_products, nstates, block)
    if ctx is not None:
        ctx.close()
    sys.stdout.flush()
    os.write(json_fd, (json.dumps(out) + "\n").encode())
    os.close(json_fd)


if __name__ == "__main__":
    main()
