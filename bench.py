"""Benchmark: 3D PIC Stokes solves/hour to 1e-6 residual (mixed precision + learned Schur scaling vs plain
FP64), plus assembly GB/s, on one MI355X.

Default workload (the metric's configuration): 128x128x64 Q1/P0 elements on [0,2]x[0,2]x[0,1], 20 particles/cell
(21.0M particles, seeded MINSTD layout), log10 eta from a seeded Gaussian random field (correlation length 0.1)
mapped to [-3, 3] (contrast 1e6) and sampled at every particle (61 materials of 0.1 decade), density
proportional to -log10 eta, free-slip walls, gravity 1.  One step = particle assembly (fused rheology) + a
BSSCR solve to a relative FP64 Stokes residual <= 1e-6, inputs (swarm, materials, BCs) resident in HBM before
the timed region.  --workload spheres runs the 96^3 spheres configuration instead.

Prints one JSON line.
"""
import argparse
import ctypes as C
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MLP_WEIGHTS = os.path.join(ROOT, "earthbyte_underworld_amd", "data", "schur_mlp_3d.npz")
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def events(lib):
    e0, e1 = C.c_void_p(), C.c_void_p()
    lib.ust_event_create(C.byref(e0))
    lib.ust_event_create(C.byref(e1))
    return e0, e1


def time_on_stream(lib, stream, fn, reps):
    e0, e1 = events(lib)
    fn()
    lib.ust_stream_sync(stream)
    lib.ust_event_record(e0, stream)
    for _ in range(reps):
        fn()
    lib.ust_event_record(e1, stream)
    ms = C.c_float()
    lib.ust_event_elapsed_ms(e0, e1, C.byref(ms))
    lib.ust_event_destroy(e0)
    lib.ust_event_destroy(e1)
    return ms.value / reps


def solve_opts(ust, precision, learned, relax=0, degree=2, coarse_degree=0, cheby_ratio=12.0, relax_cap=1e-2,
               inner_fp32=0, vcycle_bf16=1):
    from earthbyte_underworld_amd import _lib as L
    # stokes_rtol: the metric's 1e-6 bar on the FP64 Stokes residual is enforced, not assumed
    return ust.default_opts(precision=precision, schur_pc=L.SCHUR_LEARNED if learned else L.SCHUR_UW,
                            outer_rtol=1e-7, inner_rtol=1e-8, max_outer=2000, inner_relax=relax,
                            inner_relax_cap=relax_cap, smoother_degree=degree, coarse_degree=coarse_degree,
                            cheby_ratio=cheby_ratio, stokes_rtol=1e-6, inner_fp32=inner_fp32,
                            vcycle_bf16=vcycle_bf16)


def _grf_oracle_sample(O, n, ppc, seed, threads):
    """The random-field configuration at n x n x n/2 elements on the oracle: the same layout rule, the field
    sampled at the nodes and interpolated at every particle, the same 61 materials.  Returns assembly and solve
    times and iteration counts."""
    from earthbyte_underworld_amd.configs import gaussian_random_field
    O.set_threads(threads)
    nel, box = (n, n, n // 2), (2.0, 2.0, 1.0)
    m = O.make_mesh(3, nel, (0, 0, 0), box)
    ne = int(np.prod(nel))
    coords = O.layout_random(m, ppc, 13)
    off = np.arange(0, ne * ppc + 1, ppc).astype(np.int32)
    xi, w = O.map_particles(m, coords, off)
    nodes = tuple(k + 1 for k in nel)
    f = gaussian_random_field(nodes, tuple(b * (k + 1) / k for b, k in zip(box, nel)), 0.1, seed)
    f = (f - f.min()) / (f.max() - f.min()) * 6.0 - 3.0
    g = coords / (np.array(box) / np.array(nel))
    i0 = np.minimum(np.floor(g).astype(int), np.array(nel) - 1)
    t = g - i0
    v = np.zeros(len(w))
    for a in range(8):
        b = [(a >> d) & 1 for d in range(3)]
        v += np.prod([t[:, d] if b[d] else 1 - t[:, d] for d in range(3)], axis=0) * \
            f[i0[:, 0] + b[0], i0[:, 1] + b[1], i0[:, 2] + b[2]]
    mat = np.clip(np.rint((v + 3.0) / 6.0 * 60), 0, 60).astype(np.int32)
    mats = [O.make_material([(O.RH_MATERIALVISCOSITY, (10.0 ** lv,))], -lv / 3.0) for lv in np.linspace(-3, 3, 61)]
    t0 = time.perf_counter()
    eta, _ = O.eval_viscosity(m, mats, off, xi, mat)
    rho = np.array([mats[k].density for k in mat])
    Ke, Ge, fe, aS = O.assemble_elements(m, off, xi, w, eta, rho=rho, gravity=1.0)
    nn = int(np.prod(nodes))
    idx = np.arange(nn)
    ijk = [idx % nodes[0], (idx // nodes[0]) % nodes[1], idx // (nodes[0] * nodes[1])]
    mask = np.zeros((nn, 3), np.uint8)
    for a in range(3):
        mask[(ijk[a] == 0) | (ijk[a] == nel[a]), a] = 1
    gs = O.GlobalSystem(m, mask.reshape(-1), np.zeros(nn * 3), Ke, Ge, fe)
    t_asm = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, _, st = gs.solve(aS, inner_rtol=1e-8, outer_rtol=1e-7, smooth_its=1)
    t_solve = time.perf_counter() - t0
    return {"n": n, "elements": ne, "threads": threads, "t_asm": t_asm, "t_solve": t_solve, "outer": st.outer_its,
            "inner": st.inner_its_total, "residual": st.stokes_rel_residual}


def cpu_baseline(nel_full, ppc, seed, outer_full, threads, sizes=(8, 12, 16, 24), single_sizes=(8, 12, 16)):
    """The FP64 CPU oracle (the same algorithm: particle assembly, CSR operators, BSSCR with Galerkin GMG and
    SSOR) built with -march=native on this host, timed on samples of the configuration at n x n x n/2.  The
    full-size estimate separates the two factors: the cost of one outer (Schur CG) iteration, fitted as
    c N^b over the samples (it contains the inner solves, so the growth of inner iterations is in b), times the
    outer iteration count of the full-size problem, which the plain FP64 GPU run of the same BSSCR measured;
    plus the assembly, fitted the same way.  Done with all threads and with one.  The sample sizes halve down
    to a small coarsest grid (as the full grid does), so that no sample is dominated by its dense coarse solve;
    the symmetric Gauss-Seidel smoother is sequential, as PETSc's SOR is within one rank, so threads speed up
    only the products and dots."""
    from oracle import oracle as O
    O.use_native()
    out = {}
    for label, nth, szs in (("all", threads, sizes), ("single", 1, single_sizes)):
        runs = [_grf_oracle_sample(O, n, ppc, seed, nth) for n in szs]
        N = np.array([r["elements"] for r in runs], float)
        per_outer = np.array([r["t_solve"] / r["outer"] for r in runs])
        b, lc = np.polyfit(np.log(N), np.log(per_outer), 1)
        ba, la = np.polyfit(np.log(N), np.log([max(r["t_asm"], 1e-6) for r in runs]), 1)
        t_full = np.exp(lc) * nel_full ** b * outer_full + np.exp(la) * nel_full ** ba
        out[label] = {"solves_per_hour": 3600.0 / t_full, "s_per_solve": t_full, "per_outer_exponent": round(b, 3),
                      "assembly_exponent": round(ba, 3), "samples": runs}
        log(f"cpu baseline ({label}, {nth} threads): {t_full:.0f} s per full-size solve "
            f"(per-outer-iteration cost ~ N^{b:.2f}, {outer_full} outer iterations)")
    a = out["all"]
    largest = a["samples"][-1]
    return {"value": a["solves_per_hour"], "unit": "solves/hour", "cores": threads, "kind": "port",
            "single_thread": {"value": out["single"]["solves_per_hour"], "cores": 1,
                              "s_per_solve": out["single"]["s_per_solve"]},
            "s_per_solve": a["s_per_solve"],
            "sample": f"FP64 C oracle (gcc -O3 -march=native, OpenMP {threads} threads) on the random-field "
                      f"configuration at n x n x n/2 elements, n = {list(sizes)} (largest: {largest['elements']} "
                      f"elements, {largest['t_asm']:.2f}s assembly + {largest['t_solve']:.2f}s solve, "
                      f"{largest['outer']} outer / {largest['inner']} inner its); full size = fitted cost per "
                      f"outer iteration (~N^{a['per_outer_exponent']}) x the {outer_full} outer iterations the "
                      f"plain FP64 GPU solve of the full problem took, + fitted assembly",
            "fits": {k: {kk: vv for kk, vv in v.items() if kk != "samples"} for k, v in out.items()},
            "samples": {k: v["samples"] for k, v in out.items()}}


def fine_level_kernels(lib, ctx, s, stream, reps=50):
    """The three operator passes every inner iteration of the headline solve makes on the finest grid, all with
    FP32 K weights: the FP32 product q = K p of the velocity PCG (EPI_Y), the first pre-smoothing step from zero
    with the multigrid residual, x = c2 D^-1 b and r = b - K x (EPI_RES0), and the prolongation + first
    post-smoothing step (EPI_PCHEBY0); and the FP64 product of the
    refinement residuals (a few per inner solve).  Each is launched reps times through the kernel ABI on the
    library stream and timed with HIP events there; algorithmic HBM bytes per launch are the weights (27 per
    element at the operator's weight precision) plus the vectors each pass must read and write once.  Returns
    the roofline object of the per-iteration pass that takes the most time and the figures of all four."""
    from earthbyte_underworld_amd import _lib as L
    g = L.Grid()
    WK, WG, WF, mask = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
    lib.ust_get_device_views(ctx, C.byref(g), C.byref(WK), C.byref(WG), C.byref(WF), C.byref(mask))
    wl = lib.ust_k_weights_len(C.byref(g))
    n = s.ndof
    nc = int(np.prod([g.nel[d] // 2 + 1 for d in range(3)])) * 3
    work = lib.ust_k_apply_work_len(C.byref(g))
    bufs = {}
    for name, nbytes in (("W32", wl * 4), ("x", n * 4), ("b", n * 4), ("r", n * 4), ("invd", n * 4), ("d", n * 4),
                         ("xc", nc * 4), ("work", work * 8), ("u64", n * 8), ("y64", n * 8)):
        bufs[name] = C.c_void_p()
        assert lib.ust_dev_malloc(C.byref(bufs[name]), nbytes) == 0
    lib.ust_k_copy_f64_to_f32(WK, bufs["W32"], wl, stream)
    rng = np.random.default_rng(0)
    for name, cnt, dt in (("x", n, np.float32), ("b", n, np.float32), ("d", n, np.float32), ("xc", nc, np.float32),
                          ("u64", n, np.float64)):
        a = (1e-3 * rng.normal(size=cnt)).astype(dt)
        lib.ust_memcpy_h2d(bufs[name], a.ctypes.data_as(C.c_void_p), a.nbytes, stream)
    a = np.full(n, 1e-3, np.float32)  # a contracting step keeps x bounded over the repetitions
    lib.ust_memcpy_h2d(bufs["invd"], a.ctypes.data_as(C.c_void_p), n * 4, stream)
    B = bufs
    runs = {
        "k_applyK3_col<float,float,EPI_Y> + k_gather_tiles (FP32 PCG product)": (
            lambda: lib.ust_k_apply_K_f32(C.byref(g), B["W32"], mask, B["x"], B["r"], B["work"], stream),
            s.nel * 27 * 4 + n * (2 * 4 + 1), True),
        "k_applyK3_col<float,float,EPI_RES0> + k_gather_tiles (FP32 first Chebyshev step + multigrid residual)": (
            lambda: lib.ust_k_apply_K_residual_first_f32(C.byref(g), B["W32"], mask, B["b"], B["invd"], B["x"],
                                                         B["r"], C.c_float(0.5), B["work"], stream),
            s.nel * 27 * 4 + n * (4 * 4 + 1), True),
        "k_applyK3_col<float,float,EPI_PCHEBY0> + k_gather_tiles (FP32 prolongation + Chebyshev step)": (
            lambda: lib.ust_k_apply_K_cheby_prolong_f32(C.byref(g), B["W32"], mask, B["x"], B["b"], B["invd"],
                                                        None, C.c_float(0.5), B["xc"], B["work"], stream),
            s.nel * 27 * 4 + n * (4 * 4 + 1) + nc * 4, True),
        "k_applyK3_col<double,double,EPI_Y> + k_gather_tiles (FP64 refinement residual product)": (
            lambda: lib.ust_k_apply_K_f64(C.byref(g), WK, mask, B["u64"], B["y64"], B["work"], stream),
            s.nel * 27 * 8 + n * (2 * 8 + 1), False),
    }
    kernels = {}
    for name, (fn, nbytes, per_iteration) in runs.items():
        ms = time_on_stream(lib, stream, fn, reps)
        achieved = nbytes / (ms * 1e-3) / 1e9
        kernels[name] = {"avg_launch_us": round(ms * 1e3, 2), "algorithmic_bytes_per_launch": int(nbytes),
                         "achieved_GBps": round(achieved, 1), "frac": round(achieved / HBM_PEAK_GBS, 4),
                         "every_inner_iteration": per_iteration}
    for p in bufs.values():
        lib.ust_dev_free(p)
    top = max((k for k in kernels if kernels[k]["every_inner_iteration"]), key=lambda k: kernels[k]["avg_launch_us"])
    k = kernels[top]
    roofline = {"bound": "hbm", "achieved": k["achieved_GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": k["frac"], "traffic": None, "kernel": top, "avg_launch_us": k["avg_launch_us"],
                "algorithmic_bytes_per_launch": k["algorithmic_bytes_per_launch"]}
    return roofline, kernels


def measured_traffic(kernel, workload_tag):
    """HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, calibrated on a plain copy) of a kernel from the committed
    counter summary profiles/operator_traffic_<workload_tag>.json, or None when it has not been measured."""
    path = os.path.join(ROOT, "profiles", f"operator_traffic_{workload_tag}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get(kernel, {}).get("hbm_bytes_per_launch")


def kernel_only(args):
    """Runs only the fine-level operator passes (50 launches each) and a 256 MiB device copy (20 launches) for
    counter collection under rocprofv3 --pmc."""
    s, _, _ = build_workload(args)
    s.assemble(gravity=1.0)
    lib, ctx = s.lib, s.ctx
    stream = lib.ust_get_stream(ctx)
    r, kern = fine_level_kernels(lib, ctx, s, stream)
    # element stiffness matrices for the per-element callbacks (FP64 MFMA), 65536 elements per launch
    ne = min(65536, s.nel)
    ke = C.c_void_p()
    assert lib.ust_dev_malloc(C.byref(ke), ne * s.nd * s.nd * 8) == 0
    for first in (0, ne, 2 * ne):
        if first + ne <= s.nel:
            assert lib.ust_get_element_matrices_dev(ctx, first, ne, ke) == 0
    lib.ust_stream_sync(stream)
    lib.ust_dev_free(ke)
    nb = 256 << 20
    a, b = C.c_void_p(), C.c_void_p()
    lib.ust_dev_malloc(C.byref(a), nb)
    lib.ust_dev_malloc(C.byref(b), nb)
    for _ in range(20):
        lib.ust_memcpy_d2d(b, a, nb, stream)
    lib.ust_stream_sync(stream)
    print(json.dumps({"kernel_only": kern, "copy_bytes": nb}), flush=True)


def build_workload(args):
    """The benchmark's system and a description of it.  grf (default): the random-field configuration on
    held-out field seed 5000 (the learned scaling is trained on seeds 1000+); spheres: 96^3 x 30."""
    from earthbyte_underworld_amd.configs import random_field_3d, spheres_3d
    if args.workload == "grf":
        n = args.size or 128
        ppc = args.ppc or 20
        seed = 5000 if args.seed < 0 else args.seed
        log(f"building {n}x{n}x{n // 2} x {ppc} ppc random-field workload (seed {seed})")
        s = random_field_3d(nel=(n, n, n // 2), ppc=ppc, seed=seed)
        desc = (f"3D Gaussian random field: {n}x{n}x{n // 2} Q1/P0 on [0,2]x[0,2]x[0,1], {ppc} particles/cell, "
                f"log10 eta in [-3,3] (contrast 1e6) per particle, correlation length 0.1, field seed {seed}, "
                f"free-slip, BSSCR to 1e-6 relative Stokes residual")
    else:
        n = args.size or 96
        ppc = args.ppc or 30
        seed = 2024 if args.seed < 0 else args.seed
        log(f"building {n}^3 x {ppc} ppc spheres workload")
        s = spheres_3d(nel=(n, n, n), ppc=ppc, nsph=args.nsph, seed=seed)
        desc = (f"3D spheres: {n}^3 Q1/P0, {ppc} particles/cell, {args.nsph} spheres log10 eta U[-2,4], "
                f"free-slip, BSSCR to 1e-6 relative Stokes residual")
    args.ppc, args.seed = ppc, seed
    return s, desc, n


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["grf", "spheres"], default="grf",
                    help="grf: 128x128x64 Gaussian random field (the metric's configuration); spheres: 96^3")
    ap.add_argument("--size", type=int, default=0, help="elements per axis (grf: n x n x n/2); 0: the default")
    ap.add_argument("--ppc", type=int, default=0, help="particles per cell; 0: 20 (grf) / 30 (spheres)")
    ap.add_argument("--nsph", type=int, default=40)
    ap.add_argument("--seed", type=int, default=-1, help="field / sphere-set seed; -1: 5000 (grf) / 2024 (spheres)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="plain FP64 solve only (profiling runs)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mlp", default=MLP_WEIGHTS)
    ap.add_argument("--relax", type=int, default=1, help="relaxed inner tolerance on the mixed+learned path")
    ap.add_argument("--kernel-only", action="store_true", help="dominant kernel + copy only (counter runs)")
    ap.add_argument("--uzawa", action="store_true", help="also time the Uzawa solver (reference defaults)")
    ap.add_argument("--advect", action="store_true", help="also time one swarm advection step + re-sort")
    ap.add_argument("--dvc", action="store_true", help="also time DVC weights (10^3 sub-cells) on the swarm")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU path is design-only.",
              file=sys.stderr)
        sys.exit(2)

    if args.kernel_only:
        kernel_only(args)
        return
    import earthbyte_underworld_amd as ust
    from earthbyte_underworld_amd import _lib as L
    s, workload, n = build_workload(args)
    lib, ctx = s.lib, s.ctx
    stream = lib.ust_get_stream(ctx)
    learned = os.path.exists(args.mlp)
    if learned:
        d = np.load(args.mlp)
        s.mlp_set_params(d["w"], d["b"])
        log("learned Schur scaling loaded from", os.path.relpath(args.mlp, ROOT))
    else:
        log("no trained MLP found: the mixed run uses the plain approxS scaling")
    # headline path: Chebyshev degree 1 on the finest level (cheaper V-cycles outweigh the extra iterations
    # there), degree 2 on the coarser levels (short passes whose extra smoothing cuts the inner iterations);
    # the plain path keeps the default degree 2 everywhere.  Both use the default Chebyshev interval
    # [lambda_max / 12, 1.1 lambda_max] (128x128x64: mixed 0.459 s against 0.485 s at 30, flat from 8 to 15;
    # plain 1.686 s against 1.684 s at 30).  The velocity PCG itself runs in FP32 (FP32 K weights on every
    # level, FP32 Krylov vectors) inside FP64 iterative refinement (0.547 -> 0.487 s at ratio 30 against the
    # FP64 PCG around the FP32 V-cycle).
    opts_mixed = solve_opts(ust, L.PREC_MIXED, learned, args.relax, degree=1, coarse_degree=2, cheby_ratio=12.0,
                            relax_cap=0.1, inner_fp32=1, vcycle_bf16=0)
    opts_plain = solve_opts(ust, L.PREC_FP64, False, cheby_ratio=12.0)

    def step(opts):
        s.assemble(gravity=1.0)
        return s.solve(opts)

    if args.no_mixed:
        args.warmup, args.steps = 0, 0
    for i in range(args.warmup):
        st = step(opts_mixed)
        log(f"warmup {i}: {st['outer_its']} outer its, residual {st['stokes_residual']:.2e}")
    lib.ust_stream_sync(stream)
    t0 = time.perf_counter()
    stats = []
    for i in range(args.steps):
        ts = time.perf_counter()
        stats.append(step(opts_mixed))
        log(f"step {i}: {time.perf_counter() - ts:.2f}s, {stats[-1]['outer_its']} outer / {stats[-1]['inner_its']} "
            f"inner its, residual {stats[-1]['stokes_residual']:.2e}, fallback {stats[-1]['fallback_used']}")
    lib.ust_stream_sync(stream)
    t_mixed = (time.perf_counter() - t0) / args.steps if args.steps else None
    worst_res = max(st["stokes_residual"] for st in stats) if stats else None
    sol_mixed = s.solution() if stats else None

    plain = None
    if not args.no_plain:
        lib.ust_stream_sync(stream)
        t0 = time.perf_counter()
        stp = step(opts_plain)
        lib.ust_stream_sync(stream)
        tp = time.perf_counter() - t0
        plain = {"solves_per_hour": 3600.0 / tp, "s_per_solve": tp, "outer_its": stp["outer_its"],
                 "inner_its": stp["inner_its"], "stokes_residual": stp["stokes_residual"]}
        if sol_mixed is not None:  # the headline's solution against the FP64 one (relative L2)
            up, pp = s.solution()
            plain["mixed_velocity_rel_err"] = float(np.linalg.norm(sol_mixed[0] - up) / np.linalg.norm(up))
            plain["mixed_pressure_rel_err"] = float(np.linalg.norm(sol_mixed[1] - pp) / np.linalg.norm(pp))
        log(f"plain FP64: {tp:.2f}s, {stp['outer_its']} outer its, residual {stp['stokes_residual']:.2e}")

    uzawa = None
    if args.uzawa:
        # Stokes_SLE_UzawaSolver with the reference's defaults (tolerance 1e-5 on the relative Uzawa residual,
        # approxS preconditioner), from p = 0, velocity solves in FP64 and in the mixed precision path
        uzawa = {}
        for name, vo in (("fp64", opts_plain), ("mixed", opts_mixed)):
            s.assemble(gravity=1.0)
            s.set_solution(np.zeros(s.ndof), np.zeros(s.nel))
            lib.ust_stream_sync(stream)
            t0 = time.perf_counter()
            stu = s.solve_uzawa(tolerance=1e-5, opts=vo)
            lib.ust_stream_sync(stream)
            tu = time.perf_counter() - t0
            uzawa[name] = {"s_per_solve": round(tu, 3), "solves_per_hour": round(3600.0 / tu, 2),
                           "uzawa_its": stu["outer_its"], "inner_its": stu["inner_its"],
                           "stokes_residual": stu["stokes_residual"]}
            log(f"uzawa ({name} velocity solves): {tu:.2f}s, {stu['outer_its']} its, "
                f"residual {stu['stokes_residual']:.2e}")

    # ---- assembly throughput (particle kernel + derived element terms) ----
    t_asm = time_on_stream(lib, stream, lambda: lib.ust_assemble(ctx, C.c_double(1.0)), 3)
    np_ = lib.ust_swarm_count(ctx)
    asm_bytes = np_ * (3 * 8 + 8 + 1 + 8) + s.nel * (27 + 8 + 8) * 8
    roofline, kernels = fine_level_kernels(lib, ctx, s, stream)
    roofline["traffic"] = measured_traffic(roofline["kernel"], f"{args.workload}{n}")

    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        try:
            if args.workload != "grf":
                raise RuntimeError("the CPU baseline samples the random-field configuration only")
            if plain is None:
                raise RuntimeError("the CPU baseline needs the plain FP64 run's outer iteration count")
            cpu = cpu_baseline(s.nel, args.ppc, args.seed, plain["outer_its"], threads)
            log("cpu baseline", cpu["sample"])
        except Exception as ex:  # the CPU leg never blocks the GPU result
            log("cpu baseline failed:", ex)

    advect = None
    if args.advect:
        # one SwarmAdvector step (midpoint rule) through the last solution, CFL 0.5, and the re-sort (last:
        # nothing after it uses the benchmark swarm)
        u, _ = s.solution()
        dt = 0.5 / n / max(float(np.max(np.abs(u))), 1e-300)
        npart = lib.ust_swarm_count(ctx)
        lib.ust_stream_sync(stream)
        t0 = time.perf_counter()
        nr, ne = s.advect(dt, order=2, allowFallbackToFirstOrder=True)
        lib.ust_stream_sync(stream)
        ta = time.perf_counter() - t0
        advect = {"ms": round(ta * 1e3, 2), "particles": int(npart), "removed": nr, "empty_cells": ne,
                  "particles_per_s": round(npart / ta, 1)}
        log(f"advect: {ta * 1e3:.1f} ms for {npart} particles, removed {nr}, empty cells {ne}")

    dvc = None
    if args.dvc:
        # DVCWeights on the benchmark swarm (last: it moves the points to their region centroids)
        npart = lib.ust_swarm_count(ctx)
        lib.ust_stream_sync(stream)
        t0 = time.perf_counter()
        try:
            s.dvc_weights(10)
            lib.ust_stream_sync(stream)
            td = time.perf_counter() - t0
            dvc = {"ms": round(td * 1e3, 2), "particles": int(npart), "resolution": 10,
                   "particles_per_s": round(npart / td, 1)}
            log(f"dvc weights: {td * 1e3:.1f} ms for {npart} particles")
        except Exception as ex:  # e.g. an element emptied by the advection leg
            log("dvc weights failed:", ex)

    if t_mixed is None:  # --no-mixed: the plain solve alone (profiling)
        log("plain FP64 only:", json.dumps(plain))
        return
    value = 3600.0 / t_mixed
    out = {
        "metric": "3D Stokes solves/hour to 1e-6 residual, mixed+learned vs plain; assembly GB/s",
        "value": round(value, 2),
        "unit": "solves/hour",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_mixed * 1e3, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 velocity PCG and V-cycle (f32 K weights) inside f64 iterative refinement; f64 outer CG, "
                 "residuals, coarse solve",
        "data": "synthetic (seeded swarm and viscosity field)",
        "config": {"workload": workload,
                   "elements": s.nel, "particles": int(np_), "velocity_dofs": s.ndof,
                   "learned_schur_pc": learned, "inner_tolerance_relaxed": bool(args.relax),
                   "inner_pcg": {"mixed": "f32 with f64 iterative refinement", "plain": "f64"},
                   "smoother_degree": {"mixed": {"finest": 1, "coarser": 2}, "plain": 2},
                   "chebyshev_ratio": {"mixed": 12.0, "plain": 12.0}},
        "stokes_residual": worst_res,
        "outer_its": stats[-1]["outer_its"],
        "inner_its": stats[-1]["inner_its"],
        "plain_fp64": plain,
        "speedup_vs_plain": round(plain["s_per_solve"] / t_mixed, 3) if plain else None,
        "assembly_GBps": round(asm_bytes / (t_asm * 1e-3) / 1e9, 1),
        "assembly_ms": round(t_asm, 3),
        "roofline": roofline,
        "fine_level_kernels": kernels,
        "cpu_baseline": cpu,
        "uzawa": uzawa,
        "advect": advect,
        "dvc": dvc,
        "vs_cpu": round(value / cpu["value"], 2) if cpu else None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
