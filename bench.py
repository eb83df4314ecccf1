"""Benchmark: 3D PIC Stokes solves/hour to 1e-6 residual (mixed precision + learned Schur scaling vs plain
FP64), plus assembly GB/s, on one MI355X.

Workload: 96^3 Q1/P0 elements on the unit cube, 30 particles/cell (26.5M particles, seeded MINSTD layout),
background eta 1 / rho 0 with 40 seeded spheres (radius U[0.03,0.08], log10 eta U[-2,4], rho 1), free-slip
walls, gravity 1.  One step = particle assembly (fused rheology) + a BSSCR solve to a relative FP64 Stokes
residual <= 1e-6, inputs (swarm, materials, BCs) resident in HBM before the timed region.

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


def solve_opts(ust, precision, learned, relax=0, degree=2, coarse_degree=0, cheby_ratio=12.0):
    from earthbyte_underworld_amd import _lib as L
    # stokes_rtol: the metric's 1e-6 bar on the FP64 Stokes residual is enforced, not assumed
    return ust.default_opts(precision=precision, schur_pc=L.SCHUR_LEARNED if learned else L.SCHUR_UW,
                            outer_rtol=1e-7, inner_rtol=1e-8, max_outer=2000, inner_relax=relax,
                            smoother_degree=degree, coarse_degree=coarse_degree, cheby_ratio=cheby_ratio,
                            stokes_rtol=1e-6)


def cpu_baseline(size, ppc, nsph, seed, cpu_size, threads):
    """The FP64 oracle (same algorithm: particle assembly + BSSCR with GMG) on a cpu_size^3 sample of the
    same configuration, scaled to the full element count."""
    from oracle import oracle as O
    from earthbyte_underworld_amd.configs import sphere_set
    O.set_threads(threads)
    n = cpu_size
    m = O.make_mesh(3, (n, n, n))
    ne = n ** 3
    t0 = time.time()
    coords = O.layout_random(m, ppc, 13)
    off = np.arange(0, ne * ppc + 1, ppc).astype(np.int32)
    xi, w = O.map_particles(m, coords, off)
    centres, radii, logs = sphere_set(nsph, seed)
    mat = np.zeros(len(w), np.int32)
    for k in range(nsph):
        mat[np.sum((coords - centres[k]) ** 2, axis=1) <= radii[k] ** 2] = k + 1
    mats = [O.make_material([(O.RH_MATERIALVISCOSITY, (1.0,))], 0.0)] + \
           [O.make_material([(O.RH_MATERIALVISCOSITY, (10.0 ** logs[k],))], 1.0) for k in range(nsph)]
    t_setup = time.time() - t0
    t0 = time.time()
    eta, _ = O.eval_viscosity(m, mats, off, xi, mat)
    rho = np.array([mats[k].density for k in mat])
    Ke, Ge, fe, aS = O.assemble_elements(m, off, xi, w, eta, rho=rho, gravity=1.0)
    nn = (n + 1) ** 3
    idx = np.arange(nn)
    ijk = [idx % (n + 1), (idx // (n + 1)) % (n + 1), idx // (n + 1) ** 2]
    mask = np.zeros((nn, 3), np.uint8)
    for a in range(3):
        mask[(ijk[a] == 0) | (ijk[a] == n), a] = 1
    gs = O.GlobalSystem(m, mask.reshape(-1), np.zeros(nn * 3), Ke, Ge, fe)
    t_asm = time.time() - t0
    t0 = time.time()
    u, p, st = gs.solve(aS, inner_rtol=1e-8, outer_rtol=1e-7, smooth_its=1)
    t_solve = time.time() - t0
    scale = (size / cpu_size) ** 3
    t_full = (t_asm + t_solve) * scale
    return {"value": 3600.0 / t_full, "unit": "solves/hour", "cores": threads, "kind": "port",
            "sample": f"FP64 C oracle (OpenMP {threads} threads): particle assembly + BSSCR/GMG solve of the same "
                      f"configuration at {cpu_size}^3 elements ({t_asm:.2f}s assembly + {t_solve:.2f}s solve, "
                      f"{st.outer_its} outer / {st.inner_its_total} inner its, residual "
                      f"{st.stokes_rel_residual:.1e}), scaled by the element-count ratio {scale:.0f} "
                      f"(iteration growth with resolution ignored: an upper bound on the CPU rate)"}


def dominant_kernel(lib, ctx, s, stream, reps=50):
    """The fused FP32 Jacobi-Chebyshev smoothing pass of the inner solves' V-cycle on the finest grid
    (k_applyK3_tiled<float, EPI_CHEBY> + k_gather_tiles): its average time over reps launches on the library
    stream (HIP events) and its algorithmic HBM bytes."""
    from earthbyte_underworld_amd import _lib as L
    g = L.Grid()
    WK, WG, WF, mask = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
    lib.ust_get_device_views(ctx, C.byref(g), C.byref(WK), C.byref(WG), C.byref(WF), C.byref(mask))
    wl = lib.ust_k_weights_len(C.byref(g))
    n = s.ndof
    bufs = {}
    for name, nbytes in (("W", wl * 4), ("x", n * 4), ("b", n * 4), ("invd", n * 4), ("d", n * 4),
                         ("work", lib.ust_k_apply_work_len(C.byref(g)) * 4)):
        bufs[name] = C.c_void_p()
        assert lib.ust_dev_malloc(C.byref(bufs[name]), nbytes) == 0
    lib.ust_k_copy_f64_to_f32(WK, bufs["W"], wl, stream)
    rng = np.random.default_rng(0)
    for name in ("x", "b", "d"):
        a = rng.normal(size=n).astype(np.float32)
        lib.ust_memcpy_h2d(bufs[name], a.ctypes.data_as(C.c_void_p), n * 4, stream)
    a = np.full(n, 1e-3, np.float32)  # a contracting step keeps x bounded over the repetitions
    lib.ust_memcpy_h2d(bufs["invd"], a.ctypes.data_as(C.c_void_p), n * 4, stream)
    ms = time_on_stream(lib, stream, lambda: lib.ust_k_apply_K_cheby_f32(
        C.byref(g), bufs["W"], mask, bufs["x"], bufs["b"], bufs["invd"], bufs["d"], C.c_float(0.1),
        C.c_float(0.5), 0, bufs["work"], stream), reps)
    for p in bufs.values():
        lib.ust_dev_free(p)
    # weights 27 x 4 B per element; per dof: x read + write, b, invd, d read + write (4 B each), mask 1 B
    nbytes = s.nel * 27 * 4 + n * (6 * 4 + 1)
    achieved = nbytes / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "k_applyK3_tiled<float,EPI_CHEBY> + k_gather_tiles (fused FP32 K u + Chebyshev step)",
            "avg_launch_us": round(ms * 1e3, 2), "algorithmic_bytes_per_launch": int(nbytes)}


def measured_traffic():
    """HBM bytes per launch of the dominant kernel pair from the committed rocprofv3 counter summary
    (FETCH_SIZE + WRITE_SIZE, calibrated on a plain copy), or None when it has not been measured."""
    path = os.path.join(ROOT, "profiles", "k_applyK3_tiled_f32_cheby_96cube_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def kernel_only(args):
    """Runs only the dominant kernel (50 launches) and a 256 MiB device copy (20 launches) for counter
    collection under rocprofv3 --pmc."""
    from earthbyte_underworld_amd.configs import spheres_3d
    n = args.size
    s = spheres_3d(nel=(n, n, n), ppc=args.ppc, nsph=args.nsph, seed=args.seed)
    s.assemble(gravity=1.0)
    lib, ctx = s.lib, s.ctx
    stream = lib.ust_get_stream(ctx)
    r = dominant_kernel(lib, ctx, s, stream)
    nb = 256 << 20
    a, b = C.c_void_p(), C.c_void_p()
    lib.ust_dev_malloc(C.byref(a), nb)
    lib.ust_dev_malloc(C.byref(b), nb)
    for _ in range(20):
        lib.ust_memcpy_d2d(b, a, nb, stream)
    lib.ust_stream_sync(stream)
    print(json.dumps({"kernel_only": r, "copy_bytes": nb}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=96)
    ap.add_argument("--ppc", type=int, default=30)
    ap.add_argument("--nsph", type=int, default=40)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="plain FP64 solve only (profiling runs)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-size", type=int, default=16)
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
    from earthbyte_underworld_amd.configs import spheres_3d

    n = args.size
    log(f"building {n}^3 x {args.ppc} ppc spheres workload")
    s = spheres_3d(nel=(n, n, n), ppc=args.ppc, nsph=args.nsph, seed=args.seed)
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
    # the plain path keeps the default degree 2 everywhere
    # Chebyshev interval [lambda_max / 30, 1.1 lambda_max]: a wider interval than the default 12 suits the
    # degree-2 coarse smoothing here (5,217 -> 5,048 inner iterations at 96^3; flat from 30 to 150)
    opts_mixed = solve_opts(ust, L.PREC_MIXED, learned, args.relax, degree=1, coarse_degree=2, cheby_ratio=30.0)
    opts_plain = solve_opts(ust, L.PREC_FP64, False, cheby_ratio=30.0)  # the same interval (20.2 -> 19.7 s)

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

    plain = None
    if not args.no_plain:
        lib.ust_stream_sync(stream)
        t0 = time.perf_counter()
        stp = step(opts_plain)
        lib.ust_stream_sync(stream)
        tp = time.perf_counter() - t0
        plain = {"solves_per_hour": 3600.0 / tp, "s_per_solve": tp, "outer_its": stp["outer_its"],
                 "inner_its": stp["inner_its"], "stokes_residual": stp["stokes_residual"]}
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
    roofline = dominant_kernel(lib, ctx, s, stream)
    roofline["traffic"] = measured_traffic()

    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        try:
            subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
            cpu = cpu_baseline(n, args.ppc, args.nsph, args.seed, args.cpu_size, threads)
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
        "dtype": "f32 V-cycle inside f64 inner PCG; f64 outer CG, residuals, coarse solve",
        "data": "synthetic (seeded swarm and sphere set)",
        "config": {"workload": f"3D spheres: {n}^3 Q1/P0, {args.ppc} particles/cell, {args.nsph} spheres "
                               f"log10 eta U[-2,4], free-slip, BSSCR to 1e-6 relative Stokes residual",
                   "elements": s.nel, "particles": int(np_), "velocity_dofs": s.ndof,
                   "learned_schur_pc": learned, "inner_tolerance_relaxed": bool(args.relax),
                   "smoother_degree": {"mixed": {"finest": 1, "coarser": 2}, "plain": 2},
                   "chebyshev_ratio": {"mixed": 30.0, "plain": 30.0}},
        "stokes_residual": worst_res,
        "outer_its": stats[-1]["outer_its"],
        "inner_its": stats[-1]["inner_its"],
        "plain_fp64": plain,
        "speedup_vs_plain": round(plain["s_per_solve"] / t_mixed, 3) if plain else None,
        "assembly_GBps": round(asm_bytes / (t_asm * 1e-3) / 1e9, 1),
        "assembly_ms": round(t_asm, 3),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "uzawa": uzawa,
        "advect": advect,
        "dvc": dvc,
        "vs_cpu": round(value / cpu["value"], 2) if cpu else None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
