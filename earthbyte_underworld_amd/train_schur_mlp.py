"""Train the learned Schur-preconditioner scaling on the GPU from solves of the plain method.

For every training field the plain BSSCR operators are assembled and the diagonal of the Schur complement
S = G^T K^-1 G is probed with the 3^dim colouring (one A11 multigrid-CG solve per colour).  The target of a
pressure cell is log10(diag(S)_e / approxS_e), approxS being the "uw" diagonal g^T diag(K_e)^-1 g of
UzawaPreconditionerTerm.c:195-212.  The network (27 -> 64 -> 64 -> 1, BF16 weights) is then trained with
Adam by the library's own forward/backward kernels, and evaluated on held-out seeds by the outer BSSCR
iteration count against the plain approxS preconditioner.

    python -m earthbyte_underworld_amd.train_schur_mlp --size 32 --train 24 --heldout 4

writes earthbyte_underworld_amd/data/schur_mlp_3d.npz (w, b, and a JSON training record).  The protocol of the
random-field benchmark (200 training fields and 50 held-out fields at 128 x 128 x 64, 20 particles per cell,
log10 eta sampled per particle):

    python -m earthbyte_underworld_amd.train_schur_mlp --kinds grf --size 128 --ppc 20 --train 200 \
        --heldout 50 --sample-cells 65536 --eval-path headline

With --eval-path headline the held-out fields are solved on the benchmark's path (mixed precision, relaxed inner
tolerance, Chebyshev degree 1/2, ratio 30, Stokes bar 1e-6) with and without the learned scaling.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if os.path.dirname(HERE) not in sys.path:
    sys.path.insert(0, os.path.dirname(HERE))

from earthbyte_underworld_amd import _lib as L  # noqa: E402
from earthbyte_underworld_amd import configs  # noqa: E402
from earthbyte_underworld_amd.stokes import default_opts  # noqa: E402

DEFAULT_OUT = os.path.join(HERE, "data", "schur_mlp_3d.npz")


def field(kind, n, seed, ppc, grf_box=(2.0, 2.0, 1.0)):
    """kind 'spheres': 40 seeded spheres (log10 eta U[-2,4]) in the unit cube at n^3; 'grf': Gaussian random
    field (correlation length 0.1) mapped to [-3,3] on n x n x n/2 elements over grf_box."""
    if kind == "spheres":
        return configs.spheres_3d(nel=(n, n, n), ppc=ppc, nsph=40, seed=seed)
    return configs.random_field_3d(nel=(n, n, n // 2), ppc=ppc, seed=seed, box=tuple(grf_box))


def probe(s, inner_rtol, precision=L.PREC_FP64):
    s.assemble()
    o = default_opts(precision=precision, inner_rtol=inner_rtol)
    return s.mlp_features(), s.mlp_probe_target(o)


def outer_its(s, learned, rtol=1e-7, path="fp64"):
    if path == "headline":
        o = default_opts(precision=L.PREC_MIXED, schur_pc=L.SCHUR_LEARNED if learned else L.SCHUR_UW,
                         outer_rtol=rtol, inner_rtol=1e-8, max_outer=2000, inner_relax=1, smoother_degree=1,
                         coarse_degree=2, cheby_ratio=30.0, stokes_rtol=1e-6)
    else:
        o = default_opts(precision=L.PREC_FP64, schur_pc=L.SCHUR_LEARNED if learned else L.SCHUR_UW,
                         outer_rtol=rtol, inner_rtol=1e-8, max_outer=3000, stall_window=0)
    t = time.time()
    st = s.solve(o)
    st["seconds"] = time.time() - t
    return st


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--size", type=int, default=32)
    ap.add_argument("--grf-size", type=int, default=0, help="random-field grid n (n x n x n/2); 0: --size")
    ap.add_argument("--grf-box", type=float, nargs=3, default=(2.0, 2.0, 1.0))
    ap.add_argument("--ppc", type=int, default=12)
    ap.add_argument("--train", type=int, default=24, help="training fields (half spheres, half random fields)")
    ap.add_argument("--heldout", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--batch", type=int, default=2048)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--inner-rtol", type=float, default=1e-7)
    ap.add_argument("--kinds", default="spheres,grf", help="training field kinds, alternating")
    ap.add_argument("--sample-cells", type=int, default=0, help="cells kept per training field (0: all)")
    ap.add_argument("--eval-path", choices=["fp64", "headline"], default="fp64")
    ap.add_argument("--probe-precision", choices=["fp64", "mixed"], default="fp64",
                    help="A11 solves of the Schur-diagonal probe (mixed: FP32 V-cycle inside FP64 PCG)")
    ap.add_argument("--out", default=DEFAULT_OUT)
    a = ap.parse_args(argv)

    t0 = time.time()
    X, Y = [], []
    kinds = a.kinds.split(",")
    rng = np.random.default_rng(a.seed)
    for k in range(a.train):
        kind = kinds[k % len(kinds)]
        s = field(kind, (a.grf_size or a.size) if kind == "grf" else a.size, 1000 + k, a.ppc, a.grf_box)
        f, t = probe(s, a.inner_rtol, L.PREC_MIXED if a.probe_precision == "mixed" else L.PREC_FP64)
        s.close()
        if a.sample_cells and a.sample_cells < len(t):
            keep = rng.choice(len(t), a.sample_cells, replace=False)
            f, t = f[keep], t[keep]
        X.append(f)
        Y.append(t)
        print(f"[train] field {k} ({kind}): target mean {t.mean():+.3f} std {t.std():.3f} "
              f"({time.time() - t0:.1f}s)", flush=True)
    X = np.concatenate(X)
    Y = np.concatenate(Y)
    t_probe = time.time() - t0

    s = field(kinds[0], 16, 1, 4)
    s.mlp_init(seed=a.seed)
    t1 = time.time()
    losses = []
    lr = a.lr
    for stage in range(3):
        losses.append(s.mlp_train(X, Y, epochs=a.epochs // 3, batch=a.batch, lr=lr, seed=a.seed + stage))
        print(f"[train] stage {stage}: lr {lr:.1e} loss {losses[-1]:.5f} (target variance {Y.var():.5f})",
              flush=True)
        lr *= 0.3
    t_train = time.time() - t1
    w, b = s.mlp_params()
    s.close()

    # held-out evaluation: the outer iteration count of plain vs learned BSSCR
    evals = []
    for k in range(a.heldout):
        kind = kinds[k % len(kinds)]
        s = field(kind, (a.grf_size or a.size) if kind == "grf" else a.size, 5000 + k, a.ppc, a.grf_box)
        s.assemble()
        s.mlp_set_params(w, b)
        plain = outer_its(s, False, path=a.eval_path)
        s.set_solution(np.zeros(s.ndof), np.zeros(s.nel))
        learned = outer_its(s, True, path=a.eval_path)
        evals.append({"kind": kind, "seed": 5000 + k, "plain_outer": plain["outer_its"],
                      "learned_outer": learned["outer_its"], "plain_inner": plain["inner_its"],
                      "learned_inner": learned["inner_its"], "plain_residual": plain["stokes_residual"],
                      "learned_residual": learned["stokes_residual"], "plain_s": round(plain["seconds"], 3),
                      "learned_s": round(learned["seconds"], 3), "fallback": learned.get("fallback_used", 0)})
        print(f"[eval] {kind} seed {5000 + k}: outer its plain {plain['outer_its']} learned "
              f"{learned['outer_its']} ({plain['seconds']:.2f} s / {learned['seconds']:.2f} s)", flush=True)
        s.close()
    record = {"size": a.size, "grf_size": a.grf_size or a.size, "grf_box": list(a.grf_box), "ppc": a.ppc,
              "kinds": kinds, "train_fields": a.train, "sample_cells": a.sample_cells, "eval_path": a.eval_path,
              "probe_precision": a.probe_precision, "inner_rtol": a.inner_rtol,
              "samples": int(len(Y)),
              "epochs": a.epochs, "batch": a.batch, "lr": a.lr, "final_loss": float(losses[-1]),
              "target_variance": float(Y.var()), "probe_seconds": round(t_probe, 1),
              "train_seconds": round(t_train, 1), "heldout": evals}
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    np.savez(a.out, w=w.astype(np.float32), b=b.astype(np.float32), record=np.array(json.dumps(record)))
    print(json.dumps(record), flush=True)
    return record


if __name__ == "__main__":
    main()
