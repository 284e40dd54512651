"""Writes tests/golden/precision_64cube_oracle.json: the oracle's FP64 FGMRES(30)
solution of the precision form's system (3-D 27-pt convection-diffusion 64^3,
bs = 4, b = isai_gen_normal(seed 1), ISAI preconditioner with FP64 factors, 1e-10),
kept as 2048 sampled entries plus 8 projections onto seeded N(0,1) vectors.

The full solution (1M doubles) is too large for a fixture and the oracle takes
about 90 s on 8 host cores, so the GPU tests and the benchmark compare against
these samples.  Run: python tests/golden/make_precision_reference.py
"""
import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parent.parent))

import _oracle as O  # noqa: E402
from isai_kernel_amd import problems  # noqa: E402

SAMPLE_SEED, NSAMPLES, NPROJ = 12345, 2048, 8


def sample_index(n):
    return np.sort(np.random.default_rng(SAMPLE_SEED).choice(n, NSAMPLES, replace=False))


def projections(x):
    rng = np.random.default_rng(SAMPLE_SEED + 1)
    return [float(np.dot(rng.standard_normal(len(x)), x)) for _ in range(NPROJ)]


def main():
    rp, ci, val = problems.generate(2, 64, seed=0)
    n = 4 * (len(rp) - 1)
    b = problems.normal(1, n)
    _, L, U = O.isai_factors(rp, ci, val)
    P = O.make_precond(1, len(rp) - 1, 4, L=L, U=U)
    t0 = time.perf_counter()
    x, st, it, rr = O.fgmres(rp, ci, val, P, b, m=30, tol=1e-10, maxit=3000)
    secs = time.perf_counter() - t0
    idx = sample_index(n)
    out = {
        "description": "Oracle FP64 FGMRES(30) with FP64 ISAI on problem 2 (3-D 27-pt convection-diffusion "
                       "64^3, bs=4, seed 0), b = isai_gen_normal(1, n), tol 1e-10; sampled entries and "
                       "projections of the solution (written by tests/golden/make_precision_reference.py)",
        "problem": 2, "L": 64, "seed": 0, "b_seed": 1, "tol": 1e-10, "restart": 30,
        "status": int(st), "iters": int(it), "relres": float(rr), "oracle_seconds": round(secs, 1),
        "threads": O.num_threads(),
        "n": int(n), "sample_seed": SAMPLE_SEED,
        "x_norm_inf": float(np.max(np.abs(x))), "x_norm2": float(np.linalg.norm(x)),
        "sample_index": idx.tolist(), "x_samples": [float(v) for v in x[idx]],
        "projections": projections(x),
    }
    (HERE / "precision_64cube_oracle.json").write_text(json.dumps(out, indent=0))
    print(f"oracle FGMRES(30): status {st}, {it} iterations, relres {rr:.3e}, {secs:.1f} s")


if __name__ == "__main__":
    main()
