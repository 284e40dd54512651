#!/usr/bin/env python3
"""Benchmark of the block-sparse ILU(0)/ISAI preconditioning path on one MI355X.

Workload (N=1): 3-D 7-point stencil on a 96^3 grid, 4 dof per node (bs = 4,
884,736 block rows, 6,137,856 blocks), FP64, synthetic (generated in-process,
seed 0), b = A*1.  A "step" is one preconditioned solve to 1e-10 relative
residual (FP64, checked with the true residual): the mixed-precision form, an
FP64 FGMRES outer loop whose flexible preconditioner is an inner BiCGSTAB in FP32
storage (A rounded to FP32, BF16 ISAI factors, FP32 vectors, FP64 dot products).
The plain FP64 form (ISAI-preconditioned BiCGSTAB with FP64 factors), BiCGSTAB with
reduced-precision ISAI factors and exact ILU(0) applied with the sync-free
triangular solves are timed the same way for the speed-ups.

Kernel figures are hipEvent timings on the stream the kernels run on; the
roofline object is for the dominant kernel of the timed solve, the FP32 BSR
product with A and two fused dot products, with algorithmic bytes per launch
(16*4+4)*nnzb + 4*(nb+1) + 3*4*4*nb (FP32 values and x, y and the dot vector).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-forms]

The "forms" object adds the precision form (27-pt convection-diffusion 64^3,
FGMRES(30), FP64/FP32/BF16 ISAI), the batching form (256 uneven 9-pt systems with
1..16 right-hand sides, graph-replayed batched BiCGSTAB) and the differentiation
form (forward + adjoint solve and dL/dA on the 9-pt 256^2 stencil).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "solve time to 1e-10 rel-res, mixed vs FP64 ISAI; BSR SpMV/SpTRSV HBM GB/s"
HBM_PEAK_GBPS = 8000.0
TOL = 1e-10
# Bytes per launch of the FP32 A product (two dots, the handle's line-aligned FP32
# copy of A) at 96^3 beyond the L2 (HBM and Infinity Cache), from the FETCH_SIZE /
# WRITE_SIZE counter passes summarised in profiles/hbm_traffic_96cube_7pt_pmc.md
# (FETCH_SIZE x2 by the copy-kernel calibration): 228,772.9 KiB x 2 + 13,952.4 KiB.
F32_SPMV_TRAFFIC_BYTES_96CUBE = 228772.9 * 2 * 1024 + 13952.4 * 1024


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spmv_bytes(nb: int, nnzb: int, value_bytes: int = 8) -> int:
    return (16 * value_bytes + 4) * nnzb + 4 * (nb + 1) + 2 * 4 * 8 * nb


def apply_bytes(nb: int, nnzb_l: int, nnzb_u: int, value_bytes: int) -> int:
    """M_U (M_L v): both factors' values and indices once, v and the result FP64; the
    intermediate M_L v is written and gathered in FP64 with FP64 factors and in FP32
    (the precision the second product reads it in) with FP32/BF16 factors."""
    mid = 8 if value_bytes == 8 else 4
    return ((16 * value_bytes + 4) * (nnzb_l + nnzb_u) + 2 * 4 * (nb + 1) + 4 * 8 * nb
            + 2 * 4 * mid * nb + 4 * 8 * nb)


def f32_spmv_dot_bytes(nb: int, nnzb: int) -> int:
    """The mixed solve's A product: FP32 values + int32 indices once, the FP32 x
    gathered once, the FP32 result written and the FP32 dot vector read."""
    return (16 * 4 + 4) * nnzb + 4 * (nb + 1) + 3 * 4 * 4 * nb


def apply_bytes_baseline_formula(nb: int, nnzb_l: int, nnzb_u: int, value_bytes: int) -> int:
    """The ISAI-apply byte count as the benchmark targets state it: two products with
    FP64 input and output vectors, 2 x [(16 s_v + 4) nnzb_T + 4 (nb + 1) + 2*4*8*nb]."""
    return ((16 * value_bytes + 4) * (nnzb_l + nnzb_u) + 2 * 4 * (nb + 1) + 2 * 2 * 4 * 8 * nb)


def sptrsv_bytes(nb: int, nnzb_l: int, nnzb_u: int) -> int:
    return (16 * 8 + 4) * (nnzb_l + nnzb_u) + 2 * 4 * (nb + 1) + 4 * 4 * 8 * nb + 2 * 4 * nb


class Timer:
    """hipEvent pair on one stream."""

    def __init__(self, lib, stream):
        self.lib, self.stream = lib, stream
        self.e0, self.e1 = ctypes.c_void_p(), ctypes.c_void_p()
        lib.isai_event_create(ctypes.byref(self.e0))
        lib.isai_event_create(ctypes.byref(self.e1))

    def time(self, fn, iters: int) -> float:
        """Average milliseconds of fn() over iters calls."""
        from isai_kernel_amd import check
        check(self.lib.isai_stream_synchronize(self.stream), "sync")
        check(self.lib.isai_event_record(self.e0, self.stream), "event_record")
        for _ in range(iters):
            fn()
        check(self.lib.isai_event_record(self.e1, self.stream), "event_record")
        check(self.lib.isai_event_synchronize(self.e1), "event_sync")
        ms = ctypes.c_float()
        check(self.lib.isai_event_elapsed_ms(ctypes.byref(ms), self.e0, self.e1), "elapsed")
        return ms.value / iters


def cpu_baseline(rp, ci, val, b, budget_s: float = 40.0):
    """The oracle's plain FP64 method (ILU(0), ISAI(L/U), ISAI-preconditioned
    BiCGSTAB to 1e-10) on the full matrix, on the host cores."""
    sys.path.insert(0, str(ROOT / "tests"))
    import _oracle as O  # the checker, run here only as the CPU baseline

    nb = len(rp) - 1
    t0 = time.perf_counter()
    lu, st, _ = O.bsrilu0(rp, ci, val)
    mbL, mcL = O.bcsc_pattern(rp, ci, True)
    mbU, mcU = O.bcsc_pattern(rp, ci, False)
    pp, pl, ps = O.pairs(rp, ci, mbL, mcL)
    ML = O.isai_lower(lu, mbL, mcL, pp, pl, ps, O.build_dinv(rp, ci, lu, True))
    MU = O.isai_upper(rp, ci, lu, mbU, mcU, O.build_dinv(rp, ci, lu, False))
    tL = O.tri_bcsr_pattern(rp, ci, True)
    tU = O.tri_bcsr_pattern(rp, ci, False)
    _, mL = O.map_bcsc2bcsr(mbL, mcL, *tL)
    _, mU = O.map_bcsc2bcsr(mbU, mcU, *tU)
    L = (tL[0], tL[1], O.reorder(mL, ML))
    U = (tU[0], tU[1], O.reorder(mU, MU))
    t_setup = time.perf_counter() - t0
    P = O.make_precond(1, nb, 4, L=L, U=U)
    t1 = time.perf_counter()
    x, st, it, rr = O.bicgstab(rp, ci, val, P, b, tol=TOL, maxit=1000)
    t_solve = time.perf_counter() - t1
    return dict(setup_s=t_setup, solve_s=t_solve, iters=it, relres=rr, cores=O.num_threads(), status=st)


def precision_reference():
    """The oracle's FP64 FGMRES(30) solution of the precision form's system, as the
    sampled entries and projections that tests/golden/make_precision_reference.py
    recorded (the oracle itself takes about 90 s on 8 host cores)."""
    path = ROOT / "tests" / "golden" / "precision_64cube_oracle.json"
    return json.loads(path.read_text()) if path.exists() else None


def error_vs_reference(x, ref) -> dict:
    idx = np.asarray(ref["sample_index"])
    xs = np.asarray(ref["x_samples"])
    rng = np.random.default_rng(ref["sample_seed"] + 1)
    proj = [float(np.dot(rng.standard_normal(len(x)), x)) for _ in range(len(ref["projections"]))]
    pe = max(abs(a - b) / max(abs(b), 1e-300) for a, b in zip(proj, ref["projections"]))
    return {"max_rel_err_samples_vs_oracle": float(np.max(np.abs(x[idx] - xs)) / ref["x_norm_inf"]),
            "max_rel_err_projections_vs_oracle": float(pe)}


def bench_precision(timer, stream, steps: int) -> dict:
    """Precision form: 3-D 27-pt convection-diffusion 64^3 (bs = 4, 262,144 block rows),
    to 1e-10 with the ISAI factors stored in FP64 / FP32 / BF16: FP64 FGMRES(30) with
    the ISAI apply as its preconditioner, and the mixed form (FP64 FGMRES(30) outer
    loop around an inner FP32-storage BiCGSTAB, or around inner FP32-storage GMRES
    cycles of up to 40 steps).  Errors are against the oracle's
    FP64 FGMRES(30) solution (sampled, precision_reference)."""
    from isai_kernel_amd import DeviceArray, call, problems
    from isai_kernel_amd.solver import Handle, Matrix

    rp, ci, val = problems.generate(2, 64, seed=0)
    nb, nnzb = len(rp) - 1, len(ci)
    n = 4 * nb
    A = Matrix(rp, ci, val)
    d_b = DeviceArray.from_numpy(problems.normal(1, n))
    d_x = DeviceArray((n,), np.float64)
    h = Handle(A, precond="isai", precision="fp64", stream=stream)
    ref = precision_reference()
    out = {"workload": "3-D 27-pt convection-diffusion 64^3, bs=4, FGMRES(30) to 1e-10, b ~ N(0,1)",
           "nb": nb, "nnzb": nnzb,
           "oracle": ({"iters": ref["iters"], "relres": ref["relres"], "seconds": ref["oracle_seconds"],
                       "threads": ref["threads"]} if ref else None)}
    xs = {}
    legs = [("fp64", "fp64", "fgmres"), ("fp32", "fp32", "fgmres"), ("bf16", "bf16", "fgmres"),
            ("mixed_fp32", "fp32", "fgmres_mixed"), ("mixed_bf16", "bf16", "fgmres_mixed"),
            ("mixed_gmres_fp32", "fp32", "fgmres_mixed_gmres"), ("mixed_gmres_bf16", "bf16", "fgmres_mixed_gmres")]
    for key, prec, solver in legs:
        h.set_preconditioner("isai", prec)
        state = {}

        def one():
            call("isai_fill_f64_async", d_x.ptr, 0.0, n, stream)
            st, it, rr = h.solve(d_b, d_x, solver, tol=TOL, maxit=3000, restart=30, check=False)
            state.update(st=st, it=it, rr=rr)
        one()
        ms = timer.time(one, steps)
        xs[key] = d_x.to_numpy()
        out[key] = {"ms": round(ms, 3), "iters": state["it"], "relres": state["rr"], "status": state["st"],
                    "solver": solver}
        if solver.startswith("fgmres_mixed"):
            out[key]["outer_steps"] = h.last_stats()[0]
        if ref:
            out[key].update(error_vs_reference(xs[key], ref))
        log(f"precision form {key}: {ms:.2f} ms ({state['it']} it, relres {state['rr']:.2e})")
    for key, _, _ in legs[1:]:
        out[key]["max_rel_diff_vs_fp64"] = float(np.max(np.abs(xs[key] - xs["fp64"])) / np.max(np.abs(xs["fp64"])))
        out[key]["speedup_vs_fp64"] = round(out["fp64"]["ms"] / out[key]["ms"], 3)
    h.close()
    return out


def bench_batching(stream, steps: int) -> dict:
    """Batching form: 256 independent 2-D 9-pt systems, L uniform in [24, 200], 1..16
    right-hand sides each (numpy default_rng(0)), FP64 Krylov with FP32 ISAI, one
    batched setup and one batched BiCGSTAB replayed as a hipGraph."""
    from isai_kernel_amd import DeviceArray, call, problems
    from isai_kernel_amd.solver import Batch, Matrix

    t0 = time.perf_counter()
    rp, ci, val, sys_ptr, Ls, nrhs = problems.batch(256, 24, 200, seed=0, rhs_max=16)
    gen_s = time.perf_counter() - t0
    A = Matrix(rp, ci, val)
    t0 = time.perf_counter()
    bt = Batch(A, sys_ptr, nrhs, precision="fp32", stream=stream)
    setup_s = time.perf_counter() - t0
    ms_ilu, ms_isai = bt.setup_times()
    variant, tuned_us, ncand = bt.tuning()
    table = {"A": [round(u, 1) for u in bt.tuning_table(0)], "isai": [round(u, 1) for u in bt.tuning_table(1)]}
    rng = np.random.default_rng(1)
    B = np.zeros(bt.length)
    for s in range(len(nrhs)):
        m = 4 * int(sys_ptr[s + 1] - sys_ptr[s])
        for j in range(nrhs[s]):
            o = bt.offset(s, j)
            B[o:o + m] = rng.standard_normal(m)
    d_B = DeviceArray.from_numpy(B)
    d_X = DeviceArray((bt.length,), np.float64)

    def timed(mode):
        times, x = [], None
        for rep in range(1 + steps):
            call("isai_fill_f64_async", d_X.ptr, 0.0, bt.length, stream)
            call("isai_stream_synchronize", stream)
            t0 = time.perf_counter()
            res = bt.solve(d_B, d_X, tol=TOL, maxit=1000, use_graph=mode, check=False)
            if rep:
                times.append((time.perf_counter() - t0) * 1e3)
        return float(np.median(times)), res, d_X.to_numpy()

    # the persistent kernel (one launch per solve, grid barriers between steps),
    # then the per-step launches replayed as a graph: the same iterates bit for bit
    ms_p, (st_p, iters_p, rr_p), x_p = timed(2)
    ms, (st, iters, rr), x_g = timed(1)
    out = {"workload": "256 x 2-D 9-pt LxL (L in [24,200]), bs=4, 1..16 RHS each, BiCGSTAB to 1e-10, FP32 ISAI",
           "systems": len(nrhs), "pairs": bt.npairs, "block_rows": int(sys_ptr[-1]), "blocks": len(ci),
           "ms": round(ms, 3), "schedule": "per-step launches replayed as a hipGraph (4 iterations per replay)",
           "persistent": {"ms": round(ms_p, 3), "status": st_p, "grid": "one 1024-thread workgroup per CU",
                          "bit_identical_to_graph": bool(np.array_equal(x_p, x_g) and np.array_equal(iters_p, iters))},
           "status": st, "iters_min": int(iters.min()), "iters_mean": round(float(iters.mean()), 2),
           "iters_max": int(iters.max()), "max_relres": float(rr.max()),
           "setup_ms": {"ilu0": round(ms_ilu, 3), "isai_build": round(ms_isai, 3), "host_wall_s": round(setup_s, 2)},
           "spmm_autotune": {"variant": variant, "us": round(tuned_us, 1), "candidates": ncand,
                             "table_us": table,
                             "candidate_key": "c: lane mapping c//6 (0 row-pair, 1 column-split, 2 MFMA, "
                                              "3 whole-row: row-pair lanes for FP64 operands, MFMA lanes for FP32/BF16 factors), "
                                              "workgroup (64,128,256)[(c//2)%3], XCD remap c%2"},
           "generate_s": round(gen_s, 2)}
    log(f"batching form: {ms:.1f} ms (persistent kernel {ms_p:.1f} ms) for {bt.npairs} pairs "
        f"({int(iters.min())}-{int(iters.max())} it, max relres {rr.max():.2e}); setup ILU {ms_ilu:.1f} ms "
        f"ISAI {ms_isai:.1f} ms; SpMM {tuned_us:.0f} us")
    bt.close()
    return out


def bench_differentiation(timer, stream, steps: int) -> dict:
    """Differentiation form on the 2-D 9-pt 256^2 stencil (bs = 4, 65,536 block rows)
    with the well-posed coefficients of generator problem 3 (the specified
    coefficients make the matrix indefinite and the forward solve diverge):
    forward ISAI-BiCGSTAB to 1e-12, adjoint handle from the forward factors,
    adjoint solve to 1e-12, pattern-restricted dL/dA."""
    from isai_kernel_amd import DeviceArray, call, lib, problems
    from isai_kernel_amd.solver import Handle, Matrix

    rp, ci, val = problems.generate(3, 256, seed=0)
    nb, nnzb = len(rp) - 1, len(ci)
    n = 4 * nb
    A = Matrix(rp, ci, val)
    b, t = problems.normal(1, n), problems.normal(2, n)
    d_b = DeviceArray.from_numpy(b)
    d_x, d_lam = DeviceArray((n,), np.float64), DeviceArray((n,), np.float64)
    d_g = DeviceArray((nnzb, 4, 4), np.float64)
    h = Handle(A, precond="isai", precision="fp64", stream=stream)
    st_f = {}

    def forward():
        call("isai_fill_f64_async", d_x.ptr, 0.0, n, stream)
        st_f.update(zip(("st", "it", "rr"), h.solve(d_b, d_x, "bicgstab", tol=1e-12, maxit=2000, check=False)))
    forward()
    fwd_ms = timer.time(forward, steps)
    x = d_x.to_numpy()
    d_gx = DeviceArray.from_numpy(2.0 * (x - t))
    adj = ctypes.c_void_p()
    t0 = time.perf_counter()
    call("isai_adjoint_handle", h.h, ctypes.byref(adj), stream)
    adj_setup_ms = (time.perf_counter() - t0) * 1e3
    ah = Handle.__new__(Handle)
    ah.h, ah.stream, ah.A = adj, stream, None
    st_a = {}

    def adjoint():
        call("isai_fill_f64_async", d_lam.ptr, 0.0, n, stream)
        st_a.update(zip(("st", "it", "rr"), ah.solve(d_gx, d_lam, "bicgstab", tol=1e-12, maxit=2000, check=False)))
    adjoint()
    adj_ms = timer.time(adjoint, steps)

    def grad():
        lib.isai_bsr_outer_grad(nb, nnzb, 4, A.d_rp.ptr, A.d_ci.ptr, d_lam.ptr, d_x.ptr, -1.0, d_g.ptr, stream)
    for _ in range(5):
        grad()
    g_ms = timer.time(grad, 50)
    g_bytes = (16 * 8 + 4) * nnzb + 4 * (nb + 1) + 2 * 4 * 8 * nb
    out = {"workload": "2-D 9-pt 256^2, bs=4 (problem-3 coefficients), L = ||x - t||^2, tol 1e-12",
           "nb": nb, "nnzb": nnzb,
           "forward_ms": round(fwd_ms, 3), "forward_iters": st_f["it"], "forward_relres": st_f["rr"],
           "adjoint_handle_ms_wall": round(adj_setup_ms, 2),
           "adjoint_ms": round(adj_ms, 3), "adjoint_iters": st_a["it"], "adjoint_relres": st_a["rr"],
           "outer_grad_us": round(g_ms * 1e3, 2),
           "outer_grad_GBps": round(g_bytes / (g_ms * 1e-3) / 1e9, 1)}
    log(f"differentiation form: forward {fwd_ms:.2f} ms ({st_f['it']} it), adjoint {adj_ms:.2f} ms "
        f"({st_a['it']} it), adjoint handle {adj_setup_ms:.1f} ms wall, dL/dA {g_ms*1e3:.1f} us")
    ah.close()
    h.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed solves per form")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--kernel-iters", type=int, default=100)
    ap.add_argument("--problem", type=int, default=1)
    ap.add_argument("--chunk", type=int, default=0, help="rows per wave of the sync-free solve")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-forms", action="store_true", help="skip the precision / batching / differentiation legs")
    ap.add_argument("--leg", choices=["precision", "batching", "differentiation"],
                    help="run only this form's leg and print its object (for profiler runs)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="time the kernels only (no solves, no CPU leg): for profiler runs")
    ap.add_argument("--solve-forms", default="",
                    help="comma-separated subset of the timed solve forms (default: all): for profiler runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this path runs on one GPU; --gpus must be 1 "
                          "(the path is single-GPU: sharding across GPUs is not built)"}))
        return 2

    from isai_kernel_amd import DeviceArray, call, lib, problems
    from isai_kernel_amd.solver import Handle, Matrix, pair_stats

    name = ctypes.create_string_buffer(256)
    cus = ctypes.c_int()
    call("isai_device_info", name, 256, ctypes.byref(cus), None)
    log(f"device {name.value.decode()} CUs={cus.value}")

    if args.leg:
        stream = ctypes.c_void_p()
        call("isai_stream_create", ctypes.byref(stream))
        timer = Timer(lib, stream)
        fsteps = max(1, min(args.steps, 3))
        leg = {"precision": lambda: bench_precision(timer, stream, fsteps),
               "batching": lambda: bench_batching(stream, fsteps),
               "differentiation": lambda: bench_differentiation(timer, stream, fsteps)}[args.leg]()
        print(json.dumps({"leg": args.leg, args.leg: leg}), flush=True)
        return 0

    t0 = time.perf_counter()
    rp, ci, val = problems.generate(args.problem, 0, seed=0)
    nb, nnzb = len(rp) - 1, len(ci)
    n = nb * 4
    log(f"generated problem {args.problem}: nb={nb} nnzb={nnzb} in {time.perf_counter()-t0:.1f}s")

    stream = ctypes.c_void_p()
    call("isai_stream_create", ctypes.byref(stream))
    timer = Timer(lib, stream)
    A = Matrix(rp, ci, val)
    d_one = DeviceArray.from_numpy(np.ones(n))
    d_b = DeviceArray((n,), np.float64)
    d_x = DeviceArray((n,), np.float64)
    d_y = DeviceArray((n,), np.float64)
    A.spmv(d_one, d_b, stream)
    call("isai_stream_synchronize", stream)

    # ---- setup: ILU(0) + ISAI build (device time from the library's events)
    t0 = time.perf_counter()
    h = Handle(A, precond="isai", precision="fp64", chunk=args.chunk, stream=stream)
    setup_wall = time.perf_counter() - t0
    ms_ilu, ms_isai, ms_setup = h.times()
    nnzbL, nnzbU, nlevels = h.info()
    log(f"setup: ILU(0) {ms_ilu:.2f} ms, ISAI build {ms_isai:.2f} ms (levels {nlevels}, "
        f"nnzbL {nnzbL}, host wall incl. analysis {setup_wall:.2f} s)")
    # The first setup of a process also pays the first launch of each setup kernel
    # (code loading, about 0.5-1.5 ms per kernel); a second setup on the same matrix
    # gives the steady-state figures.
    h_warm = Handle(A, precond="isai", precision="fp64", chunk=args.chunk, stream=stream)
    ms_ilu_warm, ms_isai_warm, _ = h_warm.times()
    h_warm.close()
    log(f"setup, second handle: ILU(0) {ms_ilu_warm:.2f} ms, ISAI build {ms_isai_warm:.2f} ms")

    # ---- device CSR -> BSR of A from its CSR form (setup-side conversion, wall time)
    d_crp = DeviceArray((n + 1,), np.int32)
    d_cci = DeviceArray((nnzb * 16,), np.int32)
    d_cv = DeviceArray((nnzb * 16,), np.float64)
    call("isai_bsr_to_csr", nb, nnzb, 4, A.d_rp.ptr, A.d_ci.ptr, A.d_val.ptr, d_crp.ptr, d_cci.ptr, d_cv.ptr, stream)
    d_brp = DeviceArray((nb + 1,), np.int32)
    d_bci = DeviceArray((nnzb,), np.int32)
    d_bv = DeviceArray((nnzb, 4, 4), np.float64)
    c2b_ms = []
    for _ in range(3):
        call("isai_stream_synchronize", stream)
        t0 = time.perf_counter()
        cnt = ctypes.c_int()
        call("isai_csr_to_bsr_device_count", n, 4, d_crp.ptr, d_cci.ptr, d_brp.ptr, ctypes.byref(cnt), stream)
        call("isai_csr_to_bsr_device_fill", n, 4, d_crp.ptr, d_cci.ptr, d_cv.ptr, d_brp.ptr, d_bci.ptr, d_bv.ptr,
             stream)
        c2b_ms.append((time.perf_counter() - t0) * 1e3)
    csr2bsr_ms = min(c2b_ms)
    log(f"device csr2bsr (count + fill, {nnzb * 16} CSR entries): {csr2bsr_ms:.1f} ms wall")
    del d_crp, d_cci, d_cv, d_brp, d_bci, d_bv

    # ---- pair statistics of the ISAI(L) pattern (the input of the build plan)
    mbL, mcL = problems.bcsc_pattern(rp, ci, True)
    d_mb, d_mc = DeviceArray.from_numpy(mbL), DeviceArray.from_numpy(mcL)
    d_pp = DeviceArray((len(mcL) + 1,), np.int32)
    tot = ctypes.c_int()
    call("isai_pairs_count", nb, len(mcL), A.d_rp.ptr, A.d_ci.ptr, d_mb.ptr, d_mc.ptr, 0, d_pp.ptr,
         ctypes.byref(tot), stream)
    t0 = time.perf_counter()
    ps = pair_stats(d_mb, d_pp, nb, stream)
    stats_ms = (time.perf_counter() - t0) * 1e3
    plan = h.build_plan()
    pair_stats_l = {k: ps[k] for k in ("mnnz", "total_pairs", "n_avg", "n_max", "n_p99", "q_avg", "q_max", "q_p50",
                                       "q_p99", "qlen_avg", "qlen_max", "qlen_hist")}
    pair_stats_l["wall_ms"] = round(stats_ms, 2)
    log(f"ISAI(L) pair statistics in {stats_ms:.1f} ms: N avg {ps['n_avg']:.2f} max {ps['n_max']}, "
        f"Q avg {ps['q_avg']:.2f} max {ps['q_max']}; build plan {plan}")
    del d_mb, d_mc, d_pp

    kernels = {}
    K = args.kernel_iters

    # ---- plain copy kernel: empirical bandwidth ceiling and byte-counter reference
    ncopy = 64 * 1024 * 1024  # doubles: 512 MiB read + 512 MiB written per launch
    d_c0, d_c1 = DeviceArray((ncopy,), np.float64), DeviceArray((ncopy,), np.float64)

    def copy():
        lib.isai_copy_f64_async(d_c1.ptr, d_c0.ptr, ncopy, stream)
    for _ in range(3):
        copy()
    copy_ms = timer.time(copy, max(K // 5, 5))
    copy_gbps = 2 * 8 * ncopy / (copy_ms * 1e-3) / 1e9
    log(f"copy kernel: {copy_ms*1e3:.1f} us for {2*8*ncopy/2**20:.0f} MiB moved, {copy_gbps:.0f} GB/s")
    del d_c0, d_c1

    def record(name, ms, nbytes, **extra):
        gbps = nbytes / (ms * 1e-3) / 1e9
        kernels[name] = {"us": round(ms * 1e3, 2), "GBps": round(gbps, 1),
                         "frac_hbm": round(gbps / HBM_PEAK_GBPS, 4), "algorithmic_bytes": nbytes, **extra}
        log(f"{name}: {ms*1e3:.1f} us, {gbps:.0f} GB/s ({100*gbps/HBM_PEAK_GBPS:.1f}% of 8 TB/s)")

    # ---- BSR SpMV with A (FP64)
    def spmv():
        lib.isai_bsr_spmv(nb, nnzb, 4, A.d_rp.ptr, A.d_ci.ptr, A.d_val.ptr, d_one.ptr, d_y.ptr, stream)
    for _ in range(10):
        spmv()
    spmv_ms = timer.time(spmv, K)
    record("bsr_spmv_fp64", spmv_ms, spmv_bytes(nb, nnzb))

    # ---- the mixed solve's A product: FP32 values, FP32 x and y, two fused dots, on
    # the handle's line-aligned FP32 copy of A (what the solve runs), and the public
    # product on the unpadded FP32 values for comparison
    d_a32 = DeviceArray((nnzb * 16,), np.float32)
    call("isai_convert_values", nnzb * 16, 1, A.d_val.ptr, d_a32.ptr, stream)
    d_xf = DeviceArray.from_numpy(np.ones(n, np.float32))
    d_yf = DeviceArray((n,), np.float32)
    d_part = DeviceArray((2 * 4096,), np.float64)
    groups, rows_per_group = ctypes.c_int(), ctypes.c_int()

    def spmv32():
        lib.isai_handle_spmv_f32_dot(h.h, d_xf.ptr, d_yf.ptr, d_xf.ptr, 2, d_part.ptr, ctypes.byref(groups),
                                     ctypes.byref(rows_per_group), stream)

    def spmv32_unaligned():
        lib.isai_bsr_spmv_f32_dot(nb, nnzb, A.d_rp.ptr, A.d_ci.ptr, d_a32.ptr, d_xf.ptr, d_yf.ptr, d_xf.ptr, 2,
                                  d_part.ptr, ctypes.byref(groups), stream)
    for _ in range(10):
        spmv32()
    spmv32_ms = timer.time(spmv32, K)
    record("bsr_spmv_fp32_dot", spmv32_ms, f32_spmv_dot_bytes(nb, nnzb), rows_per_group=rows_per_group.value,
           layout="line-aligned FP32 copy of A (handle), non-temporal value loads")
    for _ in range(10):
        spmv32_unaligned()
    record("bsr_spmv_fp32_dot_unaligned", timer.time(spmv32_unaligned, K), f32_spmv_dot_bytes(nb, nnzb),
           layout="caller's FP32 values (isai_bsr_spmv_f32_dot), through L2")
    del d_a32, d_xf, d_yf, d_part

    # ---- ISAI apply (M_U (M_L v)) in each storage precision
    for prec, vb in (("fp64", 8), ("fp32", 4), ("bf16", 2)):
        h.set_preconditioner("isai", prec)

        def apply():
            lib.isai_precond_apply(h.h, d_b.ptr, d_y.ptr, stream)
        for _ in range(5):
            apply()
        ms = timer.time(apply, K)
        formula = apply_bytes_baseline_formula(nb, nnzbL, nnzbU, vb)
        record(f"isai_apply_{prec}", ms, apply_bytes(nb, nnzbL, nnzbU, vb), baseline_formula_bytes=formula,
               frac_hbm_baseline_formula=round(formula / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4))

    # ---- sync-free SpTRSV (L then U with the packed ILU(0) factor)
    # the first selection of ILU(0) runs the triangular-solve analysis (host order
    # and tables, then the pre-scaled factor on the device)
    t_sv = time.perf_counter()
    h.set_preconditioner("ilu", "fp64")
    call("isai_stream_synchronize", stream)
    ms_sv = (time.perf_counter() - t_sv) * 1e3

    def trsv():
        lib.isai_precond_apply(h.h, d_b.ptr, d_y.ptr, stream)
    for _ in range(3):
        trsv()
    tr_ms = timer.time(trsv, max(K // 4, 10))
    record("sptrsv_lu_fp64", tr_ms, sptrsv_bytes(nb, nnzbL, nnzbU))

    if args.kernels_only:
        h.close()
        print(json.dumps({"kernels": kernels, "copy_GBps": round(copy_gbps, 1),
                          "setup_ms": {"sptrsv_analysis": round(ms_sv, 3)}}), flush=True)
        return 0

    # ---- solves (the steps)
    # mixed_*: FP64 FGMRES(10) outer loop, inner FP32-storage BiCGSTAB (the mixed form);
    # isai_*: FP64 BiCGSTAB with the ISAI factors stored in that precision;
    # ilu_sptrsv_fp64: FP64 BiCGSTAB with exact ILU(0) solves
    solves = {}
    forms = {"mixed_bf16": ("isai", "bf16", "fgmres_mixed"), "mixed_fp32": ("isai", "fp32", "fgmres_mixed"),
             "isai_bf16": ("isai", "bf16", "bicgstab"), "isai_fp32": ("isai", "fp32", "bicgstab"),
             "isai_fp64": ("isai", "fp64", "bicgstab"), "ilu_sptrsv_fp64": ("ilu", "fp64", "bicgstab")}
    if args.solve_forms:
        keep = args.solve_forms.split(",")
        forms = {k: v for k, v in forms.items() if k in keep}
    for form, (pc, prec, solver) in forms.items():
        h.set_preconditioner(pc, prec)
        state = {}

        def one_solve():
            call("isai_fill_f64_async", d_x.ptr, 0.0, n, stream)
            st, it, rr = h.solve(d_b, d_x, solver, tol=TOL, maxit=1000, restart=10, check=False)
            state.update(st=st, it=it, rr=rr)
        for _ in range(args.warmup):
            one_solve()
        call("isai_stream_synchronize", stream)
        t0 = time.perf_counter()
        ms = timer.time(one_solve, args.steps)
        wall = (time.perf_counter() - t0) / args.steps * 1e3
        x = d_x.to_numpy()
        # the mixed form also converts A to FP32 once per handle (outside the timed solves)
        setup = ms_ilu + (ms_isai if pc == "isai" else ms_sv)
        outer, inner = h.last_stats()
        solves[form] = {"ms": round(ms, 3), "wall_ms": round(wall, 3), "ms_with_setup": round(ms + setup, 3),
                        "iters": state["it"], "solver": solver,
                        "relres": state["rr"], "status": state["st"],
                        "max_err_vs_ones": float(np.max(np.abs(x - 1.0)))}
        if solver == "fgmres_mixed":
            solves[form]["outer_steps"] = outer
        if solver == "bicgstab":
            # the same solve with its iterations replayed from a captured hipGraph
            h.set_graph(True)
            eager = (state["it"], state["rr"])
            one_solve()  # captures
            ms_g = timer.time(one_solve, args.steps)
            h.set_graph(False)
            solves[form]["graph"] = {"ms": round(ms_g, 3), "iters": state["it"],
                                     "bit_identical": bool((state["it"], state["rr"]) == eager
                                                           and np.array_equal(d_x.to_numpy(), x))}
            log(f"solve {form} (graph): {ms_g:.2f} ms")
        log(f"solve {form}: {ms:.2f} ms ({state['it']} it, relres {state['rr']:.2e}, status {state['st']})")

    ok = {k: v for k, v in solves.items() if v["status"] == 0 and v["relres"] <= TOL}
    mixed = "mixed_bf16" if "mixed_bf16" in ok else None
    value = ok[mixed]["ms"] if mixed else None
    speedups = {}
    for fast in ("mixed_bf16", "mixed_fp32", "isai_bf16", "isai_fp32"):
        for base in ("isai_fp64", "ilu_sptrsv_fp64"):
            if fast in ok and base in ok:
                speedups[f"{fast}_vs_{base}"] = round(ok[base]["ms"] / ok[fast]["ms"], 3)

    gbps = kernels["bsr_spmv_fp32_dot"]["GBps"]
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": value,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+bf16 inner, f64 outer",
        "data": "synthetic",
        "config": {"workload": "mixed-precision ISAI-preconditioned solve to 1e-10 (FP64 FGMRES(10) outer loop, "
                               "inner FP32-storage BiCGSTAB with BF16 ISAI factors, FP64 dot products), "
                               "3-D 7-pt 96^3, bs=4 (884736 block rows, 6137856 blocks), b = A*1",
                   "problem": args.problem, "nb": nb, "nnzb": nnzb, "mixed_form": mixed},
        "roofline": {"bound": "hbm", "achieved": gbps, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(gbps / HBM_PEAK_GBPS, 4),
                     "traffic": (int(F32_SPMV_TRAFFIC_BYTES_96CUBE)
                                 if args.problem == 1 and F32_SPMV_TRAFFIC_BYTES_96CUBE else None),
                     "kernel": (f"bsr4_spmv_f32_dot_kernel<2, 8, {kernels['bsr_spmv_fp32_dot']['rows_per_group']}, "
                                "true>"),
                     "algorithmic_bytes": f32_spmv_dot_bytes(nb, nnzb)},
        "speedup": speedups,
        "solves": solves,
        "kernels": kernels,
        "setup_ms": {"ilu0": round(ms_ilu, 3), "isai_build": round(ms_isai, 3), "ilu0_warm": round(ms_ilu_warm, 3),
                     "isai_build_warm": round(ms_isai_warm, 3), "sptrsv_analysis": round(ms_sv, 3),
                     "levels": nlevels,
                     "csr2bsr_device_wall": round(csr2bsr_ms, 2)},
        "isai_build_plan": plan,
        "pair_stats_L": pair_stats_l,
        "copy_kernel_GBps": round(copy_gbps, 1),
    }
    h.close()
    del A
    if not args.no_forms:
        fsteps = max(1, min(args.steps, 3))
        result["forms"] = {"precision": bench_precision(timer, stream, fsteps),
                           "batching": bench_batching(stream, fsteps),
                           "differentiation": bench_differentiation(timer, stream, fsteps)}
    if not args.no_cpu:
        c = cpu_baseline(rp, ci, val, d_b.to_numpy())
        result["cpu_baseline"] = {
            "value": round(c["solve_s"] * 1e3, 1), "unit": "ms", "cores": c["cores"], "kind": "port",
            "sample": (f"oracle FP64 ISAI-preconditioned BiCGSTAB on the full 96^3 matrix to 1e-10: "
                       f"{c['iters']} iterations, relres {c['relres']:.2e}; its ILU(0)+ISAI setup took "
                       f"{c['setup_s']:.1f} s (not included)"),
        }
        log(f"cpu baseline: solve {c['solve_s']*1e3:.0f} ms ({c['iters']} it), setup {c['setup_s']:.1f} s, "
            f"{c['cores']} threads")
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
