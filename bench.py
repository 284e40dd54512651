#!/usr/bin/env python3
"""Benchmark of the block-sparse ILU(0)/ISAI preconditioning path on one MI355X.

Workload: 3-D 7-point stencil on a 96^3 grid, 4 dof per node (bs = 4, 884,736
block rows, 6,137,856 blocks), FP64, generated in-process (synthetic, seed 0).

A "step" is one pass of the hot path over that matrix.  Kernel figures are
hipEvent timings on the stream the kernels run on; the roofline object is for
the dominant kernel (the BSR SpMV) with algorithmic bytes per launch
(16*8+4)*nnzb + 4*(nb+1) + 2*4*8*nb.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "solve time to 1e-10 rel-res, mixed vs FP64 ISAI; BSR SpMV/SpTRSV HBM GB/s"
HBM_PEAK_GBPS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spmv_bytes(nb: int, nnzb: int, value_bytes: int = 8) -> int:
    return (16 * value_bytes + 4) * nnzb + 4 * (nb + 1) + 2 * 4 * 8 * nb


class Timer:
    """hipEvent pair on one stream."""

    def __init__(self, lib, stream):
        self.lib, self.stream = lib, stream
        self.e0, self.e1 = ctypes.c_void_p(), ctypes.c_void_p()
        lib.isai_event_create(ctypes.byref(self.e0))
        lib.isai_event_create(ctypes.byref(self.e1))

    def time(self, fn, iters: int) -> float:
        """Average milliseconds of fn() over iters launches."""
        from isai_kernel_amd import check
        check(self.lib.isai_event_record(self.e0, self.stream), "event_record")
        for _ in range(iters):
            fn()
        check(self.lib.isai_event_record(self.e1, self.stream), "event_record")
        check(self.lib.isai_event_synchronize(self.e1), "event_sync")
        ms = ctypes.c_float()
        check(self.lib.isai_event_elapsed_ms(ctypes.byref(ms), self.e0, self.e1), "elapsed")
        return ms.value / iters


def cpu_spmv_baseline(rp, ci, val, x, budget_s: float = 12.0):
    """Oracle SpMV on the host cores over the full matrix, repeated within a budget."""
    sys.path.insert(0, str(ROOT / "tests"))
    import _oracle as O  # the checker, used here only as the CPU baseline

    nb = len(rp) - 1
    reps, t0 = 0, time.perf_counter()
    while True:
        O.spmv(rp, ci, val, x)
        reps += 1
        el = time.perf_counter() - t0
        if el > budget_s or reps >= 50:
            break
    ms = el / reps * 1e3
    return ms, reps, O.num_threads(), nb


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--problem", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this path runs on one GPU; --gpus must be 1 "
                          "(multi-GPU sharding is described in DESIGN.md, not built)"}))
        return 2

    from isai_kernel_amd import DeviceArray, call, lib, problems

    name = ctypes.create_string_buffer(256)
    cus = ctypes.c_int()
    call("isai_device_info", name, 256, ctypes.byref(cus), None)
    log(f"device {name.value.decode()} CUs={cus.value}")

    t0 = time.perf_counter()
    rp, ci, val = problems.generate(args.problem, 0, seed=0)
    nb, nnzb = len(rp) - 1, len(ci)
    n = nb * 4
    log(f"generated problem {args.problem}: nb={nb} nnzb={nnzb} in {time.perf_counter()-t0:.1f}s")

    stream = ctypes.c_void_p()
    call("isai_stream_create", ctypes.byref(stream))
    timer = Timer(lib, stream)

    d_rp = DeviceArray.from_numpy(rp)
    d_ci = DeviceArray.from_numpy(ci)
    d_val = DeviceArray.from_numpy(val)
    x = np.ones(n)
    d_x = DeviceArray.from_numpy(x)
    d_y = DeviceArray((n,), np.float64)

    def spmv():
        lib.isai_bsr_spmv(nb, nnzb, 4, d_rp.ptr, d_ci.ptr, d_val.ptr, d_x.ptr, d_y.ptr, stream)

    for _ in range(args.warmup):
        spmv()
    call("isai_stream_synchronize", stream)
    spmv_ms = timer.time(spmv, args.steps)
    bytes_spmv = spmv_bytes(nb, nnzb)
    spmv_gbps = bytes_spmv / (spmv_ms * 1e-3) / 1e9
    log(f"SpMV fp64: {spmv_ms*1e3:.1f} us/launch, {spmv_gbps:.0f} GB/s "
        f"({100*spmv_gbps/HBM_PEAK_GBPS:.1f}% of 8 TB/s)")

    result = {
        "metric": METRIC,
        "value": round(spmv_gbps, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(spmv_ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "bsr_spmv 3-D 7-pt 96^3 bs=4 (884736 block rows, 6137856 blocks)",
                   "problem": args.problem, "nb": nb, "nnzb": nnzb},
        "roofline": {"bound": "hbm", "achieved": round(spmv_gbps, 2), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(spmv_gbps / HBM_PEAK_GBPS, 4), "traffic": None,
                     "kernel": "bsr4_spmv_f64_kernel", "algorithmic_bytes": bytes_spmv},
    }
    if not args.no_cpu:
        ms, reps, cores, _ = cpu_spmv_baseline(rp, ci, val, x)
        result["cpu_baseline"] = {
            "value": round(bytes_spmv / (ms * 1e-3) / 1e9, 3), "unit": "GB/s", "cores": cores,
            "kind": "port",
            "sample": f"oracle BSR SpMV over the full 96^3 matrix, {reps} repetitions, {ms:.1f} ms each",
        }
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
