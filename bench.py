"""Benchmark: FullCI time-to-solution on one MI355X, with the sigma roofline and
the CPU baseline.

Workload (default): FullCI, nbasis 12, 6 alpha + 6 beta electrons (853,776
determinants; 1,553,018,544 stored nonzeros), seeded integrals (48 Cholesky-like
vectors, seed 1001), ecore 0, nroot 1.

A step is one pass of the hot path with integrals and determinants already in
HBM: CSR build (pyci_op_create) + Davidson solve (pyci_op_solve) until the
energy is within 1e-8 Eh of the FP64 reference energy. `value` is the mean
time-to-solution of the timed steps (seconds, lower is better).

The FP64 reference energy of this workload comes from an untimed FP64 solve at
tol 1e-12 of the same operator (the CSR builder and Davidson are parity-checked
against the CPU oracle on smaller systems by the test suite).

roofline: the dominant kernel is the sigma SpMV (Y = H x) over the stored CSR
(int32 columns + values, int64 row pointers). Its algorithmic bytes per launch
are 12*nnz + 8*(nrow+1) + 8*(ncol+nrow) for FP64 values; its average duration
is measured with HIP events on the engine's stream (pyci_op_bench_sigma).

cpu_baseline: the CPU oracle (a C restatement of the reference's serial
SparseOp::update + perform_op) timed on the host cores on a bounded sample of
rows of the same operator, extrapolated to the full build plus the same number
of sigma products the GPU solve needed.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from pyci_amd.synthetic import config_integrals  # noqa: E402

HBM_PEAK_GBPS = 8000.0
WORKLOADS = {
    "fci12": dict(name="fci12", nbasis=12, nup=6, ndn=6, integrals="fci12", nroot=1),
    "fci8": dict(name="fci8", nbasis=8, nup=4, ndn=4, integrals="fci8", nroot=1),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Stats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int64), ("sigma_calls", ctypes.c_int64), ("sigma_calls_f32", ctypes.c_int64),
                ("residual", ctypes.c_double * 64)]


def engine():
    L = ctypes.CDLL(os.path.join(ROOT, "pyci_amd", "libpyci_engine.so"))
    vp, i64, dp = ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_double)
    L.pyci_last_error.restype = ctypes.c_char_p
    L.pyci_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
    L.pyci_ctx_synchronize.argtypes = [vp]
    L.pyci_ham_create.argtypes = [vp, i64, ctypes.c_double, dp, dp, ctypes.POINTER(vp)]
    L.pyci_wfn_create.argtypes = [vp, ctypes.c_int, i64, i64, i64, i64, ctypes.c_void_p, ctypes.POINTER(vp)]
    L.pyci_fill_all_dets.argtypes = [vp, ctypes.c_int, i64, i64, i64, ctypes.c_void_p]
    L.pyci_op_create.argtypes = [vp, vp, i64, i64, ctypes.c_int, ctypes.POINTER(vp)]
    L.pyci_op_destroy.argtypes = [vp]
    L.pyci_op_info.argtypes = [vp, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                               ctypes.c_void_p, ctypes.c_void_p]
    L.pyci_op_solve.argtypes = [vp, i64, dp, i64, i64, ctypes.c_double, ctypes.c_int, dp, dp, ctypes.POINTER(Stats)]
    L.pyci_op_bench_sigma.argtypes = [vp, i64, i64, ctypes.c_int, ctypes.POINTER(ctypes.c_float), dp]
    return L


def ok(L, rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {L.pyci_last_error().decode()}")


def dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def measured_traffic(name):
    """HBM bytes per sigma launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary (profiles/sigma_<workload>_hbm_traffic.md), or None."""
    path = os.path.join(ROOT, "profiles", f"sigma_{name}_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("| traffic = read + write |"):
            return float(line.split("|")[2].split()[0]) * 1e9
    return None


def device_copy(nbytes):
    """One device-to-device copy of nbytes through the HIP runtime (a plain copy
    kernel of known size, used to calibrate the HBM byte counters)."""
    hip = ctypes.CDLL("libamdhip64.so")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    t0 = time.perf_counter()
    assert hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    t = time.perf_counter() - t0
    hip.hipFree(src)
    hip.hipFree(dst)
    return t


def cpu_baseline(wl, one, two, dets, sigma_calls, budget_s):
    """Oracle (serial C restatement of the reference path) on a bounded row sample."""
    from oracle import oracle as O

    O.build()
    nd = len(dets)
    rows = 256
    t_row = None
    while True:
        r0 = (nd - rows) // 2  # rows from the middle of the matrix
        t0 = time.perf_counter()
        O.csr_build("fullci", wl["nbasis"], wl["nup"], wl["ndn"], dets, one, two, nrow=r0 + rows, symmetric=True,
                    row0=r0)
        t_build = time.perf_counter() - t0
        # time of the determinant index construction alone (paid once per build)
        t0 = time.perf_counter()
        O.csr_build("fullci", wl["nbasis"], wl["nup"], wl["ndn"], dets, one, two, nrow=1, symmetric=True)
        t_index = time.perf_counter() - t0
        t_row = max(t_build - t_index, 1e-9) / rows
        if t_build > budget_s / 4 or rows >= nd:
            break
        rows = min(nd, int(rows * max(2.0, (budget_s / 4) / max(t_build, 1e-3))))
    # serial SpMV on the sampled lower-triangle rows (16 B per stored entry, as the reference)
    x = np.random.default_rng(0).normal(size=nd)
    L = O.lib()
    y = np.zeros(nd)
    r0 = (nd - rows) // 2
    t0 = time.perf_counter()
    reps = 0
    h = L.oracle_csr_build(1, wl["nbasis"], wl["nup"], wl["ndn"], nd, dets.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                           dp(np.ascontiguousarray(one)), dp(np.ascontiguousarray(two)), r0, r0 + rows, nd, 1)
    t0 = time.perf_counter()
    while True:
        L.oracle_csr_matvec(h, 1, dp(x), dp(y))
        reps += 1
        if time.perf_counter() - t0 > budget_s / 4:
            break
    t_spmv_rows = (time.perf_counter() - t0) / reps
    nnz_sample = L.oracle_csr_nnz(h)
    L.oracle_csr_free(h)
    total_lower = None
    return dict(t_row=t_row, t_index=t_index, rows=rows, t_spmv_sample=t_spmv_rows, nnz_sample=nnz_sample,
                sigma_calls=sigma_calls, total_lower=total_lower)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="fci12", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="f64", choices=["f64", "mixed"])
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--cpu-seconds", type=float, default=16.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sigma-only", action="store_true",
                    help="build the operator, run a few sigma launches and a device copy of known size "
                         "(for hardware-counter runs), print their timings and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    wl = WORKLOADS[args.workload]
    L = engine()
    ctx = ctypes.c_void_p()
    ok(L, L.pyci_ctx_create(0, ctypes.byref(ctx)), "pyci_ctx_create")
    ecore, one, two = config_integrals(wl["integrals"])
    one, two = np.ascontiguousarray(one), np.ascontiguousarray(two)
    nb, nu, nd_ = wl["nbasis"], wl["nup"], wl["ndn"]
    from math import comb

    nd = comb(nb, nu) * comb(nb, nd_)
    dets = np.zeros((nd, 2), dtype=np.uint64)
    t0 = time.perf_counter()
    ok(L, L.pyci_fill_all_dets(ctx, 1, nb, nu, nd_, dets.ctypes.data_as(ctypes.c_void_p)), "fill_all_dets")
    t_fill = time.perf_counter() - t0
    ham, wfn = ctypes.c_void_p(), ctypes.c_void_p()
    ok(L, L.pyci_ham_create(ctx, nb, ecore, dp(one), dp(two), ctypes.byref(ham)), "ham_create")
    ok(L, L.pyci_wfn_create(ctx, 1, nb, nu, nd_, nd, dets.ctypes.data_as(ctypes.c_void_p), ctypes.byref(wfn)),
       "wfn_create")
    log(f"workload {wl['name']}: {nd} determinants resident in HBM (colex fill {t_fill*1e3:.1f} ms)")
    prec = 0 if args.precision == "f64" else 1
    nroot = wl["nroot"]

    def build():
        op = ctypes.c_void_p()
        ok(L, L.pyci_op_create(ham, wfn, -1, -1, 1, ctypes.byref(op)), "op_create")
        return op

    def info(op):
        v = [ctypes.c_int64() for _ in range(4)]
        sym, ec = ctypes.c_int(), ctypes.c_double()
        ok(L, L.pyci_op_info(op, *(ctypes.byref(x) for x in v), ctypes.byref(sym), ctypes.byref(ec)), "op_info")
        return [x.value for x in v]

    def solve(op, p, tol):
        ev = np.zeros(nroot)
        vec = np.zeros((nroot, nd))
        st = Stats()
        ok(L, L.pyci_op_solve(op, nroot, None, -1, -1, tol, p, dp(ev), dp(vec), ctypes.byref(st)), "op_solve")
        return ev, st

    if args.sigma_only:
        op = build()
        nrow, ncol, size_ref, nnz = info(op)
        ms, nbytes = ctypes.c_float(), ctypes.c_double()
        ok(L, L.pyci_op_bench_sigma(op, 1, 4, 0, ctypes.byref(ms), ctypes.byref(nbytes)), "bench_sigma")
        copy_bytes = 4 << 30
        t_copy = device_copy(copy_bytes)
        print(json.dumps({"sigma_avg_ms": ms.value, "sigma_bytes": nbytes.value, "nnz": nnz,
                          "copy_bytes": copy_bytes, "copy_ms": t_copy * 1e3}), flush=True)
        return
    # FP64 reference energy (untimed, tight tolerance)
    op = build()
    nrow, ncol, size_ref, nnz = info(op)
    e_ref, st_ref = solve(op, 0, 1e-12)
    log(f"reference FP64 energy {e_ref[0]:.12f} ({st_ref.iterations} iterations); stored nnz {nnz}, "
        f"reference-format nnz {size_ref}")
    # sigma kernel on the same operator (HIP events on the engine stream)
    ms, nbytes = ctypes.c_float(), ctypes.c_double()
    ok(L, L.pyci_op_bench_sigma(op, 1, 20, 0, ctypes.byref(ms), ctypes.byref(nbytes)), "bench_sigma")
    sig = dict(avg_ms=ms.value, bytes=nbytes.value)
    ok(L, L.pyci_op_bench_sigma(op, 1, 20, 1, ctypes.byref(ms), ctypes.byref(nbytes)), "bench_sigma f32")
    sig32 = dict(avg_ms=ms.value, bytes=nbytes.value)
    ok(L, L.pyci_op_bench_sigma(op, 8, 10, 0, ctypes.byref(ms), ctypes.byref(nbytes)), "bench_sigma nv8")
    sig8 = dict(avg_ms=ms.value, bytes=nbytes.value)
    L.pyci_op_destroy(op)
    log(f"sigma f64: {sig['avg_ms']:.3f} ms, {sig['bytes']/sig['avg_ms']/1e6:.0f} GB/s; "
        f"f32: {sig32['avg_ms']:.3f} ms, {sig32['bytes']/sig32['avg_ms']/1e6:.0f} GB/s; "
        f"f64 nvec=8: {sig8['avg_ms']:.3f} ms")

    def step():
        ok(L, L.pyci_ctx_synchronize(ctx), "sync")
        t0 = time.perf_counter()
        o = build()
        ok(L, L.pyci_ctx_synchronize(ctx), "sync")
        t1 = time.perf_counter()
        e, st = solve(o, prec, args.tol)
        ok(L, L.pyci_ctx_synchronize(ctx), "sync")
        t2 = time.perf_counter()
        L.pyci_op_destroy(o)
        return t1 - t0, t2 - t1, e, st

    for w in range(args.warmup):
        tb, ts, e, st = step()
        log(f"warmup {w}: build {tb*1e3:.1f} ms, solve {ts*1e3:.1f} ms")
    ok(L, L.pyci_ctx_synchronize(ctx), "sync")
    res = []
    t_all0 = time.perf_counter()
    for k in range(args.steps):
        tb, ts, e, st = step()
        de = abs(e[0] - e_ref[0])
        res.append((tb, ts, de, st.iterations, st.sigma_calls))
        log(f"step {k}: build {tb*1e3:.1f} ms, solve {ts*1e3:.1f} ms, |dE| {de:.2e}, {st.iterations} iterations")
    ok(L, L.pyci_ctx_synchronize(ctx), "sync")
    t_all = time.perf_counter() - t_all0
    tts = t_all / args.steps
    tb = float(np.mean([r[0] for r in res]))
    ts = float(np.mean([r[1] for r in res]))
    max_de = max(r[2] for r in res)
    sigma_calls = int(np.mean([r[4] for r in res]))
    achieved = sig["bytes"] / (sig["avg_ms"] * 1e-3) / 1e9
    traffic = measured_traffic(wl["name"])
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic,
                "kernel": "k_spmv<4,double,double> (sigma, FP64 values)", "avg_ms": round(sig["avg_ms"], 4),
                "bytes_per_launch": sig["bytes"]}
    out = {
        "metric": "time-to-|dE|<1e-8 Eh vs FP64 oracle; sigma HBM GB/s; dets/s built",
        "value": round(tts, 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(tts * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == 0 else "f32-sigma+f64",
        "data": "synthetic (seeded integrals)",
        "config": {"workload": f"FullCI nbasis={nb} {nu}a+{nd_}b seeded integrals (48 vectors, seed 1001), nroot=1",
                   "ndet": nd, "nnz_stored": nnz, "nnz_reference_format": size_ref, "parallelism": "single GPU",
                   "precision": args.precision, "tol": args.tol},
        "build_ms": round(tb * 1e3, 3),
        "solve_ms": round(ts * 1e3, 3),
        "dets_per_s_built": round(nd / tb, 1),
        "nnz_per_s_built": round(nnz / tb, 1),
        "energy": float(e_ref[0]),
        "max_abs_dE": max_de,
        "iterations": int(np.mean([r[3] for r in res])),
        "sigma_calls": sigma_calls,
        "sigma_gbps": round(achieved, 1),
        "sigma_f32_gbps": round(sig32["bytes"] / (sig32["avg_ms"] * 1e-3) / 1e9, 1),
        "sigma_nvec8_ms": round(sig8["avg_ms"], 4),
        "roofline": roofline,
    }
    if not args.no_cpu:
        log("cpu baseline: oracle on a bounded row sample ...")
        cb = cpu_baseline(wl, one, two, dets, sigma_calls, args.cpu_seconds)
        frac = size_ref / max(cb["nnz_sample"], 1)  # full lower nnz / sample lower nnz
        t_build_cpu = cb["t_index"] + cb["t_row"] * nd
        t_spmv_cpu = cb["t_spmv_sample"] * frac
        t_cpu = t_build_cpu + sigma_calls * t_spmv_cpu
        out["cpu_baseline"] = {
            "value": round(t_cpu, 3),
            "unit": "s",
            "cores": 1,
            "kind": "port",
            "sample": (f"oracle CSR build of {cb['rows']} middle rows of {nd} ({cb['t_row']*1e6:.1f} us/row, "
                       f"index {cb['t_index']:.2f} s) and serial lower-triangle SpMV over them "
                       f"({cb['t_spmv_sample']*1e3:.1f} ms for {cb['nnz_sample']} nnz), extrapolated to the full "
                       f"build + {sigma_calls} sigma products"),
            "dets_per_s_built": round(1.0 / cb["t_row"], 1),
        }
        log(f"cpu baseline estimate {t_cpu:.1f} s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
