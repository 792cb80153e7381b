"""Benchmark: FullCI time-to-solution on one MI355X, with the sigma roofline and
the CPU baseline.

Workload (default): FullCI, nbasis 12, 6 alpha + 6 beta electrons (853,776
determinants; 1,553,018,544 stored nonzeros), seeded integrals (48 Cholesky-like
vectors, seed 1001), ecore 0, nroot 1.

A step is one pass of the hot path with integrals and determinants already in
HBM: sparse_op (pyci_op_create_ex, storage --storage, default "auto": the
matrix-free FullCI operator for this full space) + Davidson solve
(pyci_op_solve, precision --precision) until the energy is within 1e-8 Eh of
the FP64 reference energy. `value` is the mean time-to-solution of the timed
steps (seconds, lower is better). The plain FP64 GPU path (CSR build + FP64
solve) is timed the same way beside it ("baseline_csr_f64").

The FP64 reference energies: for fci12 the CPU oracle's (tests/golden/data/fci12.npz,
written by tests/golden/make_fci12.py: the oracle's CSR and ARPACK at tol 1e-12);
for fci8 the CPU oracle's own solve timed in the cpu_baseline leg; for fci14
(no CSR fits) an untimed FP64 matrix-free solve at tol 1e-10, an operator the
test suite checks against the oracle at fci12. An untimed FP64 CSR solve at tol
1e-12 is logged beside them as a cross-check.

roofline: the dominant kernel of one sigma (Y = H x) of the timed operator,
with `achieved` in the algorithmic unit of one sigma: for the matrix-free
operator 2 * nnz_full * nvec flops (the nonzeros the CSR would hold) over the
kernel's average duration, against the FP64 (or FP32) MFMA peak; the kernel's own
flop rate (dense pair-integral GEMM, 2 * kpad^2 per column) is reported beside it
as kernel_mfma_frac. For the CSR: bytes 8*nnz + 8*(nrow+1) + 8*ncol... as the
kernel reports them, against 8 TB/s. Durations are HIP events on the engine's
stream.

cpu_baseline: the CPU oracle (a C restatement of the reference's serial
SparseOp::update + perform_op) on the host cores. For fci12: a bounded sample
of rows timed on one core, extrapolated to the full build plus the number of
products the reference's Lanczos (ARPACK, ncv 20, tol 1e-12, recorded in the
golden file) needs. The nbasis 8 configuration (4,900 determinants) is timed in
full: oracle CSR build plus ARPACK solve, on 1 thread and on every core.
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
MFMA_PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3}
WORKLOADS = {
    "fci12": dict(name="fci12", nbasis=12, nup=6, ndn=6, integrals="fci12", nroot=1),
    "fci8": dict(name="fci8", nbasis=8, nup=4, ndn=4, integrals="fci8", nroot=1),
    # nbasis 14: the CSR (39.8 G nonzeros) does not fit in HBM; matrix-free only
    "fci14": dict(name="fci14", nbasis=14, nup=7, ndn=7, integrals="fci14", nroot=8, csr=False),
    # 64 uneven FullCI problems (nbasis 8..12, 3..6 electrons of each spin), 2 roots each
    "batch64": dict(name="batch64", nroot=2, batch=64),
    # selected CI over 48 spin-orbitals, 12 electrons: plain and learned heat-bath trajectories
    "genci24": dict(name="genci24", nroot=1, selection=True),
}


class BatchStats(ctypes.Structure):
    _fields_ = [("graph_launches", ctypes.c_int64), ("rows_per_block", ctypes.c_int64), ("converged", ctypes.c_int64),
                ("setup_ms", ctypes.c_double), ("total_ms", ctypes.c_double), ("single_path", ctypes.c_int64)]


def bench_batch(args, wl):
    """Batched uneven solves: throughput and per-problem latency, every root
    checked against a per-problem FP64 solve of the same operator."""
    from pyci_amd.synthetic import batch_problems

    L = engine()
    ctx = ctypes.c_void_p()
    ok(L, L.pyci_ctx_create(0, ctypes.byref(ctx)), "pyci_ctx_create")
    probs = batch_problems(count=wl["batch"])
    npb = len(probs)
    nroot = args.nroot or wl["nroot"]
    i64a = lambda v: np.ascontiguousarray(v, dtype=np.int64)
    nb = i64a([q["nbasis"] for q in probs])
    nu = i64a([q["nocc"] for q in probs])
    ec = np.zeros(npb)
    ones = [np.ascontiguousarray(q["one"]) for q in probs]
    twos = [np.ascontiguousarray(q["two"]) for q in probs]
    PP = ctypes.POINTER(ctypes.c_double) * npb
    op = PP(*[dp(a) for a in ones])
    tp = PP(*[dp(a) for a in twos])
    L.pyci_batch_solve.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_double,
                                   ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(BatchStats)]
    prec = 0 if args.precision == "f64" else 1

    def run():
        ev = np.zeros((npb, nroot))
        lat = np.zeros(npb)
        it = np.zeros(npb, dtype=np.int64)
        ns = np.zeros(npb, dtype=np.int64)
        cv = np.zeros(npb, dtype=np.int32)
        st = BatchStats()
        ok(L, L.pyci_batch_solve(ctx, npb, nb.ctypes.data, nu.ctypes.data, nu.ctypes.data, ec.ctypes.data, op, tp,
                                 nroot, args.tol, prec, -1, ev.ctypes.data, lat.ctypes.data, it.ctypes.data,
                                 ns.ctypes.data, cv.ctypes.data, ctypes.byref(st)), "batch_solve")
        return ev, lat, it, ns, cv, st

    # per-problem FP64 references (untimed): the matrix-free operator, tol 1e-10
    from math import comb

    e_ref = np.zeros((npb, nroot))
    L.pyci_op_create_ex.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
                                    ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    t0 = time.perf_counter()
    for p_, q in enumerate(probs):
        n_, m_ = q["nbasis"], q["nocc"]
        ndet = comb(n_, m_) ** 2
        dets = np.zeros((ndet, 2), dtype=np.uint64)
        ok(L, L.pyci_fill_all_dets(ctx, 1, n_, m_, m_, dets.ctypes.data_as(ctypes.c_void_p)), "fill")
        ham, wfn, o = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        ok(L, L.pyci_ham_create(ctx, n_, 0.0, dp(ones[p_]), dp(twos[p_]), ctypes.byref(ham)), "ham")
        ok(L, L.pyci_wfn_create(ctx, 1, n_, m_, m_, ndet, dets.ctypes.data_as(ctypes.c_void_p), ctypes.byref(wfn)),
           "wfn")
        ok(L, L.pyci_op_create_ex(ham, wfn, -1, -1, 1, 2, ctypes.byref(o)), "op")
        ev = np.zeros(nroot)
        vec = np.zeros((nroot, ndet))
        st = Stats()
        ok(L, L.pyci_op_solve(o, nroot, None, -1, -1, 1e-10, 0, dp(ev), dp(vec), ctypes.byref(st)), "solve")
        e_ref[p_] = ev
        L.pyci_op_destroy(o)
        L.pyci_wfn_destroy(wfn)
        L.pyci_ham_destroy(ham)
    log(f"per-problem FP64 references: {time.perf_counter() - t0:.1f} s")
    for w in range(args.warmup):
        run()
    res = []
    for k in range(args.steps):
        t0 = time.perf_counter()
        ev, lat, it, ns, cv, st = run()
        t = time.perf_counter() - t0
        de = float(np.max(np.abs(ev - e_ref)))
        res.append((t, ev, lat, it, ns, cv, st, de))
        log(f"step {k}: {t*1e3:.1f} ms ({st.setup_ms:.1f} ms set-up), {int(cv.sum())}/{npb} converged, "
            f"max |dE| {de:.2e}, {st.graph_launches} graph iterations, rows/block {st.rows_per_block}, "
            f"{st.single_path} problems on the single-problem solver")
    tt = float(np.mean([r[0] for r in res]))
    lat = np.concatenate([r[2] for r in res])
    dims = [comb(q["nbasis"], q["nocc"]) ** 2 for q in probs]
    out = {
        "metric": "batched uneven FullCI solves: problems/s; per-problem latency p50/p99",
        "value": round(npb / tt, 3),
        "unit": "problems/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(tt * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == 0 else "f32-sigma+f64",
        "data": "synthetic (seeded integral scan)",
        "config": {"workload": f"{npb} FullCI problems, nbasis 8..12, 3..6 electrons of each spin, "
                               f"dimensions {min(dims)}..{max(dims)} (sum {sum(dims)}), nroot={nroot}",
                   "precision": args.precision, "tol": args.tol, "parallelism": "single GPU"},
        "latency_ms_p50": round(float(np.percentile(lat, 50)), 3),
        "latency_ms_p99": round(float(np.percentile(lat, 99)), 3),
        "setup_ms": round(float(np.mean([r[6].setup_ms for r in res])), 3),
        "iterations_max": int(max(r[6].graph_launches for r in res)),
        "single_path": int(res[-1][6].single_path),
        "rows_per_block": int(res[-1][6].rows_per_block),
        "converged": int(min(r[5].sum() for r in res)),
        "max_abs_dE": max(r[7] for r in res),
    }
    print(json.dumps(out), flush=True)


def bench_selection(args, wl):
    """Selected CI on the GenCI workload of 48 spin-orbitals and 12 electrons
    (24 spatial orbitals, 64 decaying vectors, seed 1003): from HF + singles +
    doubles (42,013 determinants), add_hci -> sparse_op -> solve at eps = 1e-3,
    5e-4, 2e-4, 1e-4, 5e-5, 2e-5, 1e-5, stopping once the set holds max_det
    determinants or the next CSR would exceed the HBM budget. Three runs: plain
    heat-bath selection, and the learned selector (network retrained on each CI
    vector) ranking a pool 5x looser than eps, by the network alone (score mode 0)
    and by network + log10 heat-bath score (1), each appending per step as many
    determinants as plain selection at that eps would. Prints the (ndet, energy)
    trajectories and the time of each step."""
    import pyci_amd as pyci
    from pyci_amd.selection import hartreefock_genci, run_selection
    from pyci_amd.synthetic import spin_orbital

    _, h, g = config_integrals("genci24")
    one, two = spin_orbital(h, g)
    ham = pyci.hamiltonian(0.0, one, two)
    hf = hartreefock_genci(24, 6, 6)
    schedule = [1e-3, 5e-4, 2e-4, 1e-4, 5e-5, 2e-5, 1e-5]
    max_det = int(os.environ.get("PYCI_BENCH_MAX_DET", 5_000_000))
    runs = {}
    for name, kw in (("heat-bath", {}), ("learned-net", dict(score_mode=0)), ("learned-net+hb", dict(score_mode=1))):
        wfn = pyci.genci_wfn(48, 12, 0)
        pyci.add_excitations(wfn, 0, 1, 2, ref=hf)
        sel = pyci.mlp_selector(48, seed=0) if kw else None
        steps = []
        t_last = [time.perf_counter()]

        def cb(t):
            now = time.perf_counter()
            t["seconds"] = round(now - t_last[0], 3)
            t_last[0] = now
            steps.append(t)
            log(f"{name}: {t['ndet']} dets, E {t['energy']:.10f} (eps {t['eps']:.0e}, +{t['added']}, "
                f"{t['seconds']:.1f} s)")

        t0 = time.perf_counter()
        traj = run_selection(ham, wfn, schedule, selector=sel, pool=5.0, train_steps=300, batch=8192,
                             tol=1e-8, max_det=max_det, callback=cb, **kw)
        runs[name] = dict(trajectory=[dict(ndet=t["ndet"], energy=round(t["energy"], 10), eps=t["eps"],
                                           seconds=t.get("seconds")) for t in traj],
                          total_s=round(time.perf_counter() - t0, 3))
    hb = runs["heat-bath"]
    out = {
        "metric": "selected CI (GenCI 48 spin-orbitals, 12 e-): plain vs learned heat-bath trajectories",
        "value": hb["total_s"],
        "unit": "s",
        "n_gpus": 1,
        "steps": len(hb["trajectory"]) - 1,
        "warmup": 0,
        "ms_per_step": round(1e3 * hb["total_s"] / max(1, len(hb["trajectory"]) - 1), 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded integrals, seed 1003)",
        "config": {"workload": "genci_wfn(48, 12, 0), HF+SD start, eps 1e-3 .. 1e-5", "max_det": max_det,
                   "parallelism": "single GPU"},
        "runs": runs,
    }
    print(json.dumps(out), flush=True)


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
    L.pyci_op_create_ex.argtypes = [vp, vp, i64, i64, ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]
    L.pyci_op_storage.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    L.pyci_op_bench_sigma_parts.argtypes = [vp, i64, i64, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                            ctypes.POINTER(ctypes.c_float), dp, dp]
    L.pyci_op_destroy.argtypes = [vp]
    L.pyci_wfn_destroy.argtypes = [vp]
    L.pyci_ham_destroy.argtypes = [vp]
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


def measured_traffic(fname):
    """HBM bytes per launch of a kernel from a committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary under profiles/, or None."""
    path = os.path.join(ROOT, "profiles", fname)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("| traffic = read + write |"):
            return float(line.split("|")[2].split()[0]) * 1e9
    return None


def device_copy(nbytes):
    """One device-to-device copy of nbytes through the HIP runtime (a plain copy
    kernel of known size, used to calibrate the HBM byte counters)."""
    hip = ctypes.CDLL("libamdhip64.so.7")  # by SONAME: the runtime the engine is bound to
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


def cpu_full_fci8(nthread):
    """configs[0] on the host, in full: the oracle's CSR build (row ranges on
    `nthread` threads) and ARPACK on its symmetric product (`nthread` threads).
    Returns (build s, solve s, energy, products)."""
    from pyci_amd.synthetic import config_integrals as ci
    from oracle import oracle as O
    from tests.golden import oracle_big as B
    import scipy.sparse.linalg as spl

    ecore, one, two = ci("fci8")
    dets = O.all_dets("fullci", 8, 4, 4)
    t0 = time.perf_counter()
    csr = B.build_lower("fullci", 8, 4, 4, dets, one, two, nthread=nthread, chunk_rows=max(64, 4900 // nthread),
                        log=None)
    t1 = time.perf_counter()
    cnt = [0]

    def mv(v):
        cnt[0] += 1
        return csr.symv(np.asarray(v).reshape(-1), nthread=nthread)

    A = spl.LinearOperator((csr.n, csr.n), matvec=mv, dtype=np.float64)
    w = spl.eigsh(A, k=1, which="SA", tol=1e-12, ncv=20, v0=np.random.default_rng(12345).standard_normal(csr.n),
                  return_eigenvectors=False)
    t2 = time.perf_counter()
    return t1 - t0, t2 - t1, float(w[0]) + ecore, cnt[0]


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
    # the same rows built on every host thread of this GPU's share (row ranges per
    # thread); the threaded product is not sampled this way (a contiguous row sample
    # lands on one thread of a row-partitioned product), its full-space timing is
    # the nbasis 8 leg
    from concurrent.futures import ThreadPoolExecutor

    nth = host_threads()
    step = (rows + nth - 1) // nth
    ranges = [(r0 + a, min(r0 + rows, r0 + a + step)) for a in range(0, rows, step)]

    def part(rr):
        O.csr_build("fullci", wl["nbasis"], wl["nup"], wl["ndn"], dets, one, two, nrow=rr[1], symmetric=True,
                    row0=rr[0])

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=nth) as ex:
        list(ex.map(part, ranges))
    t_build_mt = time.perf_counter() - t0
    # every thread rebuilt the determinant index: count it once, as a full build would
    t_row_mt = max(t_build_mt - t_index, 1e-9) / rows
    return dict(t_row=t_row, t_index=t_index, rows=rows, t_spmv_sample=t_spmv_rows, nnz_sample=nnz_sample,
                sigma_calls=sigma_calls, nthreads=nth, t_row_mt=t_row_mt)


def host_threads():
    """Host threads of one GPU's share of the box (16), or fewer where fewer cores exist."""
    return max(1, min(os.cpu_count() or 1, 16))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="fci12", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="f64", choices=["f64", "mixed"])
    ap.add_argument("--storage", default="auto", choices=["auto", "csr", "direct"])
    ap.add_argument("--nroot", type=int, default=0, help="roots (default: the workload's)")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--cpu-seconds", type=float, default=16.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--nvec", type=int, default=1, help="vectors per sigma in --sigma-only runs")
    ap.add_argument("--sigma-only", action="store_true",
                    help="build the operator, run a few sigma launches and a device copy of known size "
                         "(for hardware-counter runs), print their timings and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    wl = WORKLOADS[args.workload]
    if "batch" in wl:
        bench_batch(args, wl)
        return
    if wl.get("selection"):
        bench_selection(args, wl)
        return
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
    nroot = args.nroot or wl["nroot"]
    has_csr = wl.get("csr", True)

    storage_code = {"csr": 0, "direct": 1, "auto": 2}

    def build(storage=args.storage):
        op = ctypes.c_void_p()
        ok(L, L.pyci_op_create_ex(ham, wfn, -1, -1, 1, storage_code[storage], ctypes.byref(op)), "op_create")
        return op

    def storage_of(op):
        st = ctypes.c_int()
        ok(L, L.pyci_op_storage(op, ctypes.byref(st)), "op_storage")
        return "direct" if st.value == 1 else "csr"

    def parts(op, nvec, nrep, fp32):
        n = ctypes.c_int()
        ms = (ctypes.c_float * 4)()
        by = (ctypes.c_double * 4)()
        fl = (ctypes.c_double * 4)()
        ok(L, L.pyci_op_bench_sigma_parts(op, nvec, nrep, fp32, ctypes.byref(n), ms,
                                          ctypes.cast(by, ctypes.POINTER(ctypes.c_double)),
                                          ctypes.cast(fl, ctypes.POINTER(ctypes.c_double))), "bench_sigma_parts")
        return [dict(ms=ms[i], bytes=by[i], flops=fl[i]) for i in range(n.value)]

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
        pr = parts(op, args.nvec, 4, 1 if args.precision == "mixed" else 0)
        copy_bytes = 4 << 30
        t_copy = device_copy(copy_bytes)
        print(json.dumps({"storage": storage_of(op), "nvec": args.nvec, "parts": pr, "nnz": nnz, "copy_bytes": copy_bytes,
                          "copy_ms": t_copy * 1e3}), flush=True)
        return
    # FP64 reference energies (untimed, tight tolerance): from the CSR operator,
    # or from the matrix-free FP64 operator where the CSR cannot be stored
    op = build("csr" if has_csr else "direct")
    nrow, ncol, size_ref, nnz = info(op)
    e_gpu, st_ref = solve(op, 0, 1e-12 if has_csr else 1e-10)
    e_gpu = np.sort(e_gpu)
    log(f"GPU FP64 energies at tight tolerance {np.array2string(e_gpu, precision=10)} ({st_ref.iterations} "
        f"iterations); stored nnz {nnz}, reference-format nnz {size_ref}")
    gold_path = os.path.join(ROOT, "tests", "golden", "data", f"{wl['name']}.npz")
    cpu8 = None
    lanczos_products = None
    if os.path.exists(gold_path):
        with np.load(gold_path) as g:
            e_ref = np.asarray(g["e8"][:nroot]) + ecore
            if "lanczos_products_k1" in g:
                lanczos_products = int(g["lanczos_products_k1"])
        e_src = "CPU oracle (tests/golden/data/%s.npz: oracle CSR + ARPACK, tol 1e-12)" % wl["name"]
    elif wl["name"] == "fci8" and not args.no_cpu:
        cpu8 = {n: cpu_full_fci8(n) for n in (1, host_threads())}
        e_ref = np.array([cpu8[1][2]])
        e_src = "CPU oracle solve timed in cpu_baseline (oracle CSR + ARPACK, tol 1e-12)"
    else:
        e_ref = e_gpu
        e_src = ("GPU FP64 CSR solve at tol 1e-12" if has_csr else
                 "GPU FP64 matrix-free solve at tol 1e-10 (operator checked against the oracle at nbasis 12)")
    log(f"reference energies {np.array2string(e_ref, precision=12)} from {e_src}; GPU tight-tolerance solve "
        f"differs by {np.max(np.abs(e_gpu - e_ref)):.2e}")
    csr_parts = None
    if has_csr:
        # CSR sigma (HIP events on the engine stream): the plain path's kernel
        csr_parts = {"f64": parts(op, 1, 20, 0)[0], "f32": parts(op, 1, 20, 1)[0],
                     "f64_nvec8": parts(op, 8, 10, 0)[0]}
        for k, v in csr_parts.items():
            log(f"CSR sigma {k}: {v['ms']:.3f} ms, {v['bytes'] / v['ms'] / 1e6:.0f} GB/s")
    L.pyci_op_destroy(op)

    def step(storage, p):
        ok(L, L.pyci_ctx_synchronize(ctx), "sync")
        t0 = time.perf_counter()
        o = build(storage)
        ok(L, L.pyci_ctx_synchronize(ctx), "sync")
        t1 = time.perf_counter()
        e, st = solve(o, p, args.tol)
        ok(L, L.pyci_ctx_synchronize(ctx), "sync")
        t2 = time.perf_counter()
        L.pyci_op_destroy(o)
        return t1 - t0, t2 - t1, e, st

    base = None
    if has_csr:
        # plain FP64 GPU path: CSR build + FP64 solve (one warm-up, then timed)
        step("csr", 0)
        base = [step("csr", 0) for _ in range(2)]
        t_base = float(np.mean([b[0] + b[1] for b in base]))
        log(f"baseline CSR f64: {t_base*1e3:.1f} ms (build {np.mean([b[0] for b in base])*1e3:.1f} ms, "
            f"{base[0][3].iterations} iterations), |dE| {np.max(np.abs(np.sort(base[0][2]) - e_ref)):.2e}")
    # the timed operator's sigma kernels (one vector per root block column)
    op = build()
    storage = storage_of(op)
    fp32 = 1 if args.precision == "mixed" else 0
    tparts = parts(op, nroot, 20 if nroot == 1 else 3, fp32)
    L.pyci_op_destroy(op)
    # one vector with nup == ndn: the parity-sector sigma the solver runs (live-row product,
    # alpha-term scatter + symmetrisation)
    sector = storage == "direct" and nroot == 1 and nu == nd_
    names = ["k_spmv (CSR sigma)"] if storage == "csr" else [
        "k_mf_dprod (gather + pair-integral product, MFMA)",
        "k_mf_alpha + k_mf_symm_tiles (parity-sector scatter)" if sector else "k_mf_sigma (scatter pass)"]
    for nm, v in zip(names, tparts):
        log(f"{storage} sigma part {nm}: {v['ms']:.3f} ms, {v['bytes'] / v['ms'] / 1e6:.0f} GB/s, "
            f"{v['flops'] / v['ms'] / 1e9:.1f} TFLOP/s")

    for w in range(args.warmup):
        tb, ts, e, st = step(args.storage, prec)
        log(f"warmup {w}: build {tb*1e3:.1f} ms, solve {ts*1e3:.1f} ms")
    ok(L, L.pyci_ctx_synchronize(ctx), "sync")
    res = []
    t_all0 = time.perf_counter()
    for k in range(args.steps):
        tb, ts, e, st = step(args.storage, prec)
        de = float(np.max(np.abs(np.sort(e) - e_ref)))
        res.append((tb, ts, de, st.iterations, st.sigma_calls))
        log(f"step {k}: build {tb*1e3:.1f} ms, solve {ts*1e3:.1f} ms, |dE| {de:.2e}, {st.iterations} iterations")
    ok(L, L.pyci_ctx_synchronize(ctx), "sync")
    t_all = time.perf_counter() - t_all0
    tts = t_all / args.steps
    # the other precision of the same operator, timed alike (reported beside the line)
    other = None
    if not args.sigma_only:
        po = 1 - prec
        step(args.storage, po)
        ores = [step(args.storage, po) for _ in range(max(1, args.steps))]
        other = {"precision": "mixed" if po == 1 else "f64",
                 "tts_s": round(float(np.mean([r[0] + r[1] for r in ores])), 6),
                 "max_abs_dE": float(max(np.max(np.abs(np.sort(r[2]) - e_ref)) for r in ores)),
                 "iterations": int(ores[0][3].iterations)}
        log(f"{other['precision']}: {other['tts_s']*1e3:.1f} ms per step, |dE| {other['max_abs_dE']:.2e}")
    tb = float(np.mean([r[0] for r in res]))
    ts = float(np.mean([r[1] for r in res]))
    max_de = max(r[2] for r in res)
    sigma_calls = int(np.mean([r[4] for r in res]))
    # roofline of the dominant kernel of the timed operator's sigma
    dom = int(np.argmax([v["ms"] for v in tparts]))
    dk = tparts[dom]
    pk = "f32" if fp32 else "f64"
    ridge = MFMA_PEAK_TFLOPS[pk] * 1e12 / (HBM_PEAK_GBPS * 1e9)
    kname = names[dom]
    slug = "sigma_csr" if storage == "csr" else ["mf_dprod", "mf_scatter"][dom]
    traffic = measured_traffic(f"{slug}_{wl['name']}_{pk}_hbm_traffic.md")
    # nonzeros of the full-pattern CSR: every single and double excitation of a row
    # plus the diagonal (the full FullCI space holds them all)
    na_, va_, nb_, vb_ = nu, nb - nu, nd_, nb - nd_
    per_row = (na_ * va_ + nb_ * vb_ + comb(na_, 2) * comb(va_, 2) + comb(nb_, 2) * comb(vb_, 2) +
               na_ * va_ * nb_ * vb_ + 1)
    nnz_full = per_row * nd
    if storage == "direct":
        alg_flops = 2.0 * nnz_full * nroot
        achieved = alg_flops / (dk["ms"] * 1e-3) / 1e12
        kern = dk["flops"] / (dk["ms"] * 1e-3) / 1e12
        roofline = {"bound": "mfma", "achieved": round(achieved, 2), "peak": MFMA_PEAK_TFLOPS[pk], "unit": "TFLOP/s",
                    "frac": round(achieved / MFMA_PEAK_TFLOPS[pk], 4),
                    "algorithmic_flops_per_launch": alg_flops,
                    "kernel_mfma_tflops": round(kern, 2),
                    "kernel_mfma_frac": round(kern / MFMA_PEAK_TFLOPS[pk], 4)}
    elif dk["flops"] / dk["bytes"] > ridge:
        achieved = dk["flops"] / (dk["ms"] * 1e-3) / 1e12
        roofline = {"bound": "mfma", "achieved": round(achieved, 2), "peak": MFMA_PEAK_TFLOPS[pk], "unit": "TFLOP/s",
                    "frac": round(achieved / MFMA_PEAK_TFLOPS[pk], 4)}
    else:
        achieved = dk["bytes"] / (dk["ms"] * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBPS, 4)}
    roofline.update({"traffic": traffic, "kernel": f"{kname}, {pk}", "avg_ms": round(dk["ms"], 4),
                     "bytes_per_launch": dk["bytes"], "flops_per_launch": dk["flops"]})

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
        "config": {"workload": f"FullCI nbasis={nb} {nu}a+{nd_}b seeded integrals ({wl['integrals']}), nroot={nroot}",
                   "ndet": nd, "nnz_stored": nnz, "nnz_reference_format": size_ref, "parallelism": "single GPU",
                   "precision": args.precision, "tol": args.tol},
        "storage": storage,
        "build_ms": round(tb * 1e3, 3),
        "solve_ms": round(ts * 1e3, 3),
        "energy": [float(x) for x in e_ref],
        "energy_source": e_src,
        "max_abs_dE": max_de,
        "iterations": int(np.mean([r[3] for r in res])),
        "sigma_calls": sigma_calls,
        "sigma_parts_ms": {nm: round(v["ms"], 4) for nm, v in zip(names, tparts)},
        "other_precision": other,
        "roofline": roofline,
    }
    if has_csr:
        tbase_build = float(np.mean([b[0] for b in base]))
        sig = csr_parts["f64"]
        out.update({
            "baseline_csr_f64": {"tts_s": round(t_base, 6), "build_ms": round(tbase_build * 1e3, 3),
                                 "iterations": int(base[0][3].iterations), "speedup": round(t_base / tts, 2)},
            "dets_per_s_built": round(nd / tbase_build, 1),
            "nnz_per_s_built": round(nnz / tbase_build, 1),
            "sigma_gbps": round(sig["bytes"] / (sig["ms"] * 1e-3) / 1e9, 1),
            "sigma_f32_gbps": round(csr_parts["f32"]["bytes"] / (csr_parts["f32"]["ms"] * 1e-3) / 1e9, 1),
            "sigma_nvec8_ms": round(csr_parts["f64_nvec8"]["ms"], 4),
        })
    if not args.no_cpu and has_csr:
        log("cpu baseline: oracle on a bounded row sample ...")
        # products of the reference's Lanczos (ARPACK, ncv 20, tol 1e-12) where the
        # golden file records them, else the plain GPU path's count
        cpu_sigma = lanczos_products or int(base[0][3].sigma_calls)
        cb = cpu_baseline(wl, one, two, dets, cpu_sigma, args.cpu_seconds)
        frac = size_ref / max(cb["nnz_sample"], 1)  # full lower nnz / sample lower nnz
        t_build_cpu = cb["t_index"] + cb["t_row"] * nd
        t_spmv_cpu = cb["t_spmv_sample"] * frac
        t_cpu = t_build_cpu + cpu_sigma * t_spmv_cpu
        out["cpu_baseline"] = {
            "value": round(t_cpu, 3),
            "unit": "s",
            "cores": 1,
            "kind": "port",
            "sample": (f"oracle CSR build of {cb['rows']} middle rows of {nd} ({cb['t_row']*1e6:.1f} us/row, "
                       f"index {cb['t_index']:.2f} s) and serial lower-triangle SpMV over them "
                       f"({cb['t_spmv_sample']*1e3:.1f} ms for {cb['nnz_sample']} nnz), extrapolated to the full "
                       f"build + {cpu_sigma} sigma products"),
            "dets_per_s_built": round(1.0 / cb["t_row"], 1),
            "extrapolated": True,
            "sample_fraction": round(cb["rows"] / nd, 4),
        }
        out["cpu_baseline"][f"dets_per_s_built_threads_{cb['nthreads']}"] = round(1.0 / cb["t_row_mt"], 1)
        if cpu8 is None:
            cpu8 = {n: cpu_full_fci8(n) for n in (1, host_threads())}
        out["cpu_baseline"]["fci8_full"] = {
            "workload": "FullCI nbasis=8 4a+4b seeded integrals (fci8), 4,900 determinants, nroot=1, timed in full",
            **{f"threads_{n}": {"build_s": round(v[0], 4), "solve_s": round(v[1], 4), "total_s": round(v[0] + v[1], 4),
                                "products": v[3], "energy": v[2]} for n, v in cpu8.items()},
            "cores_available": os.cpu_count(),
        }
        log(f"cpu baseline estimate {t_cpu:.1f} s (1 thread; build {1.0 / cb['t_row']:.0f} dets/s, on "
            f"{cb['nthreads']} threads {1.0 / cb['t_row_mt']:.0f} dets/s); fci8 in full: " +
            ", ".join(f"{n} threads {v[0] + v[1]:.3f} s" for n, v in cpu8.items()))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
