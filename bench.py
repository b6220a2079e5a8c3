"""Benchmark of the hot path: H*psi on the roofline workload.

Workload: XXZ ring L=32, Delta=1 (Jxy=1, S = sigma/2 convention), in the
TranslationParityBasis N=16, k=0, p=+1 sector (D = 9 398 520), ComplexF64
vectors from a seeded Gaussian.  One step = one H*psi (the gather kernel over
all rows, zero-initialised target as in H*psi / _apply!), inputs resident in
HBM.  Prints one JSON line.

roofline: algorithmic bytes per matvec
    B_alg = D * (s_I + s_R + 2 s_v) + NNZ_off * (s_v + s_I)
(s_I = 8 representative, s_R = 8 norm, s_v = 16 ComplexF64; NNZ_off counted
exactly on the device) over the kernel's mean duration, timed with HIP events
on the stream the kernel runs on.  roofline.traffic: HBM bytes per launch from
the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary of the same kernel and
workload (profiles/matvec_rk_L32_tp_k0.md).  cpu_baseline: the oracle's threaded scatter
mul on a strided sample of columns of the same matvec, scaled to H*psi/s.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "edkit.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0
L, NUP, DELTA = 32, 16, 1.0
# rocprofv3 FETCH_SIZE/WRITE_SIZE summary of the default matvec kernel on this
# workload, scaled by the copy-probe calibration (a measured figure, not live)
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "matvec_rk_L32_tp_k0.md")


# the sources the default matvec kernel (k_matvec_rk) and its setup are compiled from
KERNEL_SOURCES = [os.path.join(ROOT, "edkit.jl_amd", "csrc", f)
                  for f in ("edk_fast_rk.hip", "edk_fast.hip", "edk_fast.h", "edk_internal.h")]


# the sources of the Abelian two-site kernel (k_matvec_perm) and the definitions it uses
PERM_SOURCES = [os.path.join(ROOT, "edkit.jl_amd", "csrc", f)
                for f in ("edk_perm.hip", "edk_fast.h", "edk_internal.h")]


def kernel_source_hash(paths=KERNEL_SOURCES):
    """sha256 over the kernel's source files (the key a profile summary records)."""
    import hashlib
    h = hashlib.sha256()
    for p in paths:
        with open(p, "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def profiled_traffic(path=TRAFFIC_PROFILE, sources=KERNEL_SOURCES):
    """HBM bytes per launch recorded in a committed rocprofv3 summary, or None
    when the summary was taken on other kernel sources than the ones built."""
    traffic = key = None
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch ="):
                    traffic = int(line.split("=")[1])
                elif line.startswith("kernel_sources_sha256 ="):
                    key = line.split("=")[1].strip()
    except OSError:
        return None
    return traffic if key is not None and key == kernel_source_hash(sources) else None


class HipEvents:
    """hipEvent timing on a given stream (ctypes over libamdhip64)."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.h.hipEventSynchronize.argtypes = [C.c_void_p]
        self.h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.h.hipEventDestroy.argtypes = [C.c_void_p]

    def event(self):
        e = C.c_void_p()
        assert self.h.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.h.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, a, b):
        assert self.h.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.h.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return ms.value

    def device_sync(self):
        assert self.h.hipDeviceSynchronize() == 0


def time_matvec(edk, H, x, y, steps, warmup, ev, stream):
    from edkit_hip.matvec import matvec_into
    for _ in range(warmup):
        matvec_into(y, H, x)
    ev.device_sync()
    e0, e1 = ev.event(), ev.event()
    t0 = time.perf_counter()
    ev.record(e0, stream)
    for _ in range(steps):
        matvec_into(y, H, x)
    ev.record(e1, stream)
    ev.device_sync()
    wall = time.perf_counter() - t0
    return ev.elapsed_ms(e0, e1) / steps, wall * 1e3 / steps


def host_threads():
    """Threads for the CPU legs: every core this process may run on, at most
    16 (a one-GPU box grants its jobs 16 cores whatever nproc reports)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(H, x, budget_s=12.0):
    """Oracle (C restatement of EDKit's threaded mul) on a strided column sample."""
    from oracle import oracle as orc
    nthreads = host_threads()
    t0 = time.perf_counter()
    ob = orc.Basis("translation_parity", L, N=NUP, k=0, p=1, nthreads=nthreads)
    t_basis = time.perf_counter() - t0
    op = orc.Op(ob, H.flatten())
    D = ob.D
    tgt = np.zeros(D, np.complex128)

    def run(n):
        cols = np.linspace(0, D - 1, n).astype(np.int64)
        tgt[:] = 0
        t0 = time.perf_counter()
        op.mul_cols_(tgt, x, cols, nthreads)
        return time.perf_counter() - t0

    # a call costs a fixed part (private buffers, their final sum) plus a
    # per-column part: fit both from two sample sizes, the larger one sized
    # to the time budget, and scale the per-column part to all D columns
    n1 = 20000
    t1 = run(n1)
    t2 = run(4 * n1)
    per_col = max((t2 - t1) / (3 * n1), 1e-9)
    n3 = int(min(D, max(4 * n1, budget_s / per_col)))
    t3 = run(n3)
    per_col = max((t3 - t1) / (n3 - n1), 1e-9)
    fixed = max(t1 - per_col * n1, 0.0)
    per_matvec = fixed + per_col * D
    return {"value": 1.0 / per_matvec, "unit": "H*psi/s", "cores": nthreads, "kind": "port",
            "host_nproc": os.cpu_count(),
            "sample": f"{n3} of {D} evenly spaced columns of one L=32 TP k=0 p=+1 ComplexF64 "
                      f"H*psi (scatter colmn! per column, threaded with private buffers) took "
                      f"{t3:.1f} s; fixed {fixed:.2f} s + {per_col * 1e6:.2f} us/column -> "
                      f"{per_matvec:.1f} s per full H*psi; oracle table build {t_basis:.1f} s"}


def lanczos_leg(edk, H, xh, m, D, b_alg):
    """Ground-state Lanczos of the same sector from the same seeded psi: m steps
    with full reorthogonalisation (classical Gram-Schmidt against every basis
    vector, repeated when the first sweep removed most of w), ComplexF64, then
    ComplexF32 vectors with FP64 scalars and a final FP64 refinement."""
    from edkit_hip.krylov import lanczos_ground
    res = {"steps": m}
    for name, dt, refine in (("c64", np.complex128, False), ("c32_refined", np.complex64, True)):
        x0 = edk.DeviceArray.from_host(xh.astype(dt))
        t0 = time.perf_counter()
        E, r, steps, _ = lanczos_ground(H, x0, m=m, tol=0.0, dtype=dt, refine=refine)
        edk.default_context().synchronize()
        secs = time.perf_counter() - t0
        s_v = 16 if dt == np.complex128 else 8
        # one Gram-Schmidt sweep per step (one pass over V for V^H w, one for
        # w -= V h); the second sweep runs only when the first removed most of
        # w, so this is a lower bound on the bytes moved
        reorth = sum(2 * j * D * s_v for j in range(1, steps + 1))
        b32 = D * (8 + 8 + 2 * s_v) + (b_alg - D * 48) // 24 * (s_v + 8)
        res[name] = {"E0": E, "resid": r, "steps": steps, "s": secs,
                     "bytes": reorth + steps * (b32 if s_v == 8 else b_alg),
                     "GBs": (reorth + steps * (b32 if s_v == 8 else b_alg)) / secs / 1e9}
        del x0
    res["dE_c32_vs_c64"] = abs(res["c32_refined"]["E0"] - res["c64"]["E0"])
    # the same run with partial reorthogonalisation (the serving default):
    # the whole basis is swept only at the steps the omega recurrence flags
    x0 = edk.DeviceArray.from_host(xh)
    info = {}
    t0 = time.perf_counter()
    E, r, steps, _ = lanczos_ground(H, x0, m=m, tol=0.0, dtype=np.complex128, refine=False, reorth="partial",
                                    info=info)
    edk.default_context().synchronize()
    secs = time.perf_counter() - t0
    res["c64_partial"] = {"E0": E, "resid": r, "steps": steps, "s": secs, "reorth_steps": info["reorth_steps"],
                          "dE_vs_full": abs(E - res["c64"]["E0"])}
    del x0
    return res


def c4_workload(args, ev):
    """Batched uneven sectors: J1-J2 ring L=28 J2=0.3 N=14, all 28 momentum
    sectors, plus the spin-1 BLBQ chain L=16 beta=0.2 Sz=0 k=0; a block of 8
    seeded complex vectors per sector, ComplexF32 matvec.  One step = H_s X_s
    for every sector s.  Timed: one vector at a time (8 x 29 launches), one
    block launch per sector (29), one graph with a launch per sector, one
    graph with the sectors of one block kernel fused into launches of up to
    12 sectors (the 26 complex-momentum sectors share three), the same after autotuning each
    sector's schedule; then once more with every sector's matrix assembled and
    cached (sparse! semantics: block products through CSR SpMM).  value is the
    fastest matrix-free graph mode (the operator applied without a stored
    matrix, as mul! does); the assembled graph is reported beside it."""
    import edkit_hip as edk
    from edkit_hip.matvec import device_operator, matvec_into
    sys.path.insert(0, ROOT)
    from tests.models import blbq, j1j2
    nb, dt = 8, np.complex64
    t0 = time.perf_counter()
    sectors = [("j1j2 L=28 N=14 k=%d" % k, j1j2(edk.TranslationalBasis(L=28, N=14, k=k), 0.3))
               for k in range(28)]
    sectors.append(("blbq L=16 Sz=0 k=0", blbq(edk.TranslationalBasis(L=16, N=16, k=0, base=3), 0.2)))
    ev.device_sync()
    t_build = time.perf_counter() - t0
    rng = np.random.default_rng(4)
    items, b_alg, dims = [], 0, []
    for name, H in sectors:
        D = H.B.D
        dims.append(D)
        X = (rng.standard_normal((D, nb)) + 1j * rng.standard_normal((D, nb))).astype(dt)
        Xd = edk.DeviceArray.from_host(X)
        Yd = edk.DeviceArray((D, nb), dt)
        cols = [edk.DeviceArray.from_host(np.ascontiguousarray(X[:, c])) for c in range(nb)]
        ycol = edk.DeviceArray((D,), dt)
        items.append((H, Xd, Yd, cols, ycol))
        nnz = device_operator(H).count_offdiag()
        b_alg += D * (8 + 8 + 2 * nb * 8) + nnz * (nb * 8 + 8)
    stream = sectors[0][1].B.ctx.stream

    def timed(fn, steps):
        for _ in range(args.warmup):
            fn()
        ev.device_sync()
        e0, e1 = ev.event(), ev.event()
        ev.record(e0, stream)
        for _ in range(steps):
            fn()
        ev.record(e1, stream)
        return ev.elapsed_ms(e0, e1) / steps

    def one_at_a_time():
        for H, _, _, cols, ycol in items:
            for c in cols:
                matvec_into(ycol, H, c)

    def per_sector_block():
        for H, Xd, Yd, _, _ in items:
            matvec_into(Yd, H, Xd)

    steps = max(1, args.steps // 10)
    res = {}
    res["one_vector_at_a_time_ms"] = timed(one_at_a_time, steps)
    res["block_per_sector_ms"] = timed(per_sector_block, steps)
    batch = edk.MatvecBatch([(H, Xd, Yd) for H, Xd, Yd, _, _ in items], streams=4)
    res["graph_4_branches_ms"] = timed(batch.launch, steps)
    batch.close()
    batch = edk.MatvecBatch([(H, Xd, Yd) for H, Xd, Yd, _, _ in items], fuse=False)
    res["graph_launch_per_sector_ms"] = timed(batch.launch, steps)
    res["graph_launch_per_sector_launches"] = batch.launches
    batch.close()
    batch = edk.MatvecBatch([(H, Xd, Yd) for H, Xd, Yd, _, _ in items])
    res["graph_ms"] = timed(batch.launch, steps)
    res["graph_launches"] = batch.launches
    batch.close()
    tuned = {}
    for H, _, _, _, _ in items:
        v, _ = device_operator(H).autotune(dt, nb, reps=3)
        tuned[v] = tuned.get(v, 0) + 1
    batch = edk.MatvecBatch([(H, Xd, Yd) for H, Xd, Yd, _, _ in items])
    res["graph_autotuned_ms"] = timed(batch.launch, steps)
    batch.close()
    # the assembled matrices (sparse! semantics): block products through CSR SpMM
    ev.device_sync()
    t0 = time.perf_counter()
    for H, _, _, _, _ in items:
        device_operator(H).assemble()
    ev.device_sync()
    res["assemble_all_s"] = time.perf_counter() - t0
    res["assembled_block_per_sector_ms"] = timed(per_sector_block, steps)
    batch = edk.MatvecBatch([(H, Xd, Yd) for H, Xd, Yd, _, _ in items])
    res["graph_assembled_ms"] = timed(batch.launch, steps)
    batch.close()
    for H, _, _, _, _ in items:
        device_operator(H).release_csr()
    from edkit_hip._lib import VARIANT_NAMES
    modes = {"graph_ms": "one graph, matrix-free (sectors of one block kernel fused, up to 12 a launch)",
             "graph_launch_per_sector_ms": "one graph, matrix-free, one launch per sector",
             "graph_autotuned_ms": "one graph, matrix-free, autotuned schedules"}
    best_mode = min(modes, key=lambda k: res[k])
    best = res[best_mode]
    out = {
        "metric": "batched sector matvecs/s (29 sectors x 8 vectors, ComplexF32)",
        "value": len(items) * nb / (best * 1e-3),
        "unit": "sector-vector H*psi/s",
        "n_gpus": 1, "steps": steps, "warmup": args.warmup, "ms_per_step": best,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "complex64",
        "data": "synthetic (seeded Gaussian blocks)",
        "config": {"workload": "J1-J2 ring L=28 J2=0.3 N=14 k=0..27 + spin-1 BLBQ L=16 beta=0.2 Sz=0 k=0, "
                               "block of 8, ComplexF32", "dims": dims, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": b_alg / (best * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": b_alg / (best * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                     "bytes_alg_per_step": b_alg},
        "modes_ms": res,
        "value_mode": modes[best_mode],
        "matrix_free": {"ms": min(res["graph_ms"], res["graph_launch_per_sector_ms"], res["graph_autotuned_ms"]),
                        "frac": b_alg / (min(res["graph_ms"], res["graph_launch_per_sector_ms"],
                                             res["graph_autotuned_ms"]) * 1e-3) / 1e9 / HBM_PEAK_GBS,
                        "fused_vs_launch_per_sector": res["graph_launch_per_sector_ms"] / res["graph_ms"]},
        "assembled": {"ms": res["graph_assembled_ms"], "assemble_all_s": res["assemble_all_s"],
                      "frac": b_alg / (res["graph_assembled_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                      "note": "one graph over the sparse! matrices (CSR SpMM); frac against the matrix-free bytes"},
        "speedup_graph_vs_one_at_a_time": res["one_vector_at_a_time_ms"] / best,
        "autotuned_schedules": {VARIANT_NAMES[k]: v for k, v in tuned.items()},
        "basis_build_s": t_build,
    }
    print(json.dumps(out))


def c5_stream(n, rate, seed):
    """Seeded serving stream: Poisson arrivals at `rate` req/s; jobs uniform
    over spin-1/2 L in {18,...,30} (XXZ Delta in [0.5, 2] or J1-J2 J2 in
    [0, 0.6]) and spin-1 BLBQ L in {12, 14, 16} (beta in [0, 0.3]), Sz = 0,
    random momentum k (and parity p at k = 0, L/2); 60 % ground state to
    residual 1e-9, 40 % Krylov evolution to t in [1, 20]; each job starts from
    the product state of a seeded random representative of its sector."""
    rng = np.random.default_rng(seed)
    t, jobs = 0.0, []
    for _ in range(n):
        t += rng.exponential(1.0 / rate)
        fam = int(rng.integers(10))
        if fam < 7:
            L, base = 18 + 2 * fam, 2
            model = ("xxz", float(rng.uniform(0.5, 2.0))) if rng.random() < 0.5 else \
                ("j1j2", float(rng.uniform(0.0, 0.6)))
        else:
            L, base = 12 + 2 * (fam - 7), 3
            model = ("blbq", float(rng.uniform(0.0, 0.3)))
        k = int(rng.integers(L))
        p = int(rng.choice([1, -1])) if k in (0, L // 2) else 0
        kind = "ground" if rng.random() < 0.6 else "evolve"
        jobs.append(dict(arrival=t, L=L, base=base, model=model, k=k, p=p, kind=kind,
                         t=float(rng.uniform(1.0, 20.0)), start=float(rng.random())))
    return jobs


def _pctl(v, q):
    return float(np.percentile(np.asarray(v), q)) if len(v) else None


def c5_run(n_requests, rate, workers, wait_us=250.0, legs=("plain", "single", "served"), log=None):
    """The serving stream three ways, on the same seeded jobs, operators and
    product-state starts:
    - plain: one request at a time through the direct entry points with the
      reference's algorithms (Lanczos with full reorthogonalisation for the
      ground state, KrylovEvolutionCache with full reorthogonalisation),
      ComplexF64 -- the one-at-a-time plain GPU solves;
    - single: one request at a time through the server (one worker) with the
      serving algorithms (partial reorthogonalisation, Simon 1984);
    - served: the server with `workers` threads, requests submitted at their
      arrival times, matvecs of jobs sharing a sector and term structure
      batched step by step (continuous batching).
    Latency is arrival -> result; for the one-at-a-time legs it is the FIFO
    queue of the same arrivals served back to back."""
    import edkit_hip as edk
    from edkit_hip._lib import check, lib
    sys.path.insert(0, ROOT)
    from tests.models import blbq, j1j2, xxz
    jobs = c5_stream(n_requests, rate, 5)
    bases = {}

    def basis_of(j):
        key = (j["L"], j["base"], j["k"], j["p"])
        if key not in bases:
            N = j["L"] // 2 if j["base"] == 2 else j["L"]
            if j["p"]:
                bases[key] = edk.TranslationParityBasis(L=j["L"], N=N, k=j["k"], p=j["p"], base=j["base"])
            else:
                bases[key] = edk.TranslationalBasis(L=j["L"], N=N, k=j["k"], base=j["base"])
        return bases[key]

    def operator_of(j):
        B = basis_of(j)
        name, c = j["model"]
        return xxz(B, c) if name == "xxz" else (j1j2(B, c) if name == "j1j2" else blbq(B, c))

    def say(msg):
        if log:
            print(msg, file=log, flush=True)

    t_build = time.perf_counter()
    ops = [operator_of(j) for j in jobs]
    for H in {id(H): H for H in ops}.values():
        _ = H @ edk.DeviceArray((H.B.D,), np.complex128)
    for B in bases.values():
        B.ctx.synchronize()
    t_build = time.perf_counter() - t_build
    starts = [min(int(j["start"] * H.B.D), H.B.D - 1) for j, H in zip(jobs, ops)]
    arrivals = np.array([j["arrival"] for j in jobs])
    out = {"requests": len(jobs), "sectors_built": len(bases), "sector_build_s": t_build}

    def fifo_latency(service):
        done, lat = 0.0, []
        for a, sv in zip(arrivals, service):
            done = max(done, a) + sv
            lat.append(done - a)
        return lat

    results = {}
    if "plain" in legs:
        ctx = edk.default_context()
        service, res = [], []
        for i, (j, H, idx) in enumerate(zip(jobs, ops, starts)):
            x0 = np.zeros(H.B.D, np.complex128)
            x0[idx] = 1.0
            x0 = edk.DeviceArray.from_host(x0, ctx=H.B.ctx)
            H.B.ctx.synchronize()
            t0 = time.perf_counter()
            if j["kind"] == "ground":
                E, r, steps, _ = edk.lanczos_ground(H, x0, m=400, tol=1e-9, refine=False, reorth="full")
                res.append({"E0": E, "resid": r, "steps": steps})
            else:
                c = edk.KrylovEvolutionCache(H, x0, tol=1e-10, m_init=30, m_max=60, reorth="full")
                y = c.evolve(j["t"])
                re_, im_ = C.c_double(), C.c_double()
                check(lib().edk_vec_dot(H.B.ctx.h, 0, x0.ptr, y.ptr, H.B.D, C.byref(re_), C.byref(im_)))
                nrm = C.c_double()
                check(lib().edk_vec_nrm2(H.B.ctx.h, 0, y.ptr, H.B.D, C.byref(nrm)))
                res.append({"overlap": complex(re_.value, im_.value), "norm": nrm.value})
                del c, y
            service.append(time.perf_counter() - t0)
            if (i + 1) % 50 == 0:
                say(f"[c5] plain {i + 1}/{len(jobs)}")
        results["plain"] = res
        lat = fifo_latency(service)
        out["plain"] = {"req_s": len(jobs) / sum(service), "busy_s": sum(service),
                        "latency_s": {"p50": _pctl(lat, 50), "p99": _pctl(lat, 99)}}
        classes = {}
        for j, sv in zip(jobs, service):
            k = f"L{j['L']}{'s1' if j['base'] == 3 else ''}-{j['kind']}"
            c = classes.setdefault(k, [0, 0.0])
            c[0] += 1
            c[1] += sv
        tot = sum(service)
        out["plain"]["time_share"] = {k: round(v[1] / tot, 4) for k, v in
                                      sorted(classes.items(), key=lambda kv: -kv[1][1])[:8]}

    def submit(srv, j, H, idx):
        if j["kind"] == "ground":
            return srv.submit_ground(H, psi0_index=idx, m=400, tol=1e-9, reorth="partial")
        return srv.submit_evolve(H, j["t"], psi0_index=idx, reorth="partial")

    def pack(r, j):
        if j["kind"] == "ground":
            return {"E0": r["E0"], "resid": r["resid"], "steps": r["steps"]}
        return {"overlap": complex(r["overlap_re"], r["overlap_im"]), "norm": r["norm"]}

    if "single" in legs:
        with edk.Server(workers=1, batch_wait_us=wait_us) as srv:
            s0 = srv.clock()
            res, service = [], []
            for i, (j, H, idx) in enumerate(zip(jobs, ops, starts)):
                r, _ = srv.wait(submit(srv, j, H, idx))
                res.append(pack(r, j))
                service.append(r["t_end"] - r["t_start"])
                if (i + 1) % 50 == 0:
                    say(f"[c5] single {i + 1}/{len(jobs)}")
            span = srv.clock() - s0
        results["single"] = res
        lat = fifo_latency(service)
        out["single"] = {"req_s": len(jobs) / span, "latency_s": {"p50": _pctl(lat, 50), "p99": _pctl(lat, 99)}}

    if "served" in legs:
        with edk.Server(workers=workers, batch_wait_us=wait_us) as srv:
            t0 = srv.clock()
            tickets = []
            for j, H, idx in zip(jobs, ops, starts):
                now = srv.clock() - t0
                if j["arrival"] > now:
                    time.sleep(j["arrival"] - now)
                tickets.append(submit(srv, j, H, idx))
            raw = []
            for i, tk in enumerate(tickets):
                r, _ = srv.wait(tk)
                raw.append(r)
                if (i + 1) % 50 == 0:
                    say(f"[c5] served {i + 1}/{len(tickets)}")
            st = srv.stats()
            launches, columns, wcols = srv.batch_stats(weighted=True)
        bad = [r["message"] for r in raw if r["status"] != 0]
        lat = [r["t_end"] - t0 - a for r, a in zip(raw, arrivals)]
        span = max(r["t_end"] for r in raw) - t0
        results["served"] = [pack(r, j) for r, j in zip(raw, jobs)]
        out["served"] = {"req_s": len(jobs) / span, "latency_s": {"p50": _pctl(lat, 50), "p99": _pctl(lat, 99)},
                         "failed": len(bad), "first_error": bad[0] if bad else None,
                         "workers": workers, "peak_concurrent_jobs": st["peak_running"],
                         "peak_inflight_GB": st["peak_inflight_bytes"] / 1e9,
                         "batched_launches": launches,
                         "columns_per_launch": columns / launches if launches else None,
                         "columns_per_launch_dim_weighted": wcols,
                         "max_ground_resid": max((r["resid"] for r, j in zip(raw, jobs) if j["kind"] == "ground"),
                                                 default=None)}
    # accuracy across legs: energies and evolved overlaps against the plain solves
    if "plain" in results:
        for leg in ("single", "served"):
            if leg not in results:
                continue
            dE = [abs(a["E0"] - b["E0"]) for a, b, j in zip(results[leg], results["plain"], jobs)
                  if j["kind"] == "ground"]
            dO = [abs(a["overlap"] - b["overlap"]) for a, b, j in zip(results[leg], results["plain"], jobs)
                  if j["kind"] == "evolve"]
            out[leg]["max_dE_vs_plain"] = max(dE, default=None)
            out[leg]["max_overlap_diff_vs_plain"] = max(dO, default=None)
    if "single" in results and "served" in results:
        out["served"]["identical_to_single"] = all(a == b for a, b in zip(results["single"], results["served"]))
    if "plain" in out and "served" in out:
        out["speedup_served_vs_plain"] = out["served"]["req_s"] / out["plain"]["req_s"]
    if "single" in out and "served" in out:
        out["speedup_served_vs_single"] = out["served"]["req_s"] / out["single"]["req_s"]
    return out


def c5_workload(args, ev):
    """Serving stream (Poisson arrivals, mixed sectors and Hamiltonians,
    ground-state and evolution requests): the served rate with continuous
    batching against one-at-a-time plain solves (c5_run)."""
    legs = ("plain", "single", "served") if not args.no_plain else ("single", "served")
    r = c5_run(args.requests, args.rate, args.workers, args.batch_wait_us, legs, log=sys.stderr)
    out = {
        "metric": "mixed-solve req/s (served) with p50/p99 latency",
        "value": r["served"]["req_s"], "unit": "req/s", "n_gpus": 1, "steps": r["requests"], "warmup": 0,
        "ms_per_step": 1e3 / r["served"]["req_s"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "complex128", "data": "synthetic (seeded Poisson stream, product-state starts)",
        "config": {"workload": f"serving stream: {r['requests']} requests, Poisson {args.rate} req/s, spin-1/2 "
                               "L=18..30 XXZ/J1-J2 and spin-1 BLBQ L=12..16, random k/p sectors, Sz=0; "
                               "60% Lanczos to residual 1e-9, 40% Krylov evolution to t in [1,20]",
                   "workers": args.workers, "parallelism": "single GPU"},
        "latency_s": r["served"]["latency_s"],
    }
    out.update({k: v for k, v in r.items() if k not in ("requests",)})
    print(json.dumps(out))


def c2_workload(args, ev):
    """Symmetry-reduced Krylov evolution: J1-J2 ring L=24 J1=1 J2=0.5,
    TranslationalBasis N=12 k=0 (D = 112 720); a seeded Gaussian psi evolved
    to t = 0.1, 0.2, ..., 10 (100 outputs) through KrylovEvolutionCache with
    Krylov dimension 30, in ComplexF64 and in ComplexF32 vectors with FP64
    Krylov scalars.  The CPU leg runs the oracle's KrylovEvolutionCache on the
    same input: its time is the baseline and its states are the reference for
    the reported errors."""
    import edkit_hip as edk
    sys.path.insert(0, ROOT)
    from tests.models import j1j2
    B = edk.TranslationalBasis(L=24, N=12, k=0)
    H = j1j2(B, 0.5)
    D = B.D
    rng = np.random.default_rng(24)
    psi = rng.standard_normal(D) + 1j * rng.standard_normal(D)
    psi /= np.linalg.norm(psi)
    ts = [0.1 * (i + 1) for i in range(100)]
    kw = dict(m_init=30, m_max=30)
    dev, states = {}, {}
    for name, dt in (("c64", np.complex128), ("c32", np.complex64)):
        outs = [edk.DeviceArray((D,), dt) for _ in ts]
        x0 = edk.DeviceArray.from_host(psi.astype(dt))
        ev.device_sync()
        t0 = time.perf_counter()
        cache = edk.KrylovEvolutionCache(H, x0, dtype=dt, **kw)
        for t, o in zip(ts, outs):
            cache.evolve(t, out=o)
        ev.device_sync()
        secs = time.perf_counter() - t0
        d = cache.diagnostics
        states[name] = np.stack([o.to_host().astype(np.complex128) for o in outs])
        dev[name] = {"s": secs, "outputs_per_s": len(ts) / secs, "matvecs": d["matvecs"],
                     "restarts": d["restarts"], "max_dim": d["max_dim_used"]}
        del cache, outs
    out = {
        "metric": "exp(-iHt)psi outputs/s (100 times, Krylov dim 30)",
        "value": dev["c64"]["outputs_per_s"], "unit": "outputs/s", "n_gpus": 1, "steps": len(ts),
        "warmup": 0, "ms_per_step": 1e3 * dev["c64"]["s"] / len(ts), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "complex128", "data": "synthetic (seeded Gaussian psi)",
        "config": {"workload": "J1-J2 ring L=24 J2=0.5, TranslationalBasis N=12 k=0, exp(-iHt)psi to t=10 "
                               "at dt=0.1, Krylov dim 30", "D": D, "parallelism": "single GPU"},
        "device": dev,
        "c32_vs_c64_max_err": float(np.max(np.linalg.norm(states["c32"] - states["c64"], axis=1))),
    }
    if not args.no_cpu:
        from oracle import oracle as orc
        nthreads = host_threads()
        ob = orc.Basis("translational", 24, N=12, k=0, nthreads=nthreads)
        op = orc.Op(ob, H.flatten())
        t0 = time.perf_counter()
        kc = orc.Krylov(op, psi, nthreads=nthreads, **kw)
        ref = np.stack([kc.evolve(t) for t in ts])
        secs = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": len(ts) / secs, "unit": "outputs/s", "cores": nthreads, "kind": "port",
                               "sample": f"the full workload: 100 outputs of the oracle's KrylovEvolutionCache "
                                         f"(threaded matvec), {secs:.2f} s"}
        out["c64_vs_oracle_max_err"] = float(np.max(np.linalg.norm(states["c64"] - ref, axis=1)))
        out["c32_vs_oracle_max_err"] = float(np.max(np.linalg.norm(states["c32"] - ref, axis=1)))
    print(json.dumps(out))


def c1_workload(args, ev):
    """Plain method, small: XXZ ring L=16 Jxy=1 Delta=1.5, TensorBasis
    (D = 65 536), ComplexF64; a seeded Gaussian psi; 100 repeated H*psi, then a
    200-step Lanczos for the lowest eigenvalue.  The CPU leg runs the same work
    through the oracle (threaded scatter mul, its Lanczos)."""
    import edkit_hip as edk
    from edkit_hip.matvec import matvec_into
    sys.path.insert(0, ROOT)
    from tests.models import xxz
    B = edk.TensorBasis(L=16)
    H = xxz(B, 1.5)
    D = B.D
    rng = np.random.default_rng(16)
    psi = rng.standard_normal(D) + 1j * rng.standard_normal(D)
    psi /= np.linalg.norm(psi)
    x = edk.DeviceArray.from_host(psi)
    y = edk.DeviceArray((D,), np.complex128)
    for _ in range(5):
        matvec_into(y, H, x)
    ev.device_sync()
    t0 = time.perf_counter()
    for _ in range(100):
        matvec_into(y, H, x)
    ev.device_sync()
    t_mv = time.perf_counter() - t0
    t0 = time.perf_counter()
    E, r, steps, _ = edk.lanczos_ground(H, x, m=200, tol=0.0, refine=False)
    t_lz = time.perf_counter() - t0
    out = {
        "metric": "100 H*psi + 200-step Lanczos, seconds", "value": t_mv + t_lz, "unit": "s", "n_gpus": 1,
        "steps": 100, "warmup": 5, "ms_per_step": 1e3 * t_mv / 100, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "complex128", "data": "synthetic (seeded Gaussian psi)",
        "config": {"workload": "XXZ ring L=16 Jxy=1 Delta=1.5, TensorBasis, 100 H*psi + 200-step Lanczos",
                   "D": D, "parallelism": "single GPU"},
        "matvec_s": t_mv, "lanczos_s": t_lz, "E0": E, "lanczos_resid": r,
    }
    if not args.no_cpu:
        from oracle import oracle as orc
        nthreads = host_threads()
        ob = orc.Basis("tensor", 16)
        op = orc.Op(ob, H.flatten())
        tgt = np.zeros(D, np.complex128)
        t0 = time.perf_counter()
        for _ in range(100):
            tgt[:] = 0
            op.mul_(tgt, psi, nthreads)
        c_mv = time.perf_counter() - t0
        t0 = time.perf_counter()
        Eo, ro, so, _ = orc.lanczos_ground(op, psi, 200, 0.0, nthreads)
        c_lz = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": c_mv + c_lz, "unit": "s", "cores": nthreads, "kind": "port",
                               "sample": f"the full workload through the oracle: 100 threaded scatter H*psi "
                                         f"{c_mv:.2f} s + 200-step Lanczos {c_lz:.2f} s"}
        out["E0_vs_oracle"] = abs(E - Eo)
    print(json.dumps(out))


def c6_workload(args, ev):
    """Open-system propagation: dissipative XXZ chain of L=10 spins-1/2
    (d = 1024, a Liouville space of 1 048 576), Jxy=1, Delta=0.5, hz=0.3, one
    decay jump sqrt(0.1) s-_i per site (10 dense jumps); a seeded random pure
    state evolved to t = 0.5, 1.0, ..., 5.0 (10 outputs) by
    LindbladArnoldiCache with the reference defaults (tol 1e-10, m_init 20,
    m_max 50).  Roofline of the Liouvillian apply: 8 d^3 (2 + 2 n_jumps) FP64
    flops (complex multiply-adds of its GEMMs) against the FP64 matrix-core
    peak, over the apply's mean duration (HIP events on the map's stream).
    The CPU leg times the oracle's apply (numpy/BLAS, threaded) and charges it
    per Liouvillian apply the device run made.  The "sparse" object runs the
    same model through the sparse form (the jumps and H as CSR), HBM-bound."""
    import edkit_hip as edk
    from edkit_hip.lindblad import _colmajor
    sys.path.insert(0, ROOT)
    from tests.lindblad_models import dissipative_xxz
    Ls, gamma = 10, 0.1
    d = 2 ** Ls
    H, J = dissipative_xxz(Ls, 1.0, 0.5, gamma)
    rng = np.random.default_rng(610)
    psi = rng.standard_normal(d) + 1j * rng.standard_normal(d)
    psi /= np.linalg.norm(psi)
    r0 = np.outer(psi, psi.conj())
    A = edk.LiouvillianMap(H, J)
    stream = A.ctx.stream
    x = edk.DeviceArray.from_host(_colmajor(r0), ctx=A.ctx)
    y = edk.DeviceArray((d, d), np.complex128, ctx=A.ctx, zero=False)
    for _ in range(3):
        A.apply_device(x, y)
    n_ap = 20
    c0, c1 = ev.event(), ev.event()
    ev.record(c0, stream)
    for _ in range(n_ap):
        A.apply_device(x, y)
    ev.record(c1, stream)
    ap_ms = ev.elapsed_ms(c0, c1) / n_ap
    flops = 8.0 * d ** 3 * (2 + 2 * len(J))
    tflops = flops / (ap_ms * 1e-3) / 1e12
    ts = [0.5 * (i + 1) for i in range(10)]
    # warm-up: a cache of the same sizes served at t = 0 (rocBLAS/rocSOLVER
    # set-up and the allocator's blocks), no restarts
    warm = edk.LindbladArnoldiCache(A, r0, record_min_eig=True)
    warm.evolve_device(0.0)
    del warm
    ev.device_sync()
    t0 = time.perf_counter()
    cache = edk.LindbladArnoldiCache(A, r0)
    outs = [cache.evolve_device(t) for t in ts]
    ev.device_sync()
    secs = time.perf_counter() - t0
    dg = cache.diagnostics
    last = outs[-1].to_host().T
    out = {
        "metric": "Lindblad outputs/s (dissipative XXZ L=10, d=1024, adaptive Arnoldi)",
        "value": len(ts) / secs, "unit": "outputs/s", "n_gpus": 1, "steps": len(ts), "warmup": 1,
        "ms_per_step": 1e3 * secs / len(ts), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "complex128", "data": "synthetic (seeded random pure state)",
        "config": {"workload": "dissipative XXZ chain L=10 (d=1024) Delta=0.5 hz=0.3, 10 decay jumps "
                               "sqrt(0.1) s-_i, rho(t) at t=0.5..5 by LindbladArnoldiCache(tol=1e-10, "
                               "m_init=20, m_max=50)", "d": d, "n_jumps": len(J), "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "achieved": tflops, "peak": 78.6, "unit": "TFLOP/s",
                     "frac": tflops / 78.6, "traffic": None, "flops_per_apply": flops, "apply_ms": ap_ms},
        "diagnostics": {k: dg[k] for k in ("basis_builds", "basis_extensions", "restarts", "liouvillian_applies",
                                           "max_dim_used", "max_trace_drift", "max_hermiticity_drift")},
        "trace_last": float(np.real(np.trace(last))),
        "excited_last": float(np.real(np.trace(last @ sum(np.kron(np.kron(np.eye(2 ** i), np.diag([1.0, 0.0])),
                                                                   np.eye(2 ** (Ls - i - 1))) for i in range(Ls))))),
    }
    # the same model through the sparse form (CSR row/column passes over rho)
    import scipy.sparse as sp
    As = edk.LiouvillianMap(sp.csr_matrix(H), [sp.csr_matrix(Lk) for Lk in J])
    nnz = sum(int(m.nnz) for m in [sp.csr_matrix(Lk) for Lk in J])
    ys = edk.DeviceArray((d, d), np.complex128, ctx=As.ctx, zero=False)
    for _ in range(3):
        As.apply_device(x, ys)
    ev.record(c0, stream)
    for _ in range(n_ap):
        As.apply_device(x, ys)
    ev.record(c1, stream)
    sp_ms = ev.elapsed_ms(c0, c1) / n_ap
    # algorithmic bytes: every pass reads its dense operand once and writes
    # (left: K1 X, L_k X) or updates (right: X K2, T L_k^H) the d x d output,
    # plus the CSR entries (4 + 16 bytes) and row pointers it reads
    Dm = sum((sp.csr_matrix(Lk).conj().T @ sp.csr_matrix(Lk) for Lk in J), sp.csr_matrix((d, d), dtype=complex))
    k1nnz = int((sp.csr_matrix(H) * -1j - 0.5 * Dm).nnz)
    d2 = 16.0 * d * d
    sp_bytes = (2 * d2 + 3 * d2) * (1 + len(J)) + 20.0 * (2 * k1nnz + 2 * nnz) + 8.0 * (d + 1) * (2 + 2 * len(J))
    ev.device_sync()
    t0 = time.perf_counter()
    cs = edk.LindbladArnoldiCache(As, r0)
    outs_s = [cs.evolve_device(t) for t in ts]
    ev.device_sync()
    secs_s = time.perf_counter() - t0
    dgs = cs.diagnostics
    out["sparse"] = {
        "apply_ms": sp_ms, "bytes_per_apply": sp_bytes, "GBs": sp_bytes / (sp_ms * 1e-3) / 1e9,
        "frac_hbm": sp_bytes / (sp_ms * 1e-3) / 1e9 / 8000.0, "outputs_per_s": len(ts) / secs_s,
        "liouvillian_applies": dgs["liouvillian_applies"], "restarts": dgs["restarts"],
        "max_diff_vs_dense": float(max(np.abs(a.to_host() - b.to_host()).max() for a, b in zip(outs, outs_s))),
    }
    if not args.no_cpu:
        from oracle import lindblad as ol
        ol.liouvillian_apply(H, J, r0)
        t0 = time.perf_counter()
        reps = 2
        for _ in range(reps):
            ol.liouvillian_apply(H, J, r0)
        cpu_ap = (time.perf_counter() - t0) / reps
        out["cpu_baseline"] = {
            "value": len(ts) / (dg["liouvillian_applies"] * cpu_ap), "unit": "outputs/s",
            "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), "kind": "port",
            "sample": f"{reps} applies of the oracle's Liouvillian (numpy/BLAS zgemm) at d=1024: "
                      f"{cpu_ap:.2f} s each, charged for the device run's {dg['liouvillian_applies']} applies "
                      f"(orthogonalisation and the reduced model not charged)"}
    print(json.dumps(out))


def c7_workload(args, ev):
    """Custom generators: nearest-neighbour Heisenberg model on a 6 x 5 torus
    (L = 30), momentum (0, 0) with T_x and T_y given as site permutations
    (basis(...; symmetries); T_x is applied in the shift-and-mask form, T_y as
    a rotation of the packed word), N = 15.
    One step = one ComplexF64 H*psi through the Abelian two-site kernel
    (bit-parallel bond families, pruned group walk, rank-directory lookups);
    roofline bytes as for the default line (D (8 + 8 + 32) + NNZ_off (16 + 8)).
    The CPU leg times the oracle's threaded scatter on a column sample."""
    import edkit_hip as edk
    from edkit_hip.matvec import device_operator
    sys.path.insert(0, ROOT)
    from tests.models import heisenberg_2d, torus_translations
    Lx, Ly, N = 6, 5, 15
    Lt = Lx * Ly
    tx, ty = torus_translations(Lx, Ly)
    sym = [(tx, 0), (ty, 0)]
    t0 = time.perf_counter()
    B = edk.basis(L=Lt, N=N, symmetries=sym)
    ev.device_sync()
    t_build = time.perf_counter() - t0
    H = heisenberg_2d(B, Lx, Ly)
    D = B.D
    nnz_off = device_operator(H).count_offdiag()
    rng = np.random.default_rng(75)
    xh = rng.standard_normal(D) + 1j * rng.standard_normal(D)
    xh /= np.linalg.norm(xh)
    x = edk.DeviceArray.from_host(xh)
    y = edk.DeviceArray((D,), np.complex128)
    steps = min(args.steps, 100)
    kern_ms, step_ms = time_matvec(edk, H, x, y, steps, args.warmup, ev, B.ctx.stream)
    b_alg = D * 48 + nnz_off * 24
    achieved = b_alg / (kern_ms * 1e-3) / 1e9
    out = {
        "metric": "2D torus 6x5 (custom generators) H*psi/s", "value": 1e3 / step_ms, "unit": "H*psi/s",
        "n_gpus": 1, "steps": steps, "warmup": args.warmup, "ms_per_step": step_ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "complex128", "data": "synthetic (seeded Gaussian psi)",
        "config": {"workload": "Heisenberg model on a 6x5 torus, basis(L=30, N=15, symmetries=[(T_x, 0), "
                               "(T_y, 0)]), ComplexF64 H*psi", "D": D, "nnz_offdiag": nnz_off,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": profiled_traffic(os.path.join(ROOT, "profiles", "matvec_perm_torus6x5_custom.md"),
                                                 PERM_SOURCES),
                     "bytes_alg_per_launch": b_alg,
                     "kernel_ms": kern_ms},
        "basis_build_s": t_build,
    }
    if not args.no_cpu:
        from oracle import oracle as orc
        nthreads = host_threads()
        t0 = time.perf_counter()
        ob = orc.Basis("abelian", Lt, N=N, symmetries=sym, nthreads=nthreads)
        t_ob = time.perf_counter() - t0
        op = orc.Op(ob, H.flatten())
        tgt = np.zeros(ob.D, np.complex128)
        n = min(ob.D, 200000)
        cols = np.linspace(0, ob.D - 1, n).astype(np.int64)
        t0 = time.perf_counter()
        op.mul_cols_(tgt, xh, cols, nthreads)
        t_s = time.perf_counter() - t0
        per = t_s / n * ob.D
        out["cpu_baseline"] = {"value": 1.0 / per, "unit": "H*psi/s", "cores": nthreads, "kind": "port",
                               "sample": f"{n} of {ob.D} evenly spaced columns of the same H*psi (oracle scatter, "
                                         f"threaded) took {t_s:.2f} s -> {per:.1f} s per full H*psi; oracle table "
                                         f"build {t_ob:.1f} s"}
        out["oracle_D_equal"] = bool(ob.D == D)
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--c32", action="store_true", help="also time the ComplexF32 form")
    ap.add_argument("--copy-probe", action="store_true",
                    help="(kept for profiling commands; the copy probe always runs)")
    ap.add_argument("--workload", default="c3", choices=["c1", "c2", "c3", "c4", "c5", "c6", "c7"],
                    help="c3: L=32 H*psi roofline line (default); c1: L=16 plain method; c2: Krylov "
                         "evolution L=24; c4: batched uneven sectors; c5: serving stream; c6: Lindblad "
                         "propagation d=1024; c7: 2D torus with custom generators")
    ap.add_argument("--requests", type=int, default=2000, help="c5: requests in the stream")
    ap.add_argument("--rate", type=float, default=200.0, help="c5: mean arrival rate, req/s")
    ap.add_argument("--workers", type=int, default=32, help="c5: server worker streams")
    ap.add_argument("--batch-wait-us", type=float, default=250.0,
                    help="c5: longest wait for a group's other members before a partial block launches")
    ap.add_argument("--no-plain", action="store_true", help="c5: skip the plain one-at-a-time leg")
    ap.add_argument("--serve-requests", type=int, default=200,
                    help="c3 line: requests of its serving leg (0 skips it)")
    ap.add_argument("--lanczos-steps", type=int, default=300,
                    help="steps of the ground-state Lanczos leg (0 skips it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    import edkit_hip as edk
    from edkit_hip.matvec import device_operator
    ev = HipEvents()
    if args.workload in ("c1", "c2", "c4", "c5", "c6", "c7"):
        {"c1": c1_workload, "c2": c2_workload, "c4": c4_workload, "c5": c5_workload,
         "c6": c6_workload, "c7": c7_workload}[args.workload](args, ev)
        return
    t0 = time.perf_counter()
    B = edk.TranslationParityBasis(L=L, N=NUP, k=0, p=1)
    ev.device_sync()
    t_enum = time.perf_counter() - t0
    D = B.D
    H = edk.trans_inv_operator(edk.spin((1.0, "xx"), (1.0, "yy"), (DELTA, "zz")), [1, 2], B)
    nnz_off = device_operator(H).count_offdiag()
    rng = np.random.default_rng(20261203)
    xh = rng.standard_normal(D) + 1j * rng.standard_normal(D)
    xh /= np.linalg.norm(xh)
    x = edk.DeviceArray.from_host(xh)
    y = edk.DeviceArray((D,), np.complex128)
    stream = B.ctx.stream
    kern_ms, step_ms = time_matvec(edk, H, x, y, args.steps, args.warmup, ev, stream)
    s_I, s_R, s_v = 8, 8, 16
    b_alg = D * (s_I + s_R + 2 * s_v) + nnz_off * (s_v + s_I)
    achieved = b_alg / (kern_ms * 1e-3) / 1e9
    out = {
        "metric": "L=32 k=0 H*psi/s (% HBM roofline)",
        "value": 1e3 / step_ms,
        "unit": "H*psi/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex128",
        "data": "synthetic (seeded Gaussian psi)",
        "config": {"workload": "XXZ ring L=32 Delta=1, TranslationParityBasis N=16 k=0 p=+1, "
                               "ComplexF64 H*psi", "L": L, "D": D, "nnz_offdiag": nnz_off,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": profiled_traffic(),
                     "bytes_alg_per_launch": b_alg, "kernel_ms": kern_ms},
        "basis_build_s": t_enum,
    }
    # a plain streaming copy of known size (y = 1.0 * x over 2^26 ComplexF64,
    # 1 GiB read + 1 GiB written, four times the 256 MB last-level cache):
    # the achievable HBM bandwidth on this box, and the calibration kernel of
    # the FETCH_SIZE/WRITE_SIZE counters
    from edkit_hip._lib import check, lib
    n = (1 << 30) // 16
    src = edk.DeviceArray((n,), np.complex128)
    dst = edk.DeviceArray((n,), np.complex128)
    check(lib().edk_vec_scale(B.ctx.h, 0, dst.ptr, src.ptr, 1.0, n))
    ev.device_sync()
    c0, c1 = ev.event(), ev.event()
    ev.record(c0, stream)
    for _ in range(3):
        check(lib().edk_vec_scale(B.ctx.h, 0, dst.ptr, src.ptr, 1.0, n))
    ev.record(c1, stream)
    copy_ms = ev.elapsed_ms(c0, c1) / 3
    out["copy_probe"] = {"bytes": 32 * n, "ms": copy_ms, "GBs": 32 * n / (copy_ms * 1e-3) / 1e9}
    out["roofline"]["frac_of_copy"] = achieved / out["copy_probe"]["GBs"]
    del src, dst
    if args.lanczos_steps > 0:
        out["lanczos"] = lanczos_leg(edk, H, xh, args.lanczos_steps, D, b_alg)
    if args.c32:
        x32 = edk.DeviceArray.from_host(xh.astype(np.complex64))
        y32 = edk.DeviceArray((D,), np.complex64)
        k32, s32 = time_matvec(edk, H, x32, y32, args.steps, args.warmup, ev, stream)
        b32 = D * (s_I + s_R + 2 * 8) + nnz_off * (8 + s_I)
        out["c32"] = {"value": 1e3 / s32, "kernel_ms": k32, "achieved_GBs": b32 / (k32 * 1e-3) / 1e9}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(H, xh)
    if args.serve_requests > 0:
        # the metric's second half: a reduced serving stream (c5_run)
        del x, y
        r = c5_run(args.serve_requests, args.rate, args.workers, args.batch_wait_us, log=sys.stderr)
        out["serving"] = {
            "metric": "mixed-solve req/s (served, continuous batching) with p50/p99 latency",
            "requests": r["requests"], "req_s": r["served"]["req_s"], "latency_s": r["served"]["latency_s"],
            "plain_one_at_a_time_req_s": r["plain"]["req_s"], "plain_latency_s": r["plain"]["latency_s"],
            "single_served_req_s": r["single"]["req_s"],
            "speedup_vs_plain": r["speedup_served_vs_plain"], "speedup_vs_single": r["speedup_served_vs_single"],
            "columns_per_launch": r["served"]["columns_per_launch"], "failed": r["served"]["failed"],
            "max_ground_resid": r["served"]["max_ground_resid"],
            "max_dE_vs_plain": r["served"]["max_dE_vs_plain"],
            "max_overlap_diff_vs_plain": r["served"]["max_overlap_diff_vs_plain"],
            "identical_to_single": r["served"]["identical_to_single"],
        }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
