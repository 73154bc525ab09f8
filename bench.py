"""Headline benchmark: exact n=1024 products of signed 2048-bit entries on one
MI355X through the learned dispatch (MGX_ROUTE_AUTO with the committed cost
model, which picks the INT8-MFMA multi-modular route), timed beside the
direct CRT call and the plain GPU mpn route (Winograd over per-entry
multiplications) on the same inputs, bit for bit compared between the routes
and against the oracle on sampled entries, with mat-cpu's method 35 (GMP
mpz + Winograd job DAG + its thread scheduler) timed on the host cores as
the CPU baseline.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 1024] [--bits 2048]
                  [--no-plain] [--no-cpu]
  python bench.py --batch [--products 512] [--chunk 32] [--steps K]

Steps cycle over three input seeds (A, B from seeds 2s-1, 2s for s = 1, 2, 3);
`value` is the mean wall time of the K timed steps (whole job: feature scan,
dispatch, cached plan, stage kernels, error check; inputs resident in HBM),
and the per-seed medians are reported beside it.  The default 20 untimed
warm-up products (about 50 ms) cover the device's clock ramp: after the idle
set-up the first ten or so products run up to 30% slower per kernel.

--batch times the mixed workload instead (one more JSON line): 512
independent exact products, n uniform in [16, 512], entry bits log-uniform in
[64, 16384], 10% rectangular (K x 2K x K), routed by the committed cost model
and run as hipGraph replays of the persistent-kernel work queues.
"""
import argparse
import ctypes
import json
import os
import platform
import random
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bachelor-thesis_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0     # MI355X HBM3E, published
INT8_PEAK_TOPS = 5000.0   # MI355X dense INT8 MFMA, published
MODEL = os.path.join(ROOT, "bachelor-thesis_amd", "costmodel_params.txt")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pmc_traffic(kernel, n, bits):
    """HBM bytes per dispatch of `kernel` at this workload from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (calibrated against a plain copy);
    None when no summary covers it."""
    import glob
    import re
    pat = re.compile(r"^traffic_bytes %s n=%d bits=%d: (\d+)$" % (re.escape(kernel), n, bits))
    for f in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(f) as fh:
            for line in fh:
                m = pat.match(line.strip())
                if m:
                    return int(m.group(1)), os.path.relpath(f, ROOT)
    return None, None


def same_entries(X, Y, cap):
    sx, sy = X[..., cap].view(np.int64), Y[..., cap].view(np.int64)
    if not (sx == sy).all():
        return False
    mask = np.arange(cap)[None, None, :] < np.abs(sx)[..., None]
    return bool((np.where(mask, X[..., :cap], 0) == np.where(mask, Y[..., :cap], 0)).all())


def host_info():
    model = platform.processor() or "?"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, os.cpu_count() or 1, avail


def cpu_baseline(mgx, A, B, n, bits, cap, C_gpu, gpu_block):
    """mat-cpu's method 35 (-B 32, all usable threads) restated in the oracle:
    config 1 (n=64, 1024-bit) with -J 128 and -J 32, then this workload
    (a bounded sample of it for n > 256).  gpu_block(h) is the GPU's product
    of the top-left h x h blocks of A and B, compared with the sample."""
    import oracle
    orc = oracle.lib()
    ver = orc.or_matcpu_gmp_version()
    model, nproc, avail = host_info()
    # the box gives one GPU's share of its cores (16) to a run; use what the
    # process may run on, up to that share
    threads = max(1, min(avail, 16))

    def run(N, X, Y, c, jobsize):
        X, Y = np.ascontiguousarray(X), np.ascontiguousarray(Y)
        out = np.zeros((N, N, c + 1), dtype=np.uint64)
        t = ctypes.c_double()
        st = orc.or_matcpu_winograd_timed(N, out.ctypes.data, X.ctypes.data, Y.ctypes.data, c, 32, jobsize, threads,
                                          ctypes.byref(t))
        if st:
            raise RuntimeError("mat-cpu port failed with status %d" % st)
        return t.value, out

    # config 1
    c1 = (2 * 1024 + 8 + 63) // 64 + 1
    A1, B1 = mgx.random_matrix(64, 64, c1, 1024, 1), mgx.random_matrix(64, 64, c1, 1024, 2)
    ref1 = oracle.mat_mul(A1, B1, c1)
    cfg1 = {}
    for J in (128, 32):
        secs = []
        for _ in range(5):
            t, out = run(64, A1, B1, c1, J)
            secs.append(t)
            assert same_entries(out, ref1, c1), "mat-cpu port differs from the naive oracle at n=64"
        cfg1["J%d_s" % J] = statistics.median(secs)
    # this workload: whole when small; otherwise the top-left n/2 block
    # product with -J 64, whose job DAG has the same shape (343 leaf jobs for
    # n=1024) as the whole product's with -J 128, times 7 (method 35's top
    # level multiplies 7 such blocks; its 15 block additions are left out)
    if n <= 256:
        t, out = run(n, A, B, cap, 128)
        ok = same_entries(out, C_gpu, cap)
        value, sample = t, "the whole n=%d product (%.1f s, -J 128)" % (n, t)
    else:
        h = n // 2
        t, out = run(h, A[:h, :h], B[:h, :h], cap, 64)
        ok = same_entries(out, gpu_block(h), cap)
        value = 7.0 * t
        sample = ("the top-left n=%d block product A11*B11 with -J 64 (%.2f s), the same job-DAG shape as the whole "
                  "product with -J 128, times 7 (the products of method 35's top Winograd level; its block additions "
                  "left out)" % (h, t))
    log("cpu baseline: %.2f s (%s); config 1: -J 128 %.3f s, -J 32 %.3f s" % (value, sample, cfg1["J128_s"],
                                                                           cfg1["J32_s"]))
    return {"value": value, "unit": "s", "cores": threads, "kind": "port",
            "sample": "mat-cpu method 35 (-B 32 -J 128, %d threads) restated over GMP %s: %s; entries equal the GPU "
                      "result: %s" % (threads, ver.decode() if ver else "?", sample, ok),
            "host": {"cpu_model": model, "nproc": nproc, "usable_threads": avail, "gmp": ver.decode() if ver else None},
            "config1_n64_b1024": cfg1}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=15)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--bits", type=int, default=2048)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch", action="store_true")
    ap.add_argument("--products", type=int, default=512)
    ap.add_argument("--chunk", type=int, default=32)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path, so --gpus must be 1",
              file=sys.stderr)
        sys.exit(2)
    if args.batch:
        return run_batch(args)

    import mgx
    mgx.init(0)
    L = mgx.lib()
    n, bits = args.n, args.bits
    cap = (2 * bits + 12 + 63) // 64 + 1   # 66 limbs for 2048-bit inputs at n=1024
    W = cap + 1
    seeds = [1, 2, 3]
    log("generating %dx%d %d-bit inputs (seeds %s)" % (n, n, bits, seeds))
    host = {s: (mgx.random_matrix(n, n, cap, bits, 2 * s - 1), mgx.random_matrix(n, n, cap, bits, 2 * s))
            for s in seeds}
    dev = {s: (mgx.DeviceBuffer.from_array(a), mgx.DeviceBuffer.from_array(b)) for s, (a, b) in host.items()}
    dC = {s: mgx.DeviceBuffer(n * n * W * 8) for s in seeds}
    stream = mgx.c_vp()
    mgx.call("mgx_stream_create", ctypes.byref(stream))
    st_ = stream.value
    model = mgx.c_vp()
    mgx.check(L.mgx_costmodel_load(ctypes.byref(model), MODEL.encode()))
    mgx.check(L.mgx_dispatch_set_model(model))
    route = ctypes.c_int()
    dA1, dB1 = dev[1]
    mgx.check(L.mgx_dispatch_choose(n, n, n, dA1.ptr, n, dB1.ptr, n, cap, ctypes.byref(route)))
    log("learned dispatch picks route %d" % route.value)

    def timed(fn, K, W_):
        for i in range(W_):
            fn(seeds[i % 3])
        mgx.call("mgx_device_sync")
        per = {s: [] for s in seeds}
        t_all = time.perf_counter()
        for i in range(K):
            s = seeds[i % 3]
            t0 = time.perf_counter()
            fn(s)   # returns after its own stream synchronization (error check)
            per[s].append(time.perf_counter() - t0)
        mgx.call("mgx_device_sync")
        return (time.perf_counter() - t_all) / K, per

    def auto_step(s):
        a, b = dev[s]
        mgx.check(L.mgx_matmul_route(-1, n, n, n, dC[s].ptr, n, a.ptr, n, b.ptr, n, cap, st_))

    def direct_step(s):
        a, b = dev[s]
        mgx.call("mgx_crt_mfma_matmul", None, n, n, n, dC[s].ptr, n, a.ptr, n, b.ptr, n, cap, st_)

    crt_s, per_auto = timed(auto_step, args.steps, args.warmup)
    direct_s, per_direct = timed(direct_step, args.steps, 1)
    med_auto = {str(s): statistics.median(v) for s, v in per_auto.items() if v}
    med_direct = {str(s): statistics.median(v) for s, v in per_direct.items() if v}
    log("learned dispatch: %.3f ms per product (medians %s); direct CRT call %.3f ms"
        % (crt_s * 1e3, {k: round(v * 1e3, 3) for k, v in med_auto.items()}, direct_s * 1e3))
    for s in seeds:   # fewer than three timed steps leave a seed's product uncomputed
        if not per_auto[s]:
            auto_step(s)
    C = {s: dC[s].to_array((n, n, W)) for s in seeds}

    # ---- stage kernels timed with HIP events on the stream they run on
    plan = mgx.crt_plan(n, bits, bits)
    P = L.mgx_crt_plan_primes(plan)
    S = L.mgx_crt_plan_slices(plan)
    rp, kp = mgx.c_i64(), mgx.c_i64()
    mgx.call("mgx_crt_residue_dims", plan, n, n, ctypes.byref(rp), ctypes.byref(kp))
    rpad, kpad = rp.value, kp.value
    RA = mgx.DeviceBuffer(S * P * rpad * kpad)
    RB = mgx.DeviceBuffer(S * P * rpad * kpad)
    ystride = L.mgx_crt_y_stride(n, n)
    Y = mgx.DeviceBuffer((S * P + 3) // 4 * ystride * 4)
    ev = [mgx.c_vp() for _ in range(5)]
    for e in ev:
        mgx.call("mgx_event_create", ctypes.byref(e))
    reps = max(args.steps, 3)
    stage_ms = [[], [], [], []]
    dTmp = mgx.DeviceBuffer(n * n * W * 8)
    for it in range(reps + 1):
        a, b = dev[seeds[it % 3]]
        mgx.call("mgx_event_record", ev[0].value, st_)
        mgx.call("mgx_crt_reduce", plan, 0, n, n, a.ptr, n, cap, 0, RA.ptr, st_)
        mgx.call("mgx_event_record", ev[1].value, st_)
        mgx.call("mgx_crt_reduce", plan, 1, n, n, b.ptr, n, cap, 0, RB.ptr, st_)
        mgx.call("mgx_event_record", ev[2].value, st_)
        mgx.call("mgx_crt_residue_gemm", plan, n, n, n, RA.ptr, RB.ptr, Y.ptr, st_)
        mgx.call("mgx_event_record", ev[3].value, st_)
        mgx.call("mgx_crt_reconstruct", plan, n, n, Y.ptr, dTmp.ptr, n, cap, st_)
        mgx.call("mgx_event_record", ev[4].value, st_)
        mgx.call("mgx_check_error", st_)
        if it == 0:
            continue   # warm-up
        for k in range(4):
            ms = ctypes.c_float()
            mgx.call("mgx_event_elapsed", ctypes.byref(ms), ev[k].value, ev[k + 1].value)
            stage_ms[k].append(ms.value)
    sm = [statistics.median(v) for v in stage_ms]
    log("stage medians ms: reduceA %.3f reduceB %.3f gemm %.3f reconstruct %.3f (P=%d, S=%d)" % tuple(sm + [P, S]))
    del dTmp
    out_limbs = (2 * bits + (n - 1).bit_length() + 63) // 64
    e = n * n
    red_bytes = 2 * e * bits // 8 + 2 * e * S * P          # entry bytes in, residue digits out
    gemm_ops = 2.0 * n ** 3 * S * S * P
    rec_bytes = e * S * P + 8 * e * (out_limbs + 1)         # digit planes in, sign-magnitude entries out
    gemm_tops = gemm_ops / (sm[2] * 1e-3) / 1e12
    gemm_kernel = "crt_gemm2_kernel" if S == 2 else "crt_gemm_kernel"
    traffic, traffic_src = pmc_traffic(gemm_kernel, n, bits)

    peaks = (ctypes.c_double(), ctypes.c_double(), ctypes.c_double())
    mgx.call("mgx_measure_peaks", *[ctypes.byref(p) for p in peaks])
    mfma_meas, mad_meas, copy_meas = (p.value for p in peaks)
    log("measured: INT8 MFMA %.0f TOPS, v_mad_u64_u32 %.2f T/s, copy %.0f GB/s" % (mfma_meas / 1e12, mad_meas / 1e12,
                                                                                  copy_meas / 1e9))

    # ---- plain route (Winograd over per-entry mpn multiplication), same inputs (seed 1)
    plain = None
    bit_exact_routes = None
    if not args.no_plain:
        dP = mgx.DeviceBuffer(n * n * W * 8)
        a, b = dev[1]

        def plain_step():
            dP.zero(st_)
            mgx.call("mgx_winograd", n, dP.ptr, n, a.ptr, n, b.ptr, n, cap, st_)
        plain_step()   # warm-up
        t0 = time.perf_counter()
        plain_step()
        mgx.call("mgx_device_sync")
        plain_s = time.perf_counter() - t0
        leaf = L.mgx_get_param(0)
        levels = 0
        while (n >> levels) > leaf and ((n >> levels) & 1) == 0:
            levels += 1
        mads = n ** 3 * (7 / 8) ** levels * ((bits + 31) // 32) ** 2   # 32-bit limb products
        plain = {"s": plain_s, "winograd_levels": levels,
                 "valu": {"bound": "valu", "kernel": "plain_mul_kernel", "unit": "mad/s",
                          "achieved": mads / plain_s, "peak": mad_meas, "frac": mads / plain_s / mad_meas,
                          "note": "32x32->64-bit limb products n^3 (7/8)^L (b/32)^2 over the whole route's wall "
                                  "time; peak = measured v_mad_u64_u32 rate on this device"}}
        log("plain route: %.3f s per product" % plain_s)
        bit_exact_routes = same_entries(C[1], dP.to_array((n, n, W)), cap)
        del dP

    # ---- oracle on sampled entries (exact dot products), every seed
    import oracle
    rng = random.Random(1234)
    sample_ok = True
    for s in seeds:
        A_, B_ = host[s]
        rows, cols = rng.sample(range(n), 2), rng.sample(range(n), 2)
        sub = oracle.mat_mul(np.ascontiguousarray(A_[rows]), np.ascontiguousarray(B_[:, cols]), cap)
        sample_ok &= all(mgx.entry_to_int(sub[i, j]) == mgx.entry_to_int(C[s][r, c])
                         for i, r in enumerate(rows) for j, c in enumerate(cols))

    cpu = None
    if not args.no_cpu:
        def gpu_block(h):   # A11 * B11 of seed 1 through strided operands
            a, b = dev[1]
            dCb = mgx.DeviceBuffer(h * h * W * 8)
            mgx.check(L.mgx_matmul_route(-1, h, h, h, dCb.ptr, h, a.ptr, n, b.ptr, n, cap, st_))
            mgx.call("mgx_device_sync")
            return dCb.to_array((h, h, W))
        cpu = cpu_baseline(mgx, host[1][0], host[1][1], n, bits, cap, C[1], gpu_block)

    stage = {
        "reduce": {"ms": sm[0] + sm[1], "bound": "hbm", "bytes": red_bytes, "unit": "GB/s",
                   "achieved": red_bytes / ((sm[0] + sm[1]) * 1e-3) / 1e9, "peak": HBM_PEAK_GBS},
        "residue_gemm": {"ms": sm[2], "bound": "mfma", "ops": gemm_ops, "unit": "TOPS", "achieved": gemm_tops,
                         "peak": INT8_PEAK_TOPS},
        "reconstruct": {"ms": sm[3], "bound": "hbm", "bytes": rec_bytes, "unit": "GB/s",
                        "achieved": rec_bytes / (sm[3] * 1e-3) / 1e9, "peak": HBM_PEAK_GBS},
    }
    for v in stage.values():
        v["frac"] = v["achieved"] / v["peak"]
    line = {
        "metric": "n=1024 2048-bit matmul wall-s, bit-exact; CRT-MFMA vs plain GPU speedup",
        "value": crt_s,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": crt_s * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8 MFMA / int64",
        "data": "synthetic: seeded xoshiro256** entries, uniform %d-bit magnitudes, random signs, seeds 1-3" % bits,
        "config": {"workload": "C=A*B, n=%d square, signed %d-bit entries, learned dispatch (route %d = %s; P=%d primes "
                               "< %s, %d int8 slices)" % (n, bits, route.value,
                                                         "CRT-MFMA" if route.value == 0 else "plain", P,
                                                         "65280" if S == 2 else "2^31", S),
                   "n": n, "bits": bits, "primes": P, "slices": S, "cap_limbs": cap, "route": route.value},
        "median_s_per_seed": med_auto,
        "direct_crt_s": direct_s,
        "direct_crt_median_s_per_seed": med_direct,
        "entry_products_per_s": n ** 3 / crt_s,
        "plain_route_s": plain["s"] if plain else None,
        "speedup_vs_plain": (plain["s"] / crt_s) if plain else None,
        "bit_exact": {"crt_vs_plain_all_entries_seed1": bit_exact_routes, "crt_vs_oracle_sample_all_seeds": sample_ok},
        "stage": stage,
        "plain_roofline": plain["valu"] if plain else None,
        "measured_peaks": {"int8_tops": mfma_meas / 1e12, "v_mad_u64_u32_per_s": mad_meas,
                           "copy_gbs": copy_meas / 1e9},
        "roofline": {"bound": "mfma", "kernel": gemm_kernel, "achieved": gemm_tops, "peak": INT8_PEAK_TOPS,
                     "unit": "TOPS", "frac": gemm_tops / INT8_PEAK_TOPS, "traffic": traffic,
                     "note": "int8 ops 2*n^3*S^2*P per launch / median event-timed launch duration on its stream; "
                             "traffic = HBM bytes per launch from %s" % (traffic_src or "no PMC summary for this "
                                                                                      "workload")},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    mgx.call("mgx_crt_plan_destroy", plan)
    L.mgx_dispatch_set_model(None)
    L.mgx_costmodel_destroy(model)


def run_batch(args):
    import mgx
    mgx.init(0)
    L = mgx.bind_sched()
    rng = np.random.default_rng(20240)
    specs = []
    for i in range(args.products):
        n = int(rng.integers(16, 513))
        bits = int(round(2 ** rng.uniform(6, 14)))
        rect = rng.random() < 0.1
        K, M, N = n, (2 * n if rect else n), n
        cap = (2 * bits + int(np.ceil(np.log2(M))) + 2 + 63) // 64 + 1
        specs.append((K, M, N, bits, cap))
    model = mgx.c_vp()
    mgx.check(L.mgx_costmodel_load(ctypes.byref(model), MODEL.encode()))
    mgx.check(L.mgx_dispatch_set_model(model))
    stream = mgx.c_vp()
    mgx.call("mgx_stream_create", ctypes.byref(stream))
    s = stream.value
    total_s, routes, checked, exact = 0.0, [0, 0, 0, 0], 0, True
    entry_products = sum(K * M * N for K, M, N, _, _ in specs)
    scratch_max = 0
    for c0 in range(0, len(specs), args.chunk):
        chunk = specs[c0:c0 + args.chunk]
        bufs = []
        prods = (mgx.Product * len(chunk))()
        for j, (K, M, N, bits, cap) in enumerate(chunk):
            W = cap + 1
            dA, dB, dC = mgx.DeviceBuffer(K * M * W * 8), mgx.DeviceBuffer(M * N * W * 8), mgx.DeviceBuffer(K * N * W * 8)
            seed = 1000 + 2 * (c0 + j)
            mgx.check(L.mgx_random_matrix_device(dA.ptr, M, K, M, cap, bits, bits, 0.0, seed, s))
            mgx.check(L.mgx_random_matrix_device(dB.ptr, N, M, N, cap, bits, bits, 0.0, seed + 1, s))
            bufs.append((dA, dB, dC))
            prods[j] = mgx.Product(K, M, N, dC.ptr, N, dA.ptr, M, dB.ptr, N, cap, bits, bits, -1)
        b = mgx.c_vp()
        mgx.check(L.mgx_batch_create(ctypes.byref(b), prods, len(chunk), s))
        for j in range(len(chunk)):
            routes[L.mgx_batch_route(b, j)] += 1
        mgx.check(L.mgx_batch_run(b, s))          # warm-up replay
        mgx.call("mgx_check_error", s)
        mgx.call("mgx_stream_sync", s)
        steps = max(args.steps, 1)
        t0 = time.perf_counter()
        for _ in range(steps):
            mgx.check(L.mgx_batch_run(b, s))
        mgx.call("mgx_stream_sync", s)
        total_s += (time.perf_counter() - t0) / steps
        mgx.call("mgx_check_error", s)
        for j in range(0, len(chunk), 8):   # the batch against the single-product CRT route
            K, M, N, bits, cap = chunk[j]
            dA, dB, dC = bufs[j]
            ref = mgx.DeviceBuffer(K * N * (cap + 1) * 8)
            mgx.call("mgx_crt_mfma_matmul", None, K, M, N, ref.ptr, N, dA.ptr, M, dB.ptr, N, cap, s)
            mgx.call("mgx_stream_sync", s)
            exact = exact and same_entries(dC.to_array((K, N, cap + 1)), ref.to_array((K, N, cap + 1)), cap)
            checked += 1
            ref.free()
        scratch_max = max(scratch_max, L.mgx_batch_scratch_bytes(b))
        L.mgx_batch_destroy(b)
        for t in bufs:
            for d in t:
                d.free()
        log("products %d..%d: %.3f ms per replay so far" % (c0, c0 + len(chunk) - 1, total_s * 1e3))
    L.mgx_dispatch_set_model(None)
    L.mgx_costmodel_destroy(model)
    line = {
        "metric": "mixed batch: %d exact products wall-s" % len(specs),
        "value": total_s, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": total_s * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "int8 MFMA / int64",
        "data": "synthetic: seeded entries per product, uniform magnitudes, random signs",
        "config": {"workload": "%d products, n~U[16,512], bits log-uniform [64,16384], 10%% K x 2K x K; "
                               "learned dispatch; hipGraph replays of %d-product work-queue batches"
                               % (len(specs), args.chunk)},
        "products_per_s": len(specs) / total_s,
        "entry_products_per_s": entry_products / total_s,
        "routes": {"crt": routes[0], "plain": routes[1], "winograd1": routes[2], "winograd2": routes[3]},
        "scratch_bytes_max": scratch_max,
        "bit_exact": {"vs_single_product_crt": exact, "products_checked": checked},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
