"""Headline benchmark: one exact n=1024 product of signed 2048-bit entries on one
MI355X through the INT8-MFMA CRT route, timed beside the plain GPU mpn route
(Winograd over per-entry multiplications) on the same inputs, with the
result compared bit for bit between the two routes and against the oracle on
sampled entries.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--plain-steps P]
                  [--n 1024] [--bits 2048] [--cpu-seconds S]
  python bench.py --batch [--products 512] [--chunk 32] [--steps K]

--batch times the mixed batch instead (one more JSON line): 512 independent
exact products, n uniform in [16, 512], entry bits log-uniform in [64, 16384],
10% rectangular (K x 2K x K), routed by the committed cost model and run as
hipGraph replays of the persistent-kernel work queues, chunk products per
graph; every 8th product is compared with the single-product CRT route.
"""
import argparse
import ctypes
import json
import os
import random
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bachelor-thesis_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
INT8_PEAK_TOPS = 5000.0   # dense INT8 MFMA, MI355X


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pmc_traffic(kernel, n, bits):
    """HBM bytes per dispatch of `kernel` at this config, from the committed
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--bits", type=int, default=2048)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
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
    log("generating %dx%d %d-bit inputs (seeds 1, 2)" % (n, n, bits))
    A = mgx.random_matrix(n, n, cap, bits, 1)
    B = mgx.random_matrix(n, n, cap, bits, 2)
    dA, dB = mgx.DeviceBuffer.from_array(A), mgx.DeviceBuffer.from_array(B)
    dC = mgx.DeviceBuffer(n * n * W * 8)
    dP = mgx.DeviceBuffer(n * n * W * 8)
    stream = mgx.c_vp()
    mgx.call("mgx_stream_create", ctypes.byref(stream))
    s = stream.value

    def crt_step():
        mgx.call("mgx_crt_mfma_matmul", None, n, n, n, dC.ptr, n, dA.ptr, n, dB.ptr, n, cap, s)

    # ---- CRT-MFMA route: whole job per step (bit scan, plan, 3 kernels, error check)
    for _ in range(args.warmup):
        crt_step()
    mgx.call("mgx_device_sync")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        crt_step()
    mgx.call("mgx_device_sync")
    crt_s = (time.perf_counter() - t0) / args.steps
    log("crt route: %.3f ms per product" % (crt_s * 1e3))

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
    Y = mgx.DeviceBuffer(P * ystride * 4)
    ev = [mgx.c_vp() for _ in range(5)]
    for e in ev:
        mgx.call("mgx_event_create", ctypes.byref(e))
    reps = max(args.steps, 3)
    stage_ms = [0.0, 0.0, 0.0, 0.0]
    for it in range(reps + 1):
        mgx.call("mgx_event_record", ev[0].value, s)
        mgx.call("mgx_crt_reduce", plan, 0, n, n, dA.ptr, n, cap, 0, RA.ptr, s)
        mgx.call("mgx_event_record", ev[1].value, s)
        mgx.call("mgx_crt_reduce", plan, 1, n, n, dB.ptr, n, cap, 0, RB.ptr, s)
        mgx.call("mgx_event_record", ev[2].value, s)
        mgx.call("mgx_crt_residue_gemm", plan, n, n, n, RA.ptr, RB.ptr, Y.ptr, s)
        mgx.call("mgx_event_record", ev[3].value, s)
        mgx.call("mgx_crt_reconstruct", plan, n, n, Y.ptr, dC.ptr, n, cap, s)
        mgx.call("mgx_event_record", ev[4].value, s)
        mgx.call("mgx_check_error", s)
        if it == 0:
            continue   # warm-up
        for k in range(4):
            ms = ctypes.c_float()
            mgx.call("mgx_event_elapsed", ctypes.byref(ms), ev[k].value, ev[k + 1].value)
            stage_ms[k] += ms.value / reps
    log("stages ms: reduceA %.3f reduceB %.3f gemm %.3f reconstruct %.3f (P=%d)" % tuple(stage_ms + [P]))
    gemm_ops = 2.0 * n ** 3 * S * S * P
    gemm_tops = gemm_ops / (stage_ms[2] * 1e-3) / 1e12
    gemm_kernel = "crt_gemm2_kernel" if S == 2 else "crt_gemm_kernel"
    traffic, traffic_src = pmc_traffic(gemm_kernel, n, bits)
    C = dC.to_array((n, n, W))

    # ---- plain route (Winograd over per-entry mpn multiplication), same inputs
    plain_s = None
    bit_exact_routes = None
    if not args.no_plain:
        def plain_step():
            dP.zero(s)
            mgx.call("mgx_winograd", n, dP.ptr, n, dA.ptr, n, dB.ptr, n, cap, s)
        plain_step()
        mgx.call("mgx_device_sync")
        t0 = time.perf_counter()
        for _ in range(args.plain_steps):
            plain_step()
        mgx.call("mgx_device_sync")
        plain_s = (time.perf_counter() - t0) / args.plain_steps
        log("plain route: %.3f s per product" % plain_s)
        Pm = dP.to_array((n, n, W))
        sz = C[..., cap].view(np.int64)
        same = bool((sz == Pm[..., cap].view(np.int64)).all())
        if same:
            mask = np.arange(cap)[None, None, :] < np.abs(sz)[..., None]
            same = bool((np.where(mask, C[..., :cap], 0) == np.where(mask, Pm[..., :cap], 0)).all())
        bit_exact_routes = same

    # ---- oracle on sampled entries (exact dot products) and the CPU baseline
    import oracle
    rng = random.Random(1234)
    rows = rng.sample(range(n), 2)
    cols = rng.sample(range(n), 2)
    sub = oracle.mat_mul(np.ascontiguousarray(A[rows]), np.ascontiguousarray(B[:, cols]), cap)
    sample_ok = all(mgx.entry_to_int(sub[i, j]) == mgx.entry_to_int(C[r, c])
                    for i, r in enumerate(rows) for j, c in enumerate(cols))

    cpu = None
    if not args.no_cpu:
        orc = oracle.lib()
        ver = orc.or_matcpu_gmp_version()
        threads = min(os.cpu_count() or 1, 16)
        # bounded sample: exact entries C[y][x] by GMP dot products (mat-cpu's
        # arithmetic) on all host threads; grow the sample until ~cpu_seconds
        cnt, secs = 16, 0.0
        while True:
            yx = np.array([[rng.randrange(n), rng.randrange(n)] for _ in range(cnt)], dtype=np.int64)
            out = np.zeros((cnt, W), dtype=np.uint64)
            t = ctypes.c_double()
            st = orc.or_matcpu_sample(n, A.ctypes.data, B.ctypes.data, cap, yx.ctypes.data, cnt, out.ctypes.data,
                                      threads, ctypes.byref(t))
            secs = t.value
            if st or secs >= args.cpu_seconds / 4 or cnt >= 1 << 20:
                break
            cnt = int(cnt * max(2.0, min(8.0, args.cpu_seconds / 4 / max(secs, 1e-3))))
        ok = st == 0 and all(mgx.entry_to_int(out[i]) == mgx.entry_to_int(C[yx[i, 0], yx[i, 1]])
                             for i in range(min(cnt, 64)))
        per_entry = secs / cnt
        cpu = {"value": per_entry * n * n, "unit": "s (extrapolated full product)", "cores": threads,
               "kind": "port",
               "sample": "%d exact entries of the same n=%d %d-bit product by GMP %s mpz dot products "
                         "(mat-cpu's arithmetic, naive order), %.2f s on %d threads, scaled to n^2 entries; "
                         "entries match the GPU: %s" % (cnt, n, bits, ver.decode() if ver else "?", secs, threads, ok)}
        log("cpu baseline: %.1f s extrapolated (%d entries in %.2f s)" % (cpu["value"], cnt, secs))

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
        "data": "synthetic: seeded xoshiro256** entries, uniform %d-bit magnitudes, random signs" % bits,
        "config": {"workload": "C=A*B, n=%d square, signed %d-bit entries, CRT-MFMA route (P=%d primes < %s, "
                               "%d int8 slices)" % (n, bits, P, "65280" if S == 2 else "2^31", S),
                   "n": n, "bits": bits, "primes": P, "slices": S, "cap_limbs": cap},
        "entry_products_per_s": n ** 3 / crt_s,
        "plain_route_s": plain_s,
        "speedup_vs_plain": (plain_s / crt_s) if plain_s else None,
        "bit_exact": {"crt_vs_plain_all_entries": bit_exact_routes, "crt_vs_oracle_sample": sample_ok},
        "stage_ms": {"reduce_A": stage_ms[0], "reduce_B": stage_ms[1], "residue_gemm": stage_ms[2],
                     "reconstruct": stage_ms[3]},
        "roofline": {"bound": "mfma", "kernel": gemm_kernel, "achieved": gemm_tops, "peak": INT8_PEAK_TOPS,
                     "unit": "TFLOP/s", "frac": gemm_tops / INT8_PEAK_TOPS, "traffic": traffic,
                     "note": "int8 ops 2*n^3*S^2*P per launch / event-timed launch duration; traffic = HBM bytes "
                             "per launch from %s" % (traffic_src or "no PMC summary for this config")},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    mgx.call("mgx_crt_plan_destroy", plan)


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
    mpath = os.path.join(ROOT, "bachelor-thesis_amd", "costmodel_params.txt")
    mgx.check(L.mgx_costmodel_load(ctypes.byref(model), mpath.encode()))
    mgx.check(L.mgx_dispatch_set_model(model))
    stream = mgx.c_vp()
    mgx.call("mgx_stream_create", ctypes.byref(stream))
    s = stream.value
    total_s, routes, checked, exact = 0.0, [0, 0, 0, 0], 0, True
    entry_products = sum(K * M * N for K, M, N, _, _ in specs)
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
            x, y = dC.to_array((K, N, cap + 1)), ref.to_array((K, N, cap + 1))
            sz = x[..., cap].view(np.int64)
            ok = bool((sz == y[..., cap].view(np.int64)).all())
            if ok:
                mask = np.arange(cap)[None, None, :] < np.abs(sz)[..., None]
                ok = bool((np.where(mask, x[..., :cap], 0) == np.where(mask, y[..., :cap], 0)).all())
            exact = exact and ok
            checked += 1
            ref.free()
        scratch = L.mgx_batch_scratch_bytes(b)
        L.mgx_batch_destroy(b)
        for t in bufs:
            for d in t:
                d.free()
        log("products %d..%d: %.3f ms per replay, scratch %.2f GB" % (c0, c0 + len(chunk) - 1, total_s * 1e3,
                                                                      scratch / 1e9))
    L.mgx_dispatch_set_model(None)
    L.mgx_costmodel_destroy(model)
    line = {
        "metric": "mixed batch: %d exact products wall-s" % len(specs),
        "value": total_s, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": total_s * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "int8 MFMA / int64",
        "data": "synthetic: seeded xoshiro256** entries per product, uniform magnitudes, random signs",
        "config": {"workload": "%d products, n~U[16,512], bits log-uniform [64,16384], 10%% K x 2K x K; "
                               "learned dispatch; hipGraph replays of %d-product work-queue batches"
                               % (len(specs), args.chunk)},
        "products_per_s": len(specs) / total_s,
        "entry_products_per_s": entry_products / total_s,
        "routes": {"crt": routes[0], "plain": routes[1] + routes[2] + routes[3]},
        "bit_exact": {"vs_single_product_crt": exact, "products_checked": checked},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
