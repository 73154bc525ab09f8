"""Scheduling side on the GPU: the cost model's HIP kernels against the
float64 oracle, every route against the oracle, the batched persistent
executor (graph replays) against the oracle, and the autotuner."""
import ctypes

import numpy as np
import pytest

import oracle
from helpers import is_normalized, matmul, rand_ints, to_entries, to_ints

pytestmark = pytest.mark.gpu

vp, i64 = ctypes.c_void_p, ctypes.c_int64


def sched(mgx):
    L = mgx.lib()
    fp = ctypes.POINTER(ctypes.c_float)
    L.mgx_costmodel_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64]
    L.mgx_costmodel_destroy.argtypes = [vp]
    L.mgx_costmodel_param_count.argtypes = [vp]
    L.mgx_costmodel_param_count.restype = i64
    L.mgx_costmodel_get_params.argtypes = [vp, vp]
    L.mgx_costmodel_set_params.argtypes = [vp, vp]
    L.mgx_costmodel_set_normalisation.argtypes = [vp, vp, vp]
    L.mgx_costmodel_loss_grad.argtypes = [vp, vp, vp, vp, ctypes.c_int, vp, fp]
    L.mgx_costmodel_train.argtypes = [vp, vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, fp]
    L.mgx_costmodel_predict.argtypes = [vp, vp, ctypes.c_int, vp]
    L.mgx_matmul_route.argtypes = [ctypes.c_int, i64, i64, i64, vp, i64, vp, i64, vp, i64, i64, vp]
    L.mgx_random_matrix_device.argtypes = [vp, i64, i64, i64, i64, i64, i64, ctypes.c_double, ctypes.c_uint64, vp]
    L.mgx_tune_time_routes.argtypes = [i64, i64, i64, vp, vp, vp, i64, ctypes.c_int, vp, ctypes.POINTER(ctypes.c_uint)]
    L.mgx_dispatch_choose.argtypes = [i64, i64, i64, vp, i64, vp, i64, i64, ctypes.POINTER(ctypes.c_int)]
    L.mgx_dispatch_set_model.argtypes = [vp]
    L.mgx_tune_features_device.argtypes = [i64, i64, i64, vp, i64, vp, i64, i64, vp]
    L.mgx_batch_create.argtypes = [ctypes.POINTER(vp), vp, i64, vp]
    L.mgx_batch_run.argtypes = [vp, vp]
    L.mgx_batch_route.argtypes = [vp, i64]
    L.mgx_batch_destroy.argtypes = [vp]
    return L


class Product(ctypes.Structure):
    _fields_ = [("K", i64), ("M", i64), ("N", i64), ("C", vp), ("Crs", i64), ("A", vp), ("Ars", i64), ("B", vp),
                ("Brs", i64), ("cap", i64), ("bitsA", i64), ("bitsB", i64), ("route", ctypes.c_int)]


def test_mlp_gradient_matches_oracle(gpu):
    mgx = gpu
    L = sched(mgx)
    D, H, O, B = 13, 64, 4, 96
    cm = vp()
    mgx.check(L.mgx_costmodel_create(ctypes.byref(cm), D, H, O, 7))
    try:
        npar = L.mgx_costmodel_param_count(cm)
        rng = np.random.default_rng(3)
        X = rng.normal(size=(B, D)).astype(np.float32)
        T = rng.normal(size=(B, O)).astype(np.float32)
        mask = (rng.random((B, O)) < 0.8).astype(np.float32)
        params = np.zeros(npar, dtype=np.float32)
        mgx.check(L.mgx_costmodel_get_params(cm, params.ctypes.data))
        g = np.zeros(npar, dtype=np.float32)
        loss = ctypes.c_float()
        mgx.check(L.mgx_costmodel_loss_grad(cm, X.ctypes.data, T.ctypes.data, mask.ctypes.data, B, g.ctypes.data,
                                            ctypes.byref(loss)))
        p64, X64, T64, M64 = (a.astype(np.float64) for a in (params, X, T, mask))
        g64, Y64 = np.zeros(npar), np.zeros((B, O))
        ref = oracle.lib().or_mlp_loss_grad(D, H, O, B, p64.ctypes.data, X64.ctypes.data, T64.ctypes.data,
                                            M64.ctypes.data, g64.ctypes.data, Y64.ctypes.data)
        assert abs(loss.value - ref) <= 1e-4 * max(1.0, abs(ref))
        # FP32 kernels vs float64 reference: 1e-4 relative in norm, 1e-3 per large component
        assert np.linalg.norm(g - g64) <= 1e-4 * np.linalg.norm(g64)
        big = np.abs(g64) > 1e-3 * np.abs(g64).max()
        assert np.all(np.abs(g[big] - g64[big]) <= 1e-3 * np.abs(g64[big]) + 1e-6)
    finally:
        L.mgx_costmodel_destroy(cm)


def test_mlp_training_fits_a_smooth_target(gpu):
    mgx = gpu
    L = sched(mgx)
    D, H, O, B = 13, 64, 4, 512
    cm = vp()
    mgx.check(L.mgx_costmodel_create(ctypes.byref(cm), D, H, O, 1))
    try:
        rng = np.random.default_rng(0)
        X = rng.normal(size=(B, D)).astype(np.float32)
        T = np.stack([X[:, 0] + 0.5 * X[:, 1], X[:, 2] * X[:, 3], np.abs(X[:, 4]), X[:, 5] - X[:, 6]], 1).astype(np.float32)
        mask = np.ones((B, O), dtype=np.float32)
        l0, l1 = ctypes.c_float(), ctypes.c_float()
        mgx.check(L.mgx_costmodel_loss_grad(cm, X.ctypes.data, T.ctypes.data, mask.ctypes.data, B, None, ctypes.byref(l0)))
        mgx.check(L.mgx_costmodel_train(cm, X.ctypes.data, T.ctypes.data, mask.ctypes.data, B, 400, 3e-3, ctypes.byref(l1)))
        assert l1.value < 0.1 * l0.value
        Y = np.zeros((B, O), dtype=np.float32)
        mgx.check(L.mgx_costmodel_predict(cm, X.ctypes.data, B, Y.ctypes.data))
        assert np.mean((Y - T) ** 2) == pytest.approx(l1.value, rel=0.05)
    finally:
        L.mgx_costmodel_destroy(cm)


@pytest.mark.parametrize("route", [0, 1, 2, 3])
def test_every_route_is_exact(gpu, route):
    mgx = gpu
    L = sched(mgx)
    K, M, N, bits, cap = 12, 20, 9, 700, 26
    A, B = rand_ints(K, M, bits, 40 + route), rand_ints(M, N, bits, 50 + route)
    dA, dB = mgx.DeviceBuffer.from_array(to_entries(A, cap)), mgx.DeviceBuffer.from_array(to_entries(B, cap))
    dC = mgx.DeviceBuffer(K * N * (cap + 1) * 8)
    dC.zero()
    mgx.check(L.mgx_matmul_route(route, K, M, N, dC.ptr, N, dA.ptr, M, dB.ptr, N, cap, None))
    assert to_ints(dC.to_array((K, N, cap + 1))) == matmul(A, B)


@pytest.mark.parametrize("schedule", range(9))
def test_every_schedule_is_exact(gpu, schedule):
    """Each point of the tuner's schedule space (digit count, GEMM tile,
    Toom cutoff, Winograd depth) gives the exact product; 4000-bit operands
    are two nearly full 64-word blocks, where the Karatsuba step fires."""
    mgx = gpu
    L = mgx.lib()
    sc = (ctypes.c_int * 5)()
    mgx.check(L.mgx_schedule_get(schedule, sc))
    for (K, M, N, bits, seed) in [(20, 24, 18, 4000, 1), (130, 9, 70, 200, 2)]:
        cap = (2 * bits + 8 + 63) // 64 + 2
        A, B = rand_ints(K, M, bits, seed + 10 * schedule), rand_ints(M, N, bits, seed + 10 * schedule + 5)
        dA, dB = mgx.DeviceBuffer.from_array(to_entries(A, cap)), mgx.DeviceBuffer.from_array(to_entries(B, cap))
        dC = mgx.DeviceBuffer(K * N * (cap + 1) * 8)
        mgx.check(L.mgx_matmul_schedule(schedule, K, M, N, dC.ptr, N, dA.ptr, M, dB.ptr, N, cap, None))
        C = dC.to_array((K, N, cap + 1))
        assert to_ints(C) == matmul(A, B), (list(sc), K, M, N)
        assert is_normalized(C)


def test_random_matrix_device(gpu):
    mgx = gpu
    L = sched(mgx)
    W, H, cap = 64, 32, 20
    d = mgx.DeviceBuffer(W * H * (cap + 1) * 8)
    mgx.check(L.mgx_random_matrix_device(d.ptr, W, H, W, cap, 100, 1200, 0.25, 5, None))
    X = d.to_array((H, W, cap + 1))
    assert is_normalized(X)
    vals = [v for r in to_ints(X) for v in r]
    big = sum(abs(v).bit_length() > 100 for v in vals)
    assert 0.15 < big / len(vals) < 0.35
    assert max(abs(v).bit_length() for v in vals) <= 1200
    assert 0.4 < sum(v < 0 for v in vals) / len(vals) < 0.6
    d2 = mgx.DeviceBuffer(W * H * (cap + 1) * 8)
    mgx.check(L.mgx_random_matrix_device(d2.ptr, W, H, W, cap, 100, 1200, 0.25, 5, None))
    assert (d2.to_array((H, W, cap + 1)) == X).all()


def test_batch_graph_replays_exact(gpu):
    mgx = gpu
    L = sched(mgx)
    rng = np.random.default_rng(11)
    specs = []
    for i in range(10):
        n = int(rng.integers(3, 40))
        rect = i % 4 == 3
        K, M, N = n, (2 * n if rect else n), n
        bits = int(2 ** rng.uniform(6, 11))
        route = [0, 1, -1][i % 3]
        specs.append((K, M, N, bits, route))
    specs.append((1, 1, 1, 64, 0))
    bufs, prods, wants = [], [], []
    for i, (K, M, N, bits, route) in enumerate(specs):
        cap = (2 * bits + 16 + 63) // 64 + 1
        A, B = rand_ints(K, M, bits, 100 + i), rand_ints(M, N, bits, 200 + i)
        dA, dB = mgx.DeviceBuffer.from_array(to_entries(A, cap)), mgx.DeviceBuffer.from_array(to_entries(B, cap))
        dC = mgx.DeviceBuffer(K * N * (cap + 1) * 8)
        dC.zero()
        bufs.append((dA, dB, dC, K, N, cap))
        prods.append(Product(K, M, N, dC.ptr, N, dA.ptr, M, dB.ptr, N, cap, 0, 0, route))
        wants.append(matmul(A, B))
    arr = (Product * len(prods))(*prods)
    b = vp()
    mgx.check(L.mgx_batch_create(ctypes.byref(b), ctypes.cast(arr, vp), len(prods), None))
    try:
        routes = [L.mgx_batch_route(b, i) for i in range(len(prods))]
        assert set(routes) <= {0, 1} and 0 in routes and 1 in routes
        for rep in range(3):
            for (_, _, dC, K, N, cap) in bufs:
                dC.zero()
            mgx.check(L.mgx_batch_run(b, None))
            mgx.call("mgx_check_error", None)
            for (_, _, dC, K, N, cap), want in zip(bufs, wants):
                got = dC.to_array((K, N, cap + 1))
                assert to_ints(got) == want, rep
                assert is_normalized(got)
    finally:
        L.mgx_batch_destroy(b)


def test_tuner_and_dispatch(gpu):
    mgx = gpu
    L = sched(mgx)
    K = M = N = 32
    bits, cap = 256, 10
    dA, dB, dC = (mgx.DeviceBuffer(K * M * (cap + 1) * 8) for _ in range(3))
    mgx.check(L.mgx_random_matrix_device(dA.ptr, M, K, M, cap, bits, bits, 0.0, 1, None))
    mgx.check(L.mgx_random_matrix_device(dB.ptr, N, M, N, cap, bits, bits, 0.0, 2, None))
    t = np.zeros(4)
    est = ctypes.c_uint()
    mgx.check(L.mgx_tune_time_routes(K, M, N, dC.ptr, dA.ptr, dB.ptr, cap, 3, t.ctypes.data, ctypes.byref(est)))
    assert (t > 0).all()
    r = ctypes.c_int()
    mgx.check(L.mgx_dispatch_choose(K, M, N, dA.ptr, M, dB.ptr, N, cap, ctypes.byref(r)))
    assert 0 <= r.value < 4
    f = np.zeros(13, dtype=np.float32)
    mgx.check(L.mgx_tune_features_device(K, M, N, dA.ptr, M, dB.ptr, N, cap, f.ctypes.data))
    assert f[0] == 5 and f[5] == pytest.approx(8.0, abs=0.01) and abs(f[7:].sum() - 1) < 1e-6


def test_committed_model_loads_and_dispatches(gpu):
    """The trained model file: the C loader gives the same predictions as the
    parameters set from the JSON copy, and MGX_ROUTE_AUTO with it installed is
    exact."""
    import json
    import os
    mgx = gpu
    L = sched(mgx)
    L.mgx_costmodel_load.argtypes = [ctypes.POINTER(vp), ctypes.c_char_p]
    here = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "bachelor-thesis_amd")
    cm = vp()
    mgx.check(L.mgx_costmodel_load(ctypes.byref(cm), os.path.join(here, "costmodel_params.txt").encode()))
    with open(os.path.join(here, "costmodel_params.json")) as f:
        d = json.load(f)
    cm2 = vp()
    mgx.check(L.mgx_costmodel_create(ctypes.byref(cm2), d["D"], d["H"], d["O"], 0))
    p = np.array(d["params"], dtype=np.float32)
    mean, scale = np.array(d["mean"], dtype=np.float32), np.array(d["scale"], dtype=np.float32)
    mgx.check(L.mgx_costmodel_set_params(cm2, p.ctypes.data))
    mgx.check(L.mgx_costmodel_set_normalisation(cm2, mean.ctypes.data, scale.ctypes.data))
    X = (mean + scale * np.random.default_rng(3).standard_normal((17, d["D"]))).astype(np.float32)
    Y1, Y2 = np.zeros((17, d["O"]), dtype=np.float32), np.zeros((17, d["O"]), dtype=np.float32)
    mgx.check(L.mgx_costmodel_predict(cm, X.ctypes.data, 17, Y1.ctypes.data))
    mgx.check(L.mgx_costmodel_predict(cm2, X.ctypes.data, 17, Y2.ctypes.data))
    assert np.allclose(Y1, Y2, rtol=1e-6, atol=1e-5)
    # one sample at a time takes the fused three-layer kernel: FP32 sums in a
    # different order, so agreement to FP32 rounding of the batched kernels
    for b in range(17):
        y1 = np.zeros(d["O"], dtype=np.float32)
        mgx.check(L.mgx_costmodel_predict(cm, np.ascontiguousarray(X[b]).ctypes.data, 1, y1.ctypes.data))
        assert np.allclose(y1, Y1[b], rtol=1e-5, atol=1e-4), b
    try:
        mgx.check(L.mgx_dispatch_set_model(cm))
        K, M, N, bits, cap = 19, 23, 17, 900, 32
        A, B = rand_ints(K, M, bits, 41), rand_ints(M, N, bits, 42)
        dA, dB = mgx.DeviceBuffer.from_array(to_entries(A, cap)), mgx.DeviceBuffer.from_array(to_entries(B, cap))
        dC = mgx.DeviceBuffer(K * N * (cap + 1) * 8)
        r = ctypes.c_int()
        mgx.check(L.mgx_dispatch_choose(K, M, N, dA.ptr, M, dB.ptr, N, cap, ctypes.byref(r)))
        assert 0 <= r.value < 4
        mgx.check(L.mgx_matmul_route(-1, K, M, N, dC.ptr, N, dA.ptr, M, dB.ptr, N, cap, None))
        C = dC.to_array((K, N, cap + 1))
        assert to_ints(C) == matmul(A, B) and is_normalized(C)
    finally:
        L.mgx_dispatch_set_model(None)
        L.mgx_costmodel_destroy(cm)
        L.mgx_costmodel_destroy(cm2)


def test_device_dispatch_matches_host_inference(gpu):
    """MGX_ROUTE_AUTO runs the features and the network on the device beside
    its scan; its choice equals the host-side features + inference kernel
    (mgx_dispatch_choose_schedule) on varied products, and the result is exact."""
    import os
    mgx = gpu
    L = mgx.lib()
    here = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "bachelor-thesis_amd")
    cm = vp()
    mgx.check(L.mgx_costmodel_load(ctypes.byref(cm), os.path.join(here, "costmodel_params.txt").encode()))
    try:
        mgx.check(L.mgx_dispatch_set_model(cm))
        rng = np.random.default_rng(8)
        for i in range(16):
            K, M, N = (int(x) for x in rng.integers(2, 70, 3))
            bits = int(2 ** rng.uniform(5, 13))
            cap = (2 * bits + 8 + 63) // 64 + 2
            A, B = rand_ints(K, M, bits, 600 + i), rand_ints(M, N, bits, 700 + i)
            dA, dB = mgx.DeviceBuffer.from_array(to_entries(A, cap)), mgx.DeviceBuffer.from_array(to_entries(B, cap))
            dC = mgx.DeviceBuffer(K * N * (cap + 1) * 8)
            host = ctypes.c_int()
            mgx.check(L.mgx_dispatch_choose_schedule(K, M, N, dA.ptr, M, dB.ptr, N, cap, ctypes.byref(host)))
            mgx.check(L.mgx_matmul_route(-1, K, M, N, dC.ptr, N, dA.ptr, M, dB.ptr, N, cap, None))
            assert L.mgx_dispatch_last() == host.value, (i, K, M, N, bits)
            assert to_ints(dC.to_array((K, N, cap + 1))) == matmul(A, B)
    finally:
        L.mgx_dispatch_set_model(None)
        L.mgx_costmodel_destroy(cm)
