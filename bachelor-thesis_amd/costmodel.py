"""Train and evaluate the learned schedule dispatcher.

    python bachelor-thesis_amd/costmodel.py gen    --data timings.npz [--configs 5000]
    python bachelor-thesis_amd/costmodel.py train  [--data timings.npz] [--configs 5000] [--out costmodel_params.json]
    python bachelor-thesis_amd/costmodel.py eval   [--params costmodel_params.json] [--json out.json]

`gen` times every schedule (mgx_tune_time_schedules: the exhaustive
autotuner over route, digit count, residue-GEMM tile, Toom cutoff and
Winograd depth) on seeded synthetic products of random shape and bit-length
mix; `train` fits the FP32 MLP (3 layers, 64 hidden) with the library's own
HIP forward/backward/Adam kernels to predict log2 of each schedule's time from
the product's features.  `eval` runs held-out heavy-tailed mixes through
MGX_ROUTE_AUTO with the model installed (so the dispatch's own scan and
inference are in the timed cost), compares with the exhaustive best
(regret = t(auto)/t(best) - 1), checks the output bit for bit against the
default schedule, and reports what the network gains over fixed dispatch and
over tuning at run time.

The model only picks a schedule; every route is exact.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import mgx  # noqa: E402

vp, i64 = ctypes.c_void_p, ctypes.c_int64
D, H, O = 13, 64, 9
SCHEDULES = ["crt2", "crt2_t64", "crt4", "plain", "plain_notoom", "wino1", "wino2", "wino1_notoom", "wino2_notoom"]
DEFAULT_PARAMS = os.path.join(HERE, "costmodel_params.json")


def lib():
    L = mgx.lib()
    fp = ctypes.POINTER(ctypes.c_float)
    L.mgx_costmodel_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64]
    L.mgx_costmodel_destroy.argtypes = [vp]
    L.mgx_costmodel_param_count.argtypes = [vp]
    L.mgx_costmodel_param_count.restype = i64
    L.mgx_costmodel_get_params.argtypes = [vp, vp]
    L.mgx_costmodel_set_params.argtypes = [vp, vp]
    L.mgx_costmodel_set_normalisation.argtypes = [vp, vp, vp]
    L.mgx_costmodel_train.argtypes = [vp, vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, fp]
    L.mgx_costmodel_loss_grad.argtypes = [vp, vp, vp, vp, ctypes.c_int, vp, fp]
    L.mgx_costmodel_predict.argtypes = [vp, vp, ctypes.c_int, vp]
    L.mgx_random_matrix_device.argtypes = [vp, i64, i64, i64, i64, i64, i64, ctypes.c_double, ctypes.c_uint64, vp]
    L.mgx_tune_time_routes.argtypes = [i64, i64, i64, vp, vp, vp, i64, ctypes.c_int, vp, ctypes.POINTER(ctypes.c_uint)]
    L.mgx_tune_features_device.argtypes = [i64, i64, i64, vp, i64, vp, i64, i64, vp]
    L.mgx_dispatch_set_model.argtypes = [vp]
    L.mgx_matmul_route.argtypes = [ctypes.c_int, i64, i64, i64, vp, i64, vp, i64, vp, i64, i64, vp]
    return L


class Problem:
    """A seeded synthetic product resident on the device."""

    def __init__(self, K, M, N, blo, bhi, frac, seed):
        L = lib()
        self.K, self.M, self.N = K, M, N
        self.cap = (2 * max(blo, bhi) + 24 + 63) // 64 + 2
        W = self.cap + 1
        self.A = mgx.DeviceBuffer(K * M * W * 8)
        self.B = mgx.DeviceBuffer(M * N * W * 8)
        self.C = mgx.DeviceBuffer(K * N * W * 8)
        mgx.check(L.mgx_random_matrix_device(self.A.ptr, M, K, M, self.cap, blo, bhi, frac, seed, None))
        mgx.check(L.mgx_random_matrix_device(self.B.ptr, N, M, N, self.cap, blo, bhi, frac, seed + 7919, None))

    def features(self):
        f = np.zeros(D, dtype=np.float32)
        mgx.check(lib().mgx_tune_features_device(self.K, self.M, self.N, self.A.ptr, self.M, self.B.ptr, self.N,
                                                 self.cap, f.ctypes.data))
        return f

    def times(self, reps=1):
        t = np.zeros(O)
        est = ctypes.c_uint()
        mgx.check(lib().mgx_tune_time_schedules(self.K, self.M, self.N, self.C.ptr, self.A.ptr, self.B.ptr, self.cap,
                                                reps, t.ctypes.data, ctypes.byref(est)))
        return t, est.value

    def run(self, schedule):
        """Output of one schedule (MGX_ROUTE_AUTO = -1: the dispatch's choice)."""
        self.C.zero()
        mgx.check(lib().mgx_matmul_schedule(schedule, self.K, self.M, self.N, self.C.ptr, self.N, self.A.ptr, self.M,
                                            self.B.ptr, self.N, self.cap, None))
        return self.C.to_array((self.K, self.N, self.cap + 1))

    def time_auto(self, reps=5):
        """Median wall time of MGX_ROUTE_AUTO calls (scan, inference, product)."""
        L = lib()
        args = (self.K, self.M, self.N, self.C.ptr, self.N, self.A.ptr, self.M, self.B.ptr, self.N, self.cap, None)
        mgx.check(L.mgx_matmul_schedule(-1, *args))
        t = []
        for _ in range(reps):
            t0 = time.perf_counter()
            mgx.check(L.mgx_matmul_schedule(-1, *args))
            t.append(time.perf_counter() - t0)
        return float(np.median(t))

    def chosen(self):
        sc = ctypes.c_int()
        mgx.check(lib().mgx_dispatch_choose_schedule(self.K, self.M, self.N, self.A.ptr, self.M, self.B.ptr, self.N,
                                                     self.cap, ctypes.byref(sc)))
        return sc.value

    def free(self):
        for b in (self.A, self.B, self.C):
            b.free()


def sample_config(rng):
    n = int(round(2 ** rng.uniform(4, 10)))         # 16 .. 1024
    rect = rng.random() < 0.1
    K, M, N = n, (2 * n if rect else n), n
    blo = int(round(2 ** rng.uniform(6, 14)))       # 64 .. 16384 bits
    if rng.random() < 0.5:
        bhi, frac = blo, 0.0
    else:
        bhi = int(round(2 ** rng.uniform(math.log2(blo), 14)))
        frac = float(rng.choice([0.01, 0.05, 0.1, 0.3, 0.5]))
    # keep one problem's residues and entries well inside HBM and the timing short
    while K * M * N * (max(blo, bhi) / 2048) ** 2 > 3e9 and n > 16:
        n //= 2
        K, M, N = n, (2 * n if rect else n), n
    return K, M, N, blo, bhi, frac


def generate(nconfigs, seed, log_every=100, save=None, start=None):
    """Time every route on nconfigs seeded problems.  With `save`, the arrays
    gathered so far are written there every log_every configs, and a run
    resumes from such a file (`start` = its arrays)."""
    rng = np.random.default_rng(seed)
    X, T, flags, cfgs = [], [], [], []
    if start is not None:
        X, T, flags, cfgs = [list(a) for a in start]
    t0 = time.time()
    for i in range(nconfigs):
        cfg = sample_config(rng)   # drawn for every i, so a resumed run sees the same sequence
        if i < len(X):
            continue
        K, M, N, blo, bhi, frac = cfg
        p = Problem(K, M, N, blo, bhi, frac, 1000 + i)
        X.append(p.features())
        t, est = p.times(reps=1)
        T.append(t)
        flags.append(est)
        cfgs.append([K, M, N, blo, bhi, frac])
        p.free()
        if (i + 1) % log_every == 0:
            print("timed %d/%d configs (%.0f s)" % (i + 1, nconfigs, time.time() - t0), file=sys.stderr, flush=True)
            if save:
                np.savez(save, X=np.array(X, dtype=np.float32), T=np.array(T), flags=np.array(flags),
                         cfgs=np.array(cfgs, dtype=np.float64))
    return np.array(X, dtype=np.float32), np.array(T), np.array(flags), cfgs


def targets(T):
    """log2 of microseconds; unavailable routes (>= 1e29 s) are masked out."""
    ok = T < 1e29
    Y = np.where(ok, np.log2(np.maximum(T, 1e-7) * 1e6), 0.0).astype(np.float32)
    return Y, ok.astype(np.float32)


def train_model(X, Y, mask, epochs=3000, lr=3e-3, seed=1):
    L = lib()
    cm = vp()
    mgx.check(L.mgx_costmodel_create(ctypes.byref(cm), D, H, O, seed))
    mean = X.mean(0).astype(np.float32)
    scale = (X.std(0) + 1e-3).astype(np.float32)
    mgx.check(L.mgx_costmodel_set_normalisation(cm, mean.ctypes.data, scale.ctypes.data))
    loss = ctypes.c_float()
    mgx.check(L.mgx_costmodel_train(cm, X.ctypes.data, Y.ctypes.data, mask.ctypes.data, X.shape[0], epochs, lr,
                                    ctypes.byref(loss)))
    return cm, mean, scale, loss.value


def save_model(cm, mean, scale, path, meta):
    L = lib()
    n = L.mgx_costmodel_param_count(cm)
    p = np.zeros(n, dtype=np.float32)
    mgx.check(L.mgx_costmodel_get_params(cm, p.ctypes.data))
    with open(path, "w") as f:
        json.dump({"D": D, "H": H, "O": O, "schedules": SCHEDULES, "mean": mean.tolist(), "scale": scale.tolist(),
                   "params": p.tolist(), "meta": meta}, f)
    # the same numbers as text for mgx_costmodel_load (C callers, mat-gpu -m 0)
    with open(os.path.splitext(path)[0] + ".txt", "w") as f:
        f.write("mgx-costmodel 1 %d %d %d\n" % (D, H, O))
        for arr in (mean, scale, p):
            f.write(" ".join("%.9g" % float(x) for x in arr) + "\n")


def load_model(path=DEFAULT_PARAMS):
    L = lib()
    with open(path) as f:
        d = json.load(f)
    cm = vp()
    mgx.check(L.mgx_costmodel_create(ctypes.byref(cm), d["D"], d["H"], d["O"], 0))
    p = np.array(d["params"], dtype=np.float32)
    mean, scale = np.array(d["mean"], dtype=np.float32), np.array(d["scale"], dtype=np.float32)
    mgx.check(L.mgx_costmodel_set_params(cm, p.ctypes.data))
    mgx.check(L.mgx_costmodel_set_normalisation(cm, mean.ctypes.data, scale.ctypes.data))
    return cm


def predict(cm, X):
    Y = np.zeros((X.shape[0], O), dtype=np.float32)
    mgx.check(lib().mgx_costmodel_predict(cm, np.ascontiguousarray(X, dtype=np.float32).ctypes.data, X.shape[0],
                                          Y.ctypes.data))
    return Y


HELD_OUT = [
    # (name, K, M, N, bits_lo, bits_hi, frac_hi); the last four are small
    # products where the per-entry plain route beats the CRT's fixed costs
    ("n768_90pc256_10pc32768", 768, 768, 768, 256, 32768, 0.10),
    ("n1024_2048", 1024, 1024, 1024, 2048, 2048, 0.0),
    ("n64_64bit", 64, 64, 64, 64, 64, 0.0),
    ("n24_128bit", 24, 24, 24, 128, 128, 0.0),
    ("n200_1024_2048mix", 200, 200, 200, 1024, 2048, 0.5),
    ("n96_rect_90pc128_10pc4096", 96, 192, 96, 128, 4096, 0.10),
    ("n32_8192", 32, 32, 32, 8192, 8192, 0.0),
    ("n400_99pc64_1pc16384", 400, 400, 400, 64, 16384, 0.01),
    ("n16_16384", 16, 16, 16, 16384, 16384, 0.0),
    ("n300_512", 300, 300, 300, 512, 512, 0.0),
    ("n19_1600", 19, 19, 19, 1600, 1600, 0.0),
    ("n36_157_1pc5763", 36, 36, 36, 157, 5763, 0.01),
    ("n20_391_10pc10466", 20, 20, 20, 391, 10466, 0.10),
    ("n28_rect_187_5pc1899", 28, 56, 28, 187, 1899, 0.05),
]


def evaluate(cm, reps=3, check_exact=True, mixes=None):
    """Held-out mixes through MGX_ROUTE_AUTO with `cm` installed."""
    L = lib()
    mgx.check(L.mgx_dispatch_set_model(cm))
    rows = []
    try:
        for i, (name, K, M, N, blo, bhi, frac) in enumerate(mixes or HELD_OUT):
            p = Problem(K, M, N, blo, bhi, frac, 777 + i)
            f = p.features()
            y = predict(cm, f[None, :])[0]
            choice = p.chosen()
            t, est = p.times(reps=reps)
            t_auto = p.time_auto(reps=max(reps, 5))
            best = int(np.argmin(t))
            exact = None
            if check_exact:
                exact = bool((p.run(-1) == p.run(0)).all())
            rows.append({"mix": name, "K": K, "M": M, "N": N, "choice": SCHEDULES[choice], "best": SCHEDULES[best],
                         "times_s": [float(x) for x in t], "estimated_mask": int(est), "t_auto_s": t_auto,
                         "regret": float(t_auto / t[best] - 1.0),
                         "regret_schedule_only": float(t[choice] / t[best] - 1.0),
                         "predicted_log2_us": [float(v) for v in y], "bit_identical": exact})
            p.free()
            print("%-28s choice %-13s best %-13s regret %.3f (schedule alone %.3f)" % (
                name, SCHEDULES[choice], SCHEDULES[best], rows[-1]["regret"], rows[-1]["regret_schedule_only"]),
                file=sys.stderr, flush=True)
    finally:
        L.mgx_dispatch_set_model(None)
    tot_model = sum(r["t_auto_s"] for r in rows)
    tot_best = sum(min(r["times_s"]) for r in rows)
    tot_fixed = {SCHEDULES[k]: sum(min(r["times_s"][k], 1e9) for r in rows) for k in range(O)}
    # tuning at run time: the first call of each new shape times every
    # schedule once (warm-up + one run each) before running the best
    tot_tuning = sum(2 * sum(x for x in r["times_s"] if x < 1e29) + min(r["times_s"]) for r in rows)
    return {"mixes": rows, "max_regret": max(r["regret"] for r in rows),
            "mean_regret": float(np.mean([r["regret"] for r in rows])),
            "total_time_model_s": tot_model, "total_time_exhaustive_best_s": tot_best,
            "total_time_fixed_schedule_s": tot_fixed, "total_time_tuning_at_run_time_s": tot_tuning}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["gen", "train", "eval"])
    ap.add_argument("--configs", type=int, default=5000)
    ap.add_argument("--epochs", type=int, default=3000)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--out", default=DEFAULT_PARAMS)
    ap.add_argument("--params", default=DEFAULT_PARAMS)
    ap.add_argument("--data", default=None, help="save/load the timing data (npz; train: comma-separated list)")
    ap.add_argument("--json", default=None)
    a = ap.parse_args()
    mgx.init(0)
    if a.cmd == "gen":
        start = None
        if a.data and os.path.exists(a.data):
            z = np.load(a.data)
            start = (z["X"], z["T"], z["flags"], z["cfgs"])
            print("resuming after %d configs" % len(z["X"]), file=sys.stderr, flush=True)
        X, T, flags, cfgs = generate(a.configs, a.seed, save=a.data, start=start)
        np.savez(a.data, X=X, T=T, flags=flags, cfgs=np.array(cfgs, dtype=np.float64))
        print(json.dumps({"configs": int(X.shape[0]), "data": a.data}))
    elif a.cmd == "train":
        t0 = time.time()
        paths = a.data.split(",") if a.data else []
        if paths and all(os.path.exists(p) for p in paths):   # several timing files are concatenated
            zs = [np.load(p) for p in paths]
            X = np.concatenate([z["X"] for z in zs])
            T = np.concatenate([z["T"] for z in zs])
            flags = np.concatenate([z["flags"] for z in zs])
        else:
            X, T, flags, _ = generate(a.configs, a.seed)
            if a.data:
                np.savez(a.data, X=X, T=T, flags=flags)
        Y, mask = targets(T)
        n = X.shape[0]
        idx = np.random.default_rng(0).permutation(n)
        tr, va = idx[: int(0.9 * n)], idx[int(0.9 * n):]
        cm, mean, scale, loss = train_model(X[tr], Y[tr], mask[tr], epochs=a.epochs)
        pv = predict(cm, X[va])
        ch = np.argmin(np.where(mask[va] > 0, pv, 1e9), 1)   # infeasible schedules are masked
        tv = T[va]
        reg = tv[np.arange(len(va)), ch] / tv.min(1) - 1
        meta = {"configs": int(n), "train_loss": float(loss), "val_mean_regret": float(reg.mean()),
                "val_p90_regret": float(np.quantile(reg, 0.9)), "seconds": time.time() - t0}
        save_model(cm, mean, scale, a.out, meta)
        print(json.dumps(meta))
    else:
        cm = load_model(a.params)
        res = evaluate(cm)
        if a.json:
            with open(a.json, "w") as f:
                json.dump(res, f, indent=1)
        print(json.dumps({k: v for k, v in res.items() if k != "mixes"}))


if __name__ == "__main__":
    main()
