"""The learned dispatch against the exhaustive autotuner on held-out mixes
(shapes and entry bit-length histograms not in the training data): the
committed cost model drives MGX_ROUTE_AUTO, whose timed cost includes its
own scan and inference, and its mean regret t(auto)/t(best) - 1 over the
mixes must stay within 10%.  The test also reports what the network gains
over the default schedule and over tuning at run time, and checks that the
dispatched output equals the default schedule's bit for bit."""
import ctypes
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODEL_JSON = os.path.join(ROOT, "bachelor-thesis_amd", "costmodel_params.json")


def test_learned_dispatch_regret(gpu):
    import costmodel
    cm = costmodel.load_model(MODEL_JSON)
    try:
        res = costmodel.evaluate(cm, reps=5, check_exact=True)
    finally:
        costmodel.lib().mgx_costmodel_destroy(cm)
    print(json.dumps({k: v for k, v in res.items() if k != "mixes"}))
    for r in res["mixes"]:
        print("%-28s auto %-13s best %-13s regret %+.3f" % (r["mix"], r["choice"], r["best"], r["regret"]))
    assert all(r["bit_identical"] for r in res["mixes"])
    assert res["mean_regret"] <= 0.10, res["mean_regret"]
    # the network must not lose to always running the default schedule
    assert res["total_time_model_s"] <= 1.02 * res["total_time_fixed_schedule_s"]["crt2"]
