"""The mat-gpu command line: mat-cpu's -g text format in, -p text out,
byte-identical to the product computed with Python integers."""
import os
import subprocess

import pytest

from helpers import GOLDEN, matmul, parse_hex

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "bachelor-thesis_amd", "bin", "mat-gpu")


@pytest.mark.parametrize("method", ["1", "35", "8"])
def test_text_roundtrip(method):
    with open(os.path.join(GOLDEN, "kat_text_input.txt")) as f:
        text = f.read()
    with open(os.path.join(GOLDEN, "kat_text_product.txt")) as f:
        want = f.read()
    r = subprocess.run([BIN, "-i", "-p", "-m", method], input=text, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout == want
    assert "Time taken:" in r.stderr


def _text_cases():
    import json
    with open(os.path.join(GOLDEN, "kat_text_cases.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("method", ["1", "35", "8", "0"])
@pytest.mark.parametrize("case", _text_cases(), ids=lambda c: c["name"])
def test_text_cases(case, method):
    """Wide (3000-bit) entries with all-ones extremes, a zero row and column, a
    1 x 1 result and an empty inner dimension, through every route."""
    r = subprocess.run([BIN, "-i", "-p", "-m", method], input=case["input"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout == case["product"]


def test_generated_product_and_json():
    r = subprocess.run([BIN, "-n", "3,4,5", "-b", "300", "-m", "8", "-s", "9", "-p", "-j"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().split("\n")
    rows = [[parse_hex(t) for t in ln.split()] for ln in lines[:3]]
    assert len(rows) == 3 and all(len(x) == 5 for x in rows)
    assert lines[-1].startswith("{") and '"method": 8' in lines[-1]
    # same seed, plain route: identical text
    r2 = subprocess.run([BIN, "-n", "3,4,5", "-b", "300", "-m", "2", "-s", "9", "-p"], capture_output=True, text=True,
                        timeout=120)
    assert r2.stdout.strip().split("\n")[:3] == lines[:3]


def test_cli_learned_dispatch():
    # -m 0: the route the committed cost model picks, same product as -m 8
    args = ["-n", "5,7,6", "-b", "700", "-s", "4", "-p"]
    r0 = subprocess.run([BIN] + args + ["-m", "0"], capture_output=True, text=True, timeout=120)
    r8 = subprocess.run([BIN] + args + ["-m", "8"], capture_output=True, text=True, timeout=120)
    assert r0.returncode == 0 and r8.returncode == 0, r0.stderr + r8.stderr
    assert "No cost model" not in r0.stderr
    assert r0.stdout == r8.stdout


def test_cli_errors():
    r = subprocess.run([BIN, "-n", "2,2,2"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "No bit count" in r.stderr
    r = subprocess.run([BIN, "-i", "-n", "2,2,2"], input="", capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    r = subprocess.run([BIN, "-i", "-p"], input="1 1 1\nzz\n\n1\n\n", capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
