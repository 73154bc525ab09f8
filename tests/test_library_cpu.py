"""The C-ABI library loads and exports every declared symbol; its host-only
utilities agree with independent restatements.  No GPU calls."""
import ctypes
import glob
import os
import re

import numpy as np

from helpers import entry_int, to_entries

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(mgx_[a-z0-9_]+)\s*\(", text))
    return names


def test_exports_every_declared_symbol():
    import mgx
    L = mgx.lib()
    names = declared_symbols()
    assert len(names) > 40
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_binding_declares_every_symbol():
    """Every function the headers declare has its argument types set in the
    Python binding: ctypes would otherwise pass 64-bit pointers and sizes as
    32-bit C ints, truncating device addresses."""
    import mgx
    L = mgx.lib()
    undeclared = [n for n in sorted(declared_symbols()) if getattr(L, n).argtypes is None]
    assert not undeclared, undeclared


def _xoshiro_entries(rows, cols, cap, nbits, seed):
    M64 = (1 << 64) - 1

    def splitmix(x):
        x = (x + 0x9E3779B97F4A7C15) & M64
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        return x, z ^ (z >> 31)

    s = []
    x = seed
    for _ in range(4):
        x, z = splitmix(x)
        s.append(z)

    def rotl(v, k):
        return ((v << k) | (v >> (64 - k))) & M64

    def nxt():
        r = (rotl((s[1] * 5) & M64, 7) * 9) & M64
        t = (s[1] << 17) & M64
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]
        s[2] ^= t; s[3] = rotl(s[3], 45)
        return r

    nl = (nbits + 63) // 64
    out = []
    for _ in range(rows * cols):
        limbs = [nxt() for _ in range(nl)]
        if nbits % 64:
            limbs[-1] &= (1 << (nbits % 64)) - 1
        v = sum(l << (64 * i) for i, l in enumerate(limbs))
        out.append(-v if nxt() & 1 else v)
    return out


def test_random_matrix_stream():
    import mgx
    for nbits in (1, 64, 100, 2048):
        A = mgx.random_matrix(3, 5, 40, nbits, 77)
        got = [entry_int(A[y, x]) for y in range(3) for x in range(5)]
        assert got == _xoshiro_entries(3, 5, 40, nbits, 77)


def test_hex_roundtrip_matches_reference_format():
    import mgx
    L = mgx.lib()
    cap = 5
    for v, s in [(0, "0"), (-1, "-1"), (0xabc, "abc"), (-(1 << 200) - 5, "-" + format((1 << 200) + 5, "x"))]:
        e = to_entries([[v]], cap)[0, 0]
        buf = ctypes.create_string_buffer(300)
        assert L.mgx_entry_to_hex(buf, 300, e.ctypes.data, cap) == len(s)
        assert buf.value.decode() == s
        back = np.zeros(cap + 1, dtype=np.uint64)
        assert L.mgx_entry_from_hex(back.ctypes.data, cap, s.encode()) == 0
        assert entry_int(back) == v
    bad = np.zeros(cap + 1, dtype=np.uint64)
    assert L.mgx_entry_from_hex(bad.ctypes.data, cap, b"xyz") == 7


def test_hex_against_python_format():
    """Python's own hexadecimal formatting is the format's definition: random
    magnitudes, zero limbs inside, a zero top limb left in the size (not
    normalized), and a buffer one byte too small."""
    import random
    import mgx
    L = mgx.lib()
    cap = 8
    rng = random.Random(4)
    vals = [rng.getrandbits(rng.randint(1, 64 * cap)) * rng.choice([1, -1]) for _ in range(200)]
    vals += [1 << 128, -(1 << 448), (1 << 512) - 1, 1 << 63, -(1 << 64)]
    for v in vals:
        e = to_entries([[v]], cap)[0, 0]
        want = ("-" if v < 0 else "") + format(abs(v), "x")
        buf = ctypes.create_string_buffer(200)
        assert L.mgx_entry_to_hex(buf, 200, e.ctypes.data, cap) == len(want)
        assert buf.value.decode() == want
        assert L.mgx_entry_to_hex(buf, len(want), e.ctypes.data, cap) == -1   # no room for the NUL
    e = to_entries([[-0x1234]], cap)[0, 0]
    e[cap] = np.uint64((-3) & 0xFFFFFFFFFFFFFFFF)   # size -3, limbs 1 and 2 zero
    buf = ctypes.create_string_buffer(64)
    assert L.mgx_entry_to_hex(buf, 64, e.ctypes.data, cap) == 5 and buf.value == b"-1234"
    z = np.zeros(cap + 1, dtype=np.uint64)
    z[cap] = np.uint64((-2) & 0xFFFFFFFFFFFFFFFF)   # a negative zero prints as 0
    assert L.mgx_entry_to_hex(buf, 64, z.ctypes.data, cap) == 1 and buf.value == b"0"


def test_committed_cost_model_files_agree():
    """The text model (read by mgx_costmodel_load) and the JSON copy hold the
    same 3-layer, 64-hidden network."""
    import json
    here = os.path.join(ROOT, "bachelor-thesis_amd")
    with open(os.path.join(here, "costmodel_params.json")) as f:
        d = json.load(f)
    with open(os.path.join(here, "costmodel_params.txt")) as f:
        head = f.readline().split()
        rows = [np.array(line.split(), dtype=np.float64) for line in f]
    assert head[:2] == ["mgx-costmodel", "1"]
    D, H, O = (int(x) for x in head[2:5])
    assert (D, H, O) == (d["D"], d["H"], d["O"]) == (13, 64, 9)   # one output per schedule
    assert len(rows) == 3 and rows[0].size == D and rows[1].size == D
    assert rows[2].size == H * D + H + H * H + H + O * H + O
    assert np.allclose(rows[2], d["params"], rtol=1e-6) and np.allclose(rows[0], d["mean"], rtol=1e-6)
