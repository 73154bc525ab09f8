"""Writes the known-answer fixtures under tests/golden/ from Python's exact
integers.  The reference ships no tests or vectors, so these pin the exact
product, the text format and the CRT constants by mathematics alone.

Run in this container: python tests/golden/gen_fixtures.py
Nothing from the reference is imported or run.
"""
import json
import os
import random

import sympy

HERE = os.path.dirname(os.path.abspath(__file__))


def hexs(v):
    # mat-gpu/mpz.cu:85-104 format: lower-case hex, '-' prefix, "0"
    return ("-" if v < 0 else "") + format(abs(v), "x")


def rand_entry(rng, bits, mode):
    if mode == "zero":
        return 0
    if mode == "one":
        return rng.choice([1, -1])
    if mode == "ones":
        return rng.choice([1, -1]) * ((1 << bits) - 1)
    v = rng.getrandbits(bits) if bits else 0
    return -v if rng.getrandbits(1) else v


def matmul(A, B):
    K, M, N = len(A), len(B), len(B[0])
    return [[sum(A[y][k] * B[k][x] for k in range(M)) for x in range(N)] for y in range(K)]


def matmul_cases():
    rng = random.Random(20181203)
    cases = []
    shapes = [(1, 1, 1), (1, 4, 1), (2, 3, 4), (3, 3, 3), (4, 4, 4), (5, 7, 3), (8, 8, 8), (16, 16, 16), (4, 17, 2)]
    for bits in [1, 31, 32, 63, 64, 65, 127, 128, 129, 255, 1024, 2048]:
        for (K, M, N) in shapes:
            if bits >= 1024 and K * M * N > 512:
                continue
            for mode in ["rand", "ones", "mixed"]:
                def e():
                    m = mode if mode != "mixed" else rng.choice(["rand", "zero", "one", "ones", "rand"])
                    return rand_entry(rng, bits, m)
                A = [[e() for _ in range(M)] for _ in range(K)]
                B = [[e() for _ in range(N)] for _ in range(M)]
                cases.append({"K": K, "M": M, "N": N, "bits": bits, "mode": mode,
                              "A": [[hexs(v) for v in r] for r in A], "B": [[hexs(v) for v in r] for r in B],
                              "C": [[hexs(v) for v in r] for r in matmul(A, B)]})
    # cancellation: C = A*B with A*B having exact zeros and a one-limb result
    A = [[(1 << 200) + 1, -(1 << 200)], [7, 7]]
    B = [[(1 << 64), 1], [(1 << 64) + (1 << 64), 1]]
    cases.append({"K": 2, "M": 2, "N": 2, "bits": 201, "mode": "cancel",
                  "A": [[hexs(v) for v in r] for r in A], "B": [[hexs(v) for v in r] for r in B],
                  "C": [[hexs(v) for v in r] for r in matmul(A, B)]})
    return cases


def crt_vectors():
    # the largest primes below 2^31, from sympy (an independent primality source)
    ps = []
    p = 1 << 31
    while len(ps) < 300:
        p = sympy.prevprime(p)
        ps.append(p)
    # smallest prefix with product > 2^bits, for a few bit bounds
    counts = {}
    for bits in [1, 31, 62, 100, 1000, 2048 * 2 + 10 + 2, 4096 * 2 + 8 + 2]:
        m, n = 1, 0
        while m <= (1 << bits):
            m *= ps[n]
            n += 1
        counts[str(bits)] = n
    rng = random.Random(7)
    vals = [0, 1, -1, (1 << 2048) - 1, -((1 << 2048) - 1), ps[0], -ps[0], ps[0] - 1, -(ps[0] - 1)]
    vals += [rand_entry(rng, rng.choice([5, 64, 65, 700, 2048]), "rand") for _ in range(40)]
    residues = [{"x": hexs(v), "r": [v % q for q in ps[:8]]} for v in vals]
    # balanced radix-256 digits of centred residues
    digs = []
    for r in [0, 1, -1, 127, 128, -128, -129, 255, 256, (ps[0] - 1) // 2, -((ps[0] - 1) // 2),
              2139062143, -2155905152] + [rng.randrange(-(1 << 30), 1 << 30) for _ in range(40)]:
        d, v = [], r
        for _ in range(4):
            t = ((v + 128) & 255) - 128
            d.append(t)
            v = (v - t) // 256
        assert v == 0 and sum(t * 256 ** i for i, t in enumerate(d)) == r
        digs.append({"r": r, "d": d})
    # two-digit residues: the largest primes below 65280 and their prefix counts
    ps16, p = [], 65280
    while len(ps16) < 700:
        p = sympy.prevprime(p)
        ps16.append(p)
    counts16 = {}
    for bits in [1, 15, 100, 1000, 2048 * 2 + 10 + 2, 4096 * 2 + 8 + 2]:
        m, n = 1, 0
        while m <= (1 << bits):
            m *= ps16[n]
            n += 1
        counts16[str(bits)] = n
    return {"primes": ps, "prime_counts": counts, "residues": residues, "slices": digs,
            "primes16": ps16, "prime_counts16": counts16}


def text_case():
    # mat-cpu -g format (mat-cpu/main.c:249-276) and -p output (mat-cpu/mat.c:344-353)
    rng = random.Random(3)
    K, M, N = 3, 2, 4
    A = [[rand_entry(rng, 130, "rand") for _ in range(M)] for _ in range(K)]
    B = [[rand_entry(rng, 130, "rand") for _ in range(N)] for _ in range(M)]
    A[0][0] = 0
    gen = "%d %d %d\n" % (K, M, N)
    for mat in (A, B):
        for r in mat:
            gen += " ".join(hexs(v) for v in r) + "\n"
        gen += "\n"
    C = matmul(A, B)
    out = ""
    for r in C:
        out += " ".join(hexs(v) for v in r) + "\n"
    out += "\n"
    return gen, out


def text_cases():
    """More -g/-p pairs: wide entries (3000 bits, with all-ones extremes and
    zeros), a zero row and a zero column, a 1 x 1 result, and an empty inner
    dimension (C is all zeros)."""
    rng = random.Random(11)
    cases = []

    def render(K, M, N, A, B):
        gen = "%d %d %d\n" % (K, M, N)
        for mat in (A, B):
            for r in mat:
                gen += " ".join(hexs(v) for v in r) + "\n"
            gen += "\n"
        C = matmul(A, B) if M else [[0] * N for _ in range(K)]
        out = "".join(" ".join(hexs(v) for v in r) + "\n" for r in C) + "\n"
        return {"name": "%dx%dx%d" % (K, M, N), "input": gen, "product": out}

    K, M, N, b = 3, 5, 2, 3000
    A = [[rand_entry(rng, b, "rand") for _ in range(M)] for _ in range(K)]
    B = [[rand_entry(rng, b, "rand") for _ in range(N)] for _ in range(M)]
    A[0][1], A[2][4], B[3][0] = (1 << b) - 1, -((1 << b) - 1), 0
    cases.append(render(K, M, N, A, B))
    K, M, N = 4, 3, 3
    A = [[rand_entry(rng, 200, "rand") for _ in range(M)] for _ in range(K)]
    B = [[rand_entry(rng, 200, "rand") for _ in range(N)] for _ in range(M)]
    A[2] = [0] * M
    for r in B:
        r[1] = 0
    cases.append(render(K, M, N, A, B))
    A = [[rand_entry(rng, 700, "rand") for _ in range(7)]]
    B = [[rand_entry(rng, 700, "rand")] for _ in range(7)]
    cases.append(render(1, 7, 1, A, B))
    cases.append(render(3, 0, 2, [[] for _ in range(3)], []))
    return cases


def main():
    with open(os.path.join(HERE, "kat_matmul.json"), "w") as f:
        json.dump(matmul_cases(), f)
    with open(os.path.join(HERE, "kat_crt.json"), "w") as f:
        json.dump(crt_vectors(), f)
    with open(os.path.join(HERE, "kat_text_cases.json"), "w") as f:
        json.dump(text_cases(), f)
    gen, out = text_case()
    with open(os.path.join(HERE, "kat_text_input.txt"), "w") as f:
        f.write(gen)
    with open(os.path.join(HERE, "kat_text_product.txt"), "w") as f:
        f.write(out)


if __name__ == "__main__":
    main()
