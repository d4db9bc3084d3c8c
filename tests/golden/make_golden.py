"""Writes the golden fixtures in tests/golden/ from first principles.

Nothing here calls the oracle or the library: inputs come from a pure-Python
SplitMix64, expected values from exact rational arithmetic (fractions) rounded
once, and BF16 cases from the IEEE round-to-nearest-even definition worked by
hand.  The fixtures are therefore an independent check of oracle/oracle.c and
of the GPU kernels.

The reference (szrlee/sgemm.cu) ships no golden vectors: test.cu compares
against a time-seeded run of its naive kernel and stores nothing.  What these
fixtures pin is the semantics the reference defines:
  * C := alpha*A*B + beta*C (src/sgemm.cuh:19-33; column-major here),
  * kernel_basic's FP32 operation order (src/kernels/basic.cuh:15-25):
    acc = fma(a, b, acc) over p in order, acc *= alpha, acc = fma(c, beta, acc).

Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import struct
from fractions import Fraction

HERE = os.path.dirname(os.path.abspath(__file__))
MASK = (1 << 64) - 1
GAMMA = 0x9E3779B97F4A7C15


def splitmix64(state: int):
    while True:
        state = (state + GAMMA) & MASK
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        yield z ^ (z >> 31)


def take(seed: int, n: int) -> list[int]:
    g = splitmix64(seed)
    return [next(g) for _ in range(n)]


def uniform(seed: int, n: int) -> list[float]:
    # 2u - 1 with u = (z >> 40) / 2^24: exact in binary32.
    return [2.0 * ((z >> 40) / 16777216.0) - 1.0 for z in take(seed, n)]


def ints(seed: int, n: int, lo: int, hi: int) -> list[float]:
    span = hi - lo + 1
    return [float(lo + ((z >> 32) % span)) for z in take(seed, n)]


def f32(x: float) -> float:
    return struct.unpack("<f", struct.pack("<f", x))[0]


def round_f32(q: Fraction) -> float:
    """Correctly rounded (ties to even) binary32 value of the rational q."""
    if q == 0:
        return 0.0
    sign = -1.0 if q < 0 else 1.0
    a = abs(q)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if Fraction(2) ** e > a:
        e -= 1
    while Fraction(2) ** (e + 1) <= a:
        e += 1
    qexp = max(e, -126) - 23
    scaled = a / (Fraction(2) ** qexp)
    n = scaled.numerator // scaled.denominator
    r = scaled - n
    if r > Fraction(1, 2) or (r == Fraction(1, 2) and n % 2 == 1):
        n += 1
    val = Fraction(n) * (Fraction(2) ** qexp)
    if val >= Fraction(2) ** 128:
        return sign * float("inf")
    return sign * float(val)


def colmajor(vals, rows, cols, ld):
    """Embed a rows x cols list (column-major order) into an ld-strided buffer;
    padding is filled with a marker that must never matter."""
    buf = [777.0] * (ld * (cols - 1) + rows if rows and cols else 0)
    for j in range(cols):
        for i in range(rows):
            buf[i + j * ld] = vals[i + j * rows]
    return buf


def exact_gemm(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    out = []
    for j in range(n):
        for i in range(m):
            s = sum((Fraction(A[i + p * lda]) * Fraction(B[p + j * ldb]) for p in range(k)),
                    Fraction(0))
            v = Fraction(alpha) * s
            if beta != 0.0:
                v += Fraction(beta) * Fraction(C[i + j * ldc])
            out.append(v)
    return out  # column-major, ld = m


def basic_f32_colmajor(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    """kernel_basic's operation order, column-major operands, exact rounding
    of every step (each fma rounds once)."""
    out = []
    for j in range(n):
        for i in range(m):
            acc = 0.0
            for p in range(k):
                acc = round_f32(Fraction(A[i + p * lda]) * Fraction(B[p + j * ldb]) + Fraction(acc))
            acc = round_f32(Fraction(acc) * Fraction(f32(alpha)))
            acc = round_f32(Fraction(C[i + j * ldc]) * Fraction(f32(beta)) + Fraction(acc))
            out.append(acc)
    return out


def make_case(name, m, n, k, lda, ldb, ldc, alpha, beta, kind, seed):
    if kind == "int":
        gen = lambda s, cnt: ints(s, cnt, -4, 4)
    else:
        gen = uniform
    A = colmajor(gen(seed, m * k), m, k, lda)
    B = colmajor(gen(seed + 1, k * n), k, n, ldb)
    C = colmajor(gen(seed + 2, m * n), m, n, ldc)
    exact = exact_gemm(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)
    case = {
        "name": name, "m": m, "n": n, "k": k, "lda": lda, "ldb": ldb, "ldc": ldc,
        "alpha": alpha, "beta": beta, "kind": kind, "seed": seed,
        "A": A, "B": B, "C0": C,
        # exact result rounded once to double / to float
        "expected_f64": [float(v) for v in exact],
        "expected_f32_rounded": [round_f32(v) for v in exact],
        "expected_basic_f32": basic_f32_colmajor(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc),
    }
    return case


def bf16_cases():
    """(float32 bits, expected bf16 bits) pairs worked from the RNE rule:
    keep the top 16 bits, add 1 if the dropped 16 bits exceed 0x8000, or equal
    0x8000 with the kept LSB odd."""
    return [
        (0x3F800000, 0x3F80, "1.0 exact"),
        (0x3F808000, 0x3F80, "1+2^-8: tie, kept LSB even -> down"),
        (0x3F818000, 0x3F82, "1+3*2^-8: tie, kept LSB odd -> up"),
        (0x3F808001, 0x3F81, "just above a tie -> up"),
        (0x3F807FFF, 0x3F80, "just below a tie -> down"),
        (0xBF818000, 0xBF82, "negative tie, odd -> away from zero in magnitude"),
        (0x00000000, 0x0000, "+0"),
        (0x80000000, 0x8000, "-0"),
        (0x00010000, 0x0001, "smallest bf16 subnormal, exact"),
        (0x00008000, 0x0000, "half the smallest bf16 subnormal: tie to even 0"),
        (0x00018000, 0x0002, "1.5 bf16 subnormal ulps: tie to even 2"),
        (0x7F7F0000, 0x7F7F, "max finite bf16, exact"),
        (0x7F7FFFFF, 0x7F80, "FLT_MAX rounds up to +Inf"),
        (0x7F800000, 0x7F80, "+Inf"),
        (0xFF800000, 0xFF80, "-Inf"),
        (0x7FC00000, 0x7FC0, "quiet NaN stays NaN"),
        (0x40490FDB, 0x4049, "pi: dropped 0x0FDB < 0x8000 -> down"),
        (0x3EAAAAAB, 0x3EAB, "1/3: dropped 0xAAAB > 0x8000 -> up"),
    ]


def main():
    cases = [
        make_case("int_1x1x1", 1, 1, 1, 1, 1, 1, 1.5, 0.5, "int", 11),
        make_case("int_3x5x7_ld", 3, 5, 7, 5, 9, 4, 1.5, 0.5, "int", 21),
        make_case("int_17x13x9", 17, 13, 9, 17, 9, 17, 1.5, 0.5, "int", 31),
        make_case("int_16x16x33_beta0", 16, 16, 33, 19, 40, 16, 1.0, 0.0, "int", 41),
        make_case("int_20x6x5_neg", 20, 6, 5, 20, 5, 23, -2.0, 0.25, "int", 51),
        make_case("uni_1x1x1", 1, 1, 1, 1, 1, 1, 1.0, 0.0, "uniform", 61),
        make_case("uni_5x3x4", 5, 3, 4, 5, 4, 5, 1.5, 0.5, "uniform", 71),
        make_case("uni_9x7x12_ld", 9, 7, 12, 11, 13, 10, 1.5, 0.5, "uniform", 81),
        make_case("uni_12x12x12", 12, 12, 12, 12, 12, 12, 1.0, 0.0, "uniform", 91),
        make_case("uni_7x10x31", 7, 10, 31, 8, 31, 7, -0.75, 1.25, "uniform", 101),
    ]
    with open(os.path.join(HERE, "gemm_cases.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py", "cases": cases}, f)
    streams = {str(seed): [str(v) for v in take(seed, 8)] for seed in (0, 1, 1234567, 2**63 + 5)}
    with open(os.path.join(HERE, "splitmix64.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py", "streams": streams}, f, indent=1)
    with open(os.path.join(HERE, "bf16_rne.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py",
                   "cases": [{"f32_bits": a, "bf16_bits": b, "what": w}
                             for a, b, w in bf16_cases()]}, f, indent=1)


if __name__ == "__main__":
    main()
