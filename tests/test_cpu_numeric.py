"""Numeric primitives: Philox4x32-10 and the portable log2.

Pins: the Random123 known-answer vectors for Philox4x32-10, rocRAND's
independent open-source implementation on random counters/keys, numpy's
float64 log2 (error bound), and bit equality between the engine's header and
the oracle's separate implementation (which is what makes host and GPU
streams identical)."""
import numpy as np

from conftest import ORACLE, pipe

KAT = [  # Random123 kat_vectors, philox4x32 10 rounds
    ("00000000 00000000 00000000 00000000 00000000 00000000", "6627e8d5 e169c58d bc57ac4c 9b00dbd8"),
    ("ffffffff ffffffff ffffffff ffffffff ffffffff ffffffff", "408f276d 41c83b0e a20bc7c6 6d5451fd"),
    ("243f6a88 85a308d3 13198a2e 03707344 a4093822 299f31d0", "d16cfe09 94fdcceb 5001e420 24126ea1"),
]


def test_philox_known_answers(native_bin):
    eng = native_bin("numeric_probe")
    lines = "".join(f"P {c}\n" for c, _ in KAT)
    assert pipe(eng, lines) == [r for _, r in KAT]
    assert pipe(ORACLE, lines, "--numeric-probe") == [r for _, r in KAT]


def test_philox_matches_rocrand(native_bin):
    eng = native_bin("numeric_probe")
    roc = native_bin("rocrand_probe", compiler="/opt/rocm/bin/hipcc")
    rng = np.random.default_rng(7)
    vals = rng.integers(0, 2**32, size=(500, 6), dtype=np.uint64)
    rows = [" ".join(f"{int(x):08x}" for x in r) for r in vals]
    a = pipe(eng, "".join(f"P {r}\n" for r in rows))
    b = pipe(roc, "".join(f"{r}\n" for r in rows))
    c = pipe(ORACLE, "".join(f"P {r}\n" for r in rows), "--numeric-probe")
    assert a == b == c


def _log2_inputs():
    rng = np.random.default_rng(11)
    xs = [np.float32(1.0), np.float32(2.0), np.float32(0.5), np.float32(np.sqrt(2)),
          np.nextafter(np.float32(np.sqrt(2)), np.float32(2)), np.float32(1e-40), np.float32(3.4e38),
          np.float32(1.0000001), np.float32(0.99999994)]
    xs += list(rng.uniform(0, 1, 20000).astype(np.float32))
    xs += list(np.exp(rng.uniform(-80, 80, 20000)).astype(np.float32))
    # every exponent, random mantissas, plus the 24-bit uniforms' grid
    xs += list((np.arange(1, 2**24, 997, dtype=np.float64) / 2**24).astype(np.float32))
    return np.array([x for x in xs if x > 0], dtype=np.float32)


def test_log2_accuracy_and_host_parity(native_bin):
    xs = _log2_inputs()
    bits = xs.view(np.uint32)
    text = "".join(f"L {b:08x}\n" for b in bits)
    eng = np.array([int(x, 16) for x in pipe(native_bin("numeric_probe"), text)], dtype=np.uint32).view(np.float32)
    orc = np.array([int(x, 16) for x in pipe(ORACLE, text, "--numeric-probe")], dtype=np.uint32).view(np.float32)
    assert np.array_equal(eng.view(np.uint32), orc.view(np.uint32))
    ref = np.log2(xs.astype(np.float64))
    err = np.abs(eng.astype(np.float64) - ref)
    # about 1 ulp of the result, and absolute 2^-24 near x = 1
    tol = np.maximum(np.abs(ref) * 2.5e-7, 1.2e-7 * np.abs(np.log2(np.abs(xs.astype(np.float64)) + 1e-300) + 1))
    bad = err > np.maximum(tol, 3e-8)
    assert not bad.any(), (xs[bad][:5], eng[bad][:5], ref[bad][:5])
    assert eng[0] == 0.0 and eng[1] == 1.0 and eng[2] == -1.0


def test_uniform_grid(native_bin):
    xs = [0, 1, 255, 256, 0x7fffffff, 0xffffff00, 0xffffffff]
    text = "".join(f"U {x:08x}\n" for x in xs)
    out = np.array([int(v, 16) for v in pipe(native_bin("numeric_probe"), text)], dtype=np.uint32).view(np.float32)
    expect = np.array([((x >> 8) + 1) / 2**24 for x in xs], dtype=np.float32)
    assert np.array_equal(out, expect)
    assert out.min() > 0 and out.max() == 1.0
