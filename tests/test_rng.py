"""RNG streams: oracle pinned by the reference's known sums; device streams
bit-exact against the oracle (ranmar: egs++/egs_rndm.cpp:303-398,
ranlux: src/ranlux.mortran)."""
import json
import os

import numpy as np
import pytest

from tests import oracle_lib

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "rng_known_sums.json")))


def _check_sums(vals, sums):
    cs = np.cumsum(vals)
    for k, v in sums.items():
        # the test programs print the sum with 8 decimals
        assert abs(cs[int(k) - 1] - v) <= 5.01e-9, (k, cs[int(k) - 1], v)


def test_oracle_ranmar_known_sums():
    g = GOLD["ranmar"]
    vals = oracle_lib.ranmar_fill(g["ixx"], g["jxx"], 1, 1000000)[0]
    _check_sums(vals, g["sums"])


@pytest.mark.parametrize("lux", [1, 4])
def test_oracle_ranlux_known_sums(lux):
    g = GOLD["ranlux"]
    vals = oracle_lib.ranlux_fill(lux, g["seed"], 1, 1000000)[0]
    _check_sums(vals, g["lux%d" % lux])


def test_oracle_ranmar_invalid_seeds_fall_back_to_default():
    # EGS_Ranmar::setState resets out-of-range seeds to 1802/9373
    a = oracle_lib.ranmar_fill(0, 0, 1, 100)
    b = oracle_lib.ranmar_fill(1802, 9373, 1, 100)
    assert np.array_equal(a, b)


def test_oracle_ranmar_values_are_24bit():
    v = oracle_lib.ranmar_fill(33, 97, 4, 5000)
    assert np.all(v >= 0) and np.all(v < 1)
    assert np.array_equal(v * 2**24, np.round(v * 2**24))


def test_oracle_ranlux_values_are_float_exact():
    # ranlux forms its outputs in real*4 (ranlux.mortran); small values carry
    # the low 24 bits of the next word but stay single-precision numbers
    v = oracle_lib.ranlux_fill(1, 7, 2, 20000)
    assert np.array_equal(v.astype(np.float32).astype(np.float64), v)
    small = v[v < 2**-12]
    assert small.size > 0
    assert np.any(small * 2**24 != np.round(small * 2**24))


def test_stream_seed_mapping():
    # stream s of seeds (ixx, jxx) is createRNG(input, s) = (ixx, jxx+s)
    a = oracle_lib.ranmar_fill(1802, 9373, 3, 50)
    for s in range(3):
        b = oracle_lib.ranmar_fill(1802, 9373 + s, 1, 50)[0]
        assert np.array_equal(a[s], b)
    # beyond jxx+s = 30080 every stream still differs
    many = oracle_lib.ranmar_fill(1802, 30000, 200, 4)
    assert len({tuple(r) for r in many}) == 200


@pytest.mark.gpu
def test_device_ranmar_bit_exact(engine):
    for (ixx, jxx, ns, n) in [(1802, 9373, 1, 1000000), (33, 97, 257, 3000),
                              (1802, 30000, 300, 200)]:
        d = engine.ranmar_fill(ixx, jxx, ns, n)
        o = oracle_lib.ranmar_fill(ixx, jxx, ns, n)
        assert np.array_equal(d, o), (ixx, jxx, ns, n)


@pytest.mark.gpu
def test_device_ranmar_known_sums(engine):
    g = GOLD["ranmar"]
    _check_sums(engine.ranmar_fill(g["ixx"], g["jxx"], 1, 1000000)[0], g["sums"])


@pytest.mark.gpu
@pytest.mark.parametrize("lux", [0, 1, 2, 3, 4])
def test_device_ranlux_bit_exact(engine, lux):
    d = engine.ranlux_fill(lux, 1, 65, 2000)
    o = oracle_lib.ranlux_fill(lux, 1, 65, 2000)
    assert np.array_equal(d, o)
    if lux in (1, 4):
        g = GOLD["ranlux"]
        _check_sums(engine.ranlux_fill(lux, 1, 1, 1000000)[0], g["lux%d" % lux])
