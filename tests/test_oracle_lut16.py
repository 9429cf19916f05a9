"""The oracle's FP16 filter table (the engine's PUCKGPU_LUT_FP16 form; the
reference's table is FP32, quantization.cpp:260-270).

What pins it:
* po_half_round is numpy's float16 conversion (IEEE binary16, round to
  nearest even) on random values of every range, on exact midpoints between
  neighbouring binary16 values, at the overflow edge and in the subnormal
  range;
* po_lut16_exp picks the largest power of two 2^e with (|q| + cbn)^2 * 2^e <=
  65504, checked against that inequality in float64, with its clamps;
* po_lut16 keeps every entry within 2^-11 relative of the FP32 entry and
  leaves entries that are exact in 11 significant bits unchanged;
* a search over an index whose table entries are all exact in FP16 returns
  the FP32 search's results bit for bit (walk and budget scan), and on
  ordinary data the FP16 search's filter distances move by no more than the
  rounding bound -- a mistake such as a scale not undone fails both.
"""
import numpy as np
import pytest

from oracle import oracle as O
from puck_amd import datasets as DS
import cpu_index as CI


def test_half_round_is_binary16_nearest_even():
    rng = np.random.default_rng(3)
    v = np.concatenate([
        rng.random(4000, dtype=np.float32) * 65000,
        rng.random(4000, dtype=np.float32) * 2e-4,  # binary16 subnormals below 6.1e-5
        (np.float32(2.0) ** rng.integers(-26, 16, 4000).astype(np.float32)) * (1 + rng.random(4000, dtype=np.float32)),
    ]).astype(np.float32)
    # exact midpoints between neighbouring binary16 values (ties to even)
    h = rng.integers(1, 0x7bff, 4000).astype(np.uint16).view(np.float16)
    nxt = np.nextafter(h, np.float16(np.inf))
    mid = ((h.astype(np.float64) + nxt.astype(np.float64)) / 2).astype(np.float32)
    assert np.array_equal(mid.astype(np.float64), (h.astype(np.float64) + nxt.astype(np.float64)) / 2)
    v = np.concatenate([v, mid, -v[:500], np.float32([0.0, 65504.0, 65519.996, 65520.0, 1e30, 5.96e-8, 2.98e-8])])
    got = np.array([O.half_round(x) for x in v], np.float32)
    with np.errstate(over="ignore"):
        want = v.astype(np.float16).astype(np.float32)
    assert np.array_equal(got, want)
    # a midpoint goes to the neighbour with an even significand
    even = np.where((h.view(np.uint16) & 1) == 0, h, nxt).astype(np.float32)
    assert np.array_equal(np.array([O.half_round(x) for x in mid], np.float32), even)


def test_lut16_exp_is_largest_safe_power_of_two():
    rng = np.random.default_rng(4)
    for _ in range(3000):
        qn = np.float32(rng.random() * 10.0 ** rng.integers(-4, 9))
        cbn = np.float32(rng.random() * 10.0 ** rng.integers(-3, 4))
        e = O.lut16_exp(qn, cbn)
        B = (np.sqrt(np.float64(qn)) + np.float64(cbn)) ** 2
        if -30 < e < 30:
            assert B * 2.0 ** e <= 65504 * (1 + 1e-6), (qn, cbn, e)
            assert B * 2.0 ** (e + 1) > 65504 * (1 - 1e-6), (qn, cbn, e)
    assert O.lut16_exp(0.0, 0.0) == 30  # an all-zero table
    assert O.lut16_exp(1e30, 1e15) == -30
    assert O.lut16_exp(1e-30, 1e-20) == 30


@pytest.fixture(scope="module")
def idx_cos():
    x = DS.unit_norm_mixture(6000, 32, 64, seed=80, means_seed=81)
    q = DS.unit_norm_mixture(24, 32, 64, seed=82, means_seed=81)
    return CI.build_index(x, 8, 8, 16, whether_norm=True, train_n=4000), q


def test_lut16_entries_within_rounding(idx_cos):
    ix, q = idx_cos
    cb = ix["codebook"].reshape(-1)
    lsq = ix["raw"].shape[1] // ix["nsq"]
    cbn = float(np.sqrt((cb.reshape(-1, lsq).astype(np.float64) ** 2).sum(1).max()))
    for v in q[:8]:
        qv = O.normalize(v)
        qn = O.dot(qv, qv)
        t = O.lut(qv, cb, ix["nsq"], lsq)
        t16 = O.lut16(t, qn, np.float32(cbn))
        assert np.all(np.abs(t16.astype(np.float64) - t) <= 2.0 ** -11 * np.abs(t) + 2.0 ** -24 * 2.0 ** -O.lut16_exp(
            qn, np.float32(cbn)))
        assert np.any(t16 != t)
    # entries that fit 11 significant bits stay as they are (entries up to 32,
    # within the bound (|q| + cbn)^2 = 36)
    exact = (np.arange(1, 2049, dtype=np.float32) * np.float32(2.0 ** -6))
    assert np.array_equal(O.lut16(exact, np.float32(1.0), np.float32(5.0)), exact)


@pytest.fixture(scope="module")
def idx_exact():
    # integer data in [0, 3] and a codebook of quarter-integers in [-4.5, 4.5]:
    # every table entry (two squared differences below 7.5^2) is a multiple of
    # 1/16 below 2^7, exact in 11 significant bits at any power-of-two scale
    rng = np.random.default_rng(5)
    means = rng.integers(0, 4, (40, 16)).astype(np.float32)
    x = np.clip(means[rng.integers(0, 40, 5000)] + rng.integers(-1, 2, (5000, 16)), 0, 3).astype(np.float32)
    q = np.clip(means[rng.integers(0, 40, 24)] + rng.integers(-1, 2, (24, 16)), 0, 3).astype(np.float32)
    ix = CI.build_index(x, 6, 6, 8, whether_norm=False, train_n=4000)
    cb = np.clip(np.round(ix["codebook"] * 4) / 4, -4.5, 4.5)
    assert np.abs(cb).max() <= 4.5
    ix["codebook"] = cb.astype(np.float32)
    ix["records"] = O.encode_filter(ix["coarse"], ix["fine"], ix["codebook"], 8, x, ix["cell_assign"])
    return ix, q


def test_lut16_search_exact_table_equals_fp32(idx_exact):
    ix, q = idx_exact
    a, b = CI.oracle_index(ix), CI.oracle_index(ix)
    b.set_lut16(True)
    for scc, ftk, nf in ((6, 5000, 0), (2, 40, 0), (3, 60, 2), (6, 200, 1)):
        for v in q:
            ra = a.search_one(v, 10, scc, ftk, fine_budget=nf)
            rb = b.search_one(v, 10, scc, ftk, fine_budget=nf)
            assert ra["n"] == rb["n"] and ra["filt_n"] == rb["filt_n"]
            assert np.array_equal(ra["ids"], rb["ids"]) and np.array_equal(ra["dist"].view(np.uint32),
                                                                           rb["dist"].view(np.uint32))
            assert np.array_equal(ra["filt_val"].view(np.uint32), rb["filt_val"].view(np.uint32))
            assert np.array_equal(ra["filt_mem"], rb["filt_mem"])


def test_lut16_filter_distances_within_bound(idx_cos):
    ix, q = idx_cos
    N = ix["raw"].shape[0]
    lsq = ix["raw"].shape[1] // ix["nsq"]
    a, b = CI.oracle_index(ix), CI.oracle_index(ix)
    b.set_lut16(True)
    moved = 0
    for v in q:
        # every point of every cell: the filter set is all N points
        ra = a.search_one(v, 10, 8, N, fine_budget=0)
        rb = b.search_one(v, 10, 8, N, fine_budget=0)
        assert ra["filt_n"] == rb["filt_n"] == N
        fa = dict(zip(ra["filt_mem"].tolist(), ra["filt_val"].astype(np.float64)))
        fb = dict(zip(rb["filt_mem"].tolist(), rb["filt_val"].astype(np.float64)))
        qv = O.normalize(v)
        tmax = float(O.lut(qv, ix["codebook"].reshape(-1), ix["nsq"], lsq).max())
        bound = ix["nsq"] * 2.0 ** -11 * tmax
        d = np.array([abs(fb[m] - fa[m]) for m in fa])
        assert d.max() <= bound + 1e-6, (d.max(), bound)
        moved += int((d > 0).sum())
        # the returned distances are exact L2 in both forms
        assert np.all(rb["dist"][:rb["n"]] >= 0)
    assert moved > 0
