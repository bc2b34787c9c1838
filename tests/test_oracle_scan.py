"""Pins the oracle's scan/compaction: the reference's all-ones known answer
(test/algorithm/PrefixSumIntegerTest.java:30-36) and numpy on random data."""
import numpy as np
import pytest

from oracle import orc_py as orc
from tests.conftest import golden


def test_prefix_sum_all_ones_known_answer():
    k = golden("prefix_sum_all_ones.json")
    out, total = orc.scan_excl(np.array(k["input"], np.int32))
    np.testing.assert_array_equal(out, k["exclusive_scan"])
    assert total == k["total"]


@pytest.mark.parametrize("n", [1, 2, 127, 128, 129, 10000])
def test_scan_matches_numpy(n):
    x = np.random.default_rng(n).integers(-(1 << 31), (1 << 31) - 1, size=n, dtype=np.int64).astype(np.int32)
    out, total = orc.scan_excl(x)
    ref = (np.concatenate([[0], np.cumsum(x.astype(np.int64))[:-1]]) & 0xFFFFFFFF).astype(np.uint32).view(np.int32)
    np.testing.assert_array_equal(out, ref)
    v = int(np.sum(x.astype(np.int64))) & 0xFFFFFFFF
    assert total == (v - (1 << 32) if v >= (1 << 31) else v)


def test_empty_scan():
    out, total = orc.scan_excl(np.zeros(0, np.int32))
    assert out.size == 0 and total == 0


@pytest.mark.parametrize("p", [0.0, 0.01, 0.5, 1.0])
def test_compaction_is_stable_flatnonzero(p):
    pred = (np.random.default_rng(7).random(5000) < p).astype(np.int32)
    idx, cnt = orc.compact_stable(pred)
    np.testing.assert_array_equal(idx, np.flatnonzero(pred))
    assert cnt == pred.sum()
