"""The FP64 atomic-rate hook (egsgpu_atomic_rate) that the benchmark quotes
as the roof of the tally flushes and the gradient sweep: every add lands,
for each access pattern, and the rate is a positive finite number."""
import math

import pytest

from egsnrc_amd import _lib


@pytest.mark.gpu
@pytest.mark.parametrize("pattern,naddr", [(0, 64 ** 3), (1, 4096), (2, 1 << 16)])
def test_every_atomic_add_lands(pattern, naddr):
    rate, total, issued = _lib.atomic_rate(pattern, naddr, 1 << 20)
    assert issued >= 1 << 20
    assert total == 2 * issued
    assert math.isfinite(rate) and rate > 0


def test_atomic_rate_rejects_bad_arguments():
    """(argument checks come before any device call: runs without a GPU)"""
    with pytest.raises(_lib.EngineError):
        _lib.atomic_rate(3, 4096, 1 << 10)
    with pytest.raises(_lib.EngineError):
        _lib.atomic_rate(0, 8, 1 << 10)
