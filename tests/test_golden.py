"""The oracle still reproduces the committed golden fixtures
(tests/golden/oracle_golden.npz, written by tests/golden/make_golden.py)
bit for bit."""
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden" / "oracle_golden.npz"


def test_oracle_reproduces_golden_fixtures():
    import sys
    sys.path.insert(0, str(GOLDEN.parent))
    from make_golden import cases

    stored = np.load(GOLDEN)  # allow_pickle stays False
    now = cases()
    assert sorted(stored.files) == sorted(now)
    for k in stored.files:
        assert stored[k].dtype == now[k].dtype, k
        np.testing.assert_array_equal(stored[k], now[k], err_msg=k)


@pytest.mark.parametrize("key", ["noise_binary", "enc_u8", "filter_tri_lut"])
def test_golden_fixtures_are_byte_data(key):
    assert np.load(GOLDEN)[key].dtype == np.uint8
