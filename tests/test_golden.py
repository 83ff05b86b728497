"""The oracle reproduces its committed golden fixture bit for bit (tests/golden/make_golden.py)."""
import os

import numpy as np

from tests.golden.make_golden import compute

HERE = os.path.dirname(os.path.abspath(__file__))


def test_oracle_matches_golden_fixture():
    g = np.load(os.path.join(HERE, "golden", "config0_oracle.npz"))
    now = compute()
    assert set(g.files) == set(now)
    for k in g.files:
        np.testing.assert_array_equal(now[k], g[k], err_msg=k)
