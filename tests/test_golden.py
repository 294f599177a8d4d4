"""The oracle reproduces the committed golden fixtures (tests/golden/*.npz,
written by tests/golden/make_golden.py); a change in the oracle's arithmetic
or loop semantics shows up here. The GPU build is checked against the same
fixtures in test_gpu_golden.py."""
import os

import numpy as np
import pytest

from tests.golden import make_golden as M

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("name", sorted(M.CASES))
def test_oracle_reproduces_fixture(name):
    ref = np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False)
    new = M.CASES[name]()
    for k in ref.files:
        a, b = np.asarray(ref[k]), np.asarray(new[k])
        if a.dtype.kind in "iub":
            assert np.array_equal(a, b), k
        else:
            assert np.allclose(a, b, rtol=1e-13, atol=1e-14), k
