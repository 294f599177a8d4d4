"""The oracle reproduces the committed small golden fixtures
(tests/golden/*.npz, written by tests/golden/make_golden.py); a change in
the oracle's arithmetic or loop semantics shows up here. The GPU build is
checked against the same fixtures in test_gpu_golden.py. The fixtures of the
benchmark configurations at their own sizes (c1_129, c2_1025, c3_4097) take
the oracle minutes to write; the GPU tests of test_gpu_configs.py check
against them."""
import os

import numpy as np
import pytest

from tests.golden import make_golden as M

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("name", M.SMALL)
def test_oracle_reproduces_fixture(name):
    ref = np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False)
    new = M.CASES[name]()
    for k in ref.files:
        a, b = np.asarray(ref[k]), np.asarray(new[k])
        if a.dtype.kind in "iub":
            assert np.array_equal(a, b), k
        else:
            assert np.allclose(a, b, rtol=1e-13, atol=1e-14), k
