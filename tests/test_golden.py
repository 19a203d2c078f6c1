"""The oracle against the committed golden fixtures (tests/golden, written by
tests/golden/make_fixtures.py): the loop's output planes and the Ooura 8x8 transforms must
reproduce the stored bits, and the objective its stored value.  The GPU side of the same
fixtures is tests/test_gpu_parity.py::test_golden_fixtures_bit_exact."""
import numpy as np
import pytest

import oracle
import golden_cases
from golden_cases import dct_blocks, loop_cases

CASES = loop_cases()


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_reproduces_loop_fixture(case):
    planes = oracle.compute(case["coefs"], case["weight"], case["pweight"], case["iterations"], case["max_w"],
                            case["max_h"], norm_mode="f64")
    for c, (a, b) in enumerate(zip(planes, case["planes"])):
        assert np.array_equal(bits(a), bits(b)), f"channel {c}"
    total, terms = oracle.objective(case["coefs"], case["weight"], case["pweight"], case["max_w"], case["max_h"],
                                    planes)
    assert np.array_equal(np.array([total, *terms]), case["objective"])


def test_fixture_inputs_cover_every_sampling_path():
    seen = {tuple((k.horizontal_samp_factor, k.vertical_samp_factor) for k in c["coefs"]) for c in CASES}
    assert ((1, 1),) in seen                      # luma only
    assert ((1, 1), (2, 2), (2, 2)) in seen       # 4:2:0
    assert ((1, 1), (2, 1), (2, 1)) in seen       # 4:2:2
    assert ((1, 1), (1, 2), (1, 2)) in seen       # 4:4:0
    assert ((1, 1), (1, 1), (1, 1)) in seen       # 4:4:4
    assert {c["mode"] for c in CASES} == {"tv", "tgv"}


def test_oracle_reproduces_dct_fixture():
    d = dct_blocks()
    assert np.array_equal(bits(oracle.dct8x8(d["blocks"])), bits(d["dct"]))
    assert np.array_equal(bits(oracle.idct8x8(d["blocks"])), bits(d["idct"]))


def test_max_grid_fixture_is_current():
    """The 16384 x 16384 fixture was written for the parameters the GPU test uses, and its
    coefficient generator is deterministic (checked on a small grid)."""
    fx = golden_cases.max_grid_case()
    assert len(fx["sha256"]) == 3 and all(len(h) == 64 for h in fx["sha256"])
    a = golden_cases.random_420(64, 32, 7, 75)
    b = golden_cases.random_420(64, 32, 7, 75)
    assert [c.rounded_px_w for c in a] == [64, 32, 32]
    for x, y in zip(a, b):
        assert np.array_equal(x.dct_coefs, y.dct_coefs)
