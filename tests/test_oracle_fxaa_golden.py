"""The oracle's FXAA against hand-worked cases (tests/golden/fxaa_cases.json,
written by tests/golden/make_fxaa_cases.py from closed forms derived by hand
from the reference's formula, kernels/ssaa.cuh:141-182).

The cases pin the constants that define the reference's FXAA output: the
luma weights and their /255 scale (D14), REDUCE_MIN = 1/128, REDUCE_MUL =
1/32, SPAN_MAX = 8, the A / B selection and the clamp-to-edge taps (D13).
Each case carries the values it would take with one of those changed, and
the test checks that each is far outside the tolerance: a change of any of
them in the oracle fails this test.  The GPU's FXAA equals the oracle's bit
for bit (test_aa_passes_bit_exact), so the pins carry over to it.
"""
import json
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = json.load(open(os.path.join(HERE, "golden", "fxaa_cases.json")))["cases"]
REL = 1e-5  # float32 evaluation against closed forms in double precision


def tolerance(expected):
    return REL * max(1.0, float(np.abs(expected).max()))


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_fxaa_matches_the_hand_worked_case(case):
    img = np.array(case["image"], np.float32)
    x, y = case["pixel"]
    out = O.fxaa(img)[y, x]
    exp = np.array(case["expected"])
    assert np.abs(out - exp).max() <= tolerance(exp), (out.tolist(), exp.tolist())


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_each_case_tells_the_constants_apart(case):
    exp = np.array(case["expected"])
    assert case["altered"]
    for what, alt in case["altered"].items():
        assert np.abs(np.array(alt) - exp).max() > 100 * tolerance(exp), what


def test_cases_cover_every_constant_and_both_selections():
    altered = {k for c in CASES for k in c["altered"]}
    for needle in ("/255", "REDUCE_MIN", "REDUCE_MUL", "SPAN_MAX", "clamped", "luma weights"):
        assert any(needle in k for k in altered), needle
    names = [c["name"] for c in CASES]
    assert any(n.endswith("selects_A") for n in names) and any(n.endswith("selects_B") for n in names)
