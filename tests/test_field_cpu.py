"""The device field arithmetic (csrc/field.hpp) compiled for the host and
checked against independent big-integer definitions: Goldilocks product,
128-bit reduction, add/sub, the F_p^2 and F_p^3 prepared-multiplier products,
the F_p^3 fold form and Karatsuba's F_p^3 form, the lazy dot products, the
two-fold forms of every field and the base-field accumulators, on edge
operands and 2 * 10^6 random ones (tests/cpp/field_check.cpp)."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "field_check")


def test_device_field_arithmetic_on_the_cpu():
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/field_check is not built (make -C tests/cpp)")
    r = subprocess.run([BIN, "2000000"], capture_output=True, text=True, timeout=300)
    out = {d["name"]: d for d in map(json.loads, r.stdout.splitlines())}
    assert r.returncode == 0, r.stdout
    assert out["base"]["ok"] and out["base"]["cases"] >= 2000000
    assert out["fp2"]["ok"] and out["fp3"]["ok"]
    assert out["kdot2"]["ok"] and out["kdot3"]["ok"] and out["fold2_fp2"]["ok"] and out["fold2_fp3"]["ok"]
    # two pending folds in one 160-bit sum, the carry-chain sum of a word and
    # three products, and the column sums of up to 4096 edge-valued products
    assert out["fold2_base"]["ok"] and out["fold2_base"]["cases"] >= 500000
