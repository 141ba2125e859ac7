"""Runs the C++ driver written against the reference-shaped interface
(include/gdict.hpp): the forced-unified scenarios of
FliX/paper_lsmu_sweep.cu:441-700 with a host unordered_map expectation."""

import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_forced_unified_cpp_driver(dev):
    exe = os.path.join(HERE, "cpp", "forced_unified")
    assert os.path.exists(exe), "build() compiles tests/cpp first"
    res = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    out = res.stdout
    for scenario in ("sparse_unique", "dense_unique", "duplicate_newest_wins",
                     "tombstone_visibility"):
        assert "scenario=%s" % scenario in out
    assert "CANONICAL_MULTILEVEL_ROLLOVER_VALIDATION status=0 errors=0" in out
    assert "controlled_status=1 controlled_exception=1" in out
    assert "FAILURES 0" in out
