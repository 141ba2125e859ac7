"""Runs the C++ drivers written against the reference-shaped interface
(include/gdict.hpp): the forced-unified scenarios of
FliX/paper_lsmu_sweep.cu:441-700 with a host unordered_map expectation
(GPULSMOpt), and the FliX interface test of FliX/main.cu:98-156 plus the
reference driver's update rounds with a host std::map expectation
(cg_rtx_index_updates)."""

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


def test_flix_cpp_driver(dev):
    exe = os.path.join(HERE, "cpp", "flix_interface")
    assert os.path.exists(exe), "build() compiles tests/cpp first"
    res = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    out = res.stdout
    assert "PASS known answer after the deletes" in out
    assert "PASS key32" in out and "PASS key64" in out
    assert "PASS use before build throws std::invalid_argument" in out
    assert "FAILURES 0" in out
