"""C++ operator layer (include/velodb_amd/operators.hpp: TaskManager,
SortOperator, FilterCompactionOperator, PKFKJoinGroupSumOperator) on the
GPU: runs tests/cpp/test_operators, the ORDER BY cases of the reference's
tests/query/test_order_by.cpp plus join and queue checks.

The binary starts resident queues (cooperative launches).  On this runtime a
process cannot make a cooperative launch on a GPU once another process has
made one there, so this module must run before any test in the same pytest
process that starts a queue: tests/conftest.py puts its tests first in every
selection."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_operators")


@pytest.mark.parametrize("mode", [[], ["--persistent"]], ids=["launch", "persistent"])
def test_cpp_operator_suite(mode):
    """The whole suite with a launch-mode TaskManager, then with a persistent
    one whose resident queue kernel runs every operator's commands."""
    assert os.path.exists(BIN), "tests/cpp/test_operators is not built (__graft_entry__.build())"
    r = subprocess.run([BIN] + mode, capture_output=True, text=True, timeout=110)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert "OK: 0 failed checks" in r.stdout
