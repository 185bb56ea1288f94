"""The C++ drop-in surface (include/prandtl_operator.hpp) driven from a C++
program on the GPU, checked against the oracle inside that program."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "test_operator")


@pytest.mark.gpu
def test_cpp_operator_matches_oracle():
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/test_operator is not built (run __graft_entry__.build())")
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and r.stdout.startswith("OK")


def test_cpp_operator_binary_built():
    assert os.path.exists(BIN)
