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


MFEM_BIN = os.path.join(HERE, "cpp", "test_mfem_binding")


@pytest.mark.gpu
def test_mfem_binding_matches_oracle():
    """INTEGRATION.md's DGSEMOperatorHIP (an mfem::TimeDependentOperator over
    the library) compiled against the MFEM stand-in, Mult and
    mfem::RK3SSPSolver steps on 3-D Navier-Stokes against the oracle."""
    if not os.path.exists(MFEM_BIN):
        pytest.fail("tests/cpp/test_mfem_binding is not built (run __graft_entry__.build())")
    r = subprocess.run([MFEM_BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and r.stdout.rstrip().endswith("OK")


def test_mfem_binding_built():
    assert os.path.exists(MFEM_BIN)
