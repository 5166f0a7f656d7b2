"""The C ABI from a plain C program (no Python between caller and library):
tests/c_abi/c_abi_check.c checks pf_apply against pf_op_connectivity,
pf_filter against the Chebyshev identity T_2(x) = 2x^2 - 1, and the error path."""
import importlib.util
import pathlib
import subprocess

import pytest

pytestmark = pytest.mark.gpu
HERE = pathlib.Path(__file__).resolve().parent / "c_abi"


def test_plain_c_caller():
    binary = HERE / "c_abi_check"
    if not binary.exists():  # built by __graft_entry__.build(); build it here otherwise
        spec = importlib.util.spec_from_file_location("_c_abi_build", HERE / "build.py")
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        mod.build()
    res = subprocess.run([str(binary)], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr)
    assert res.returncode == 0, res.stderr
    assert "c_abi_check: ok" in res.stdout
