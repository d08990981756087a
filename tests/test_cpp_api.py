"""Runs the C++ API tests of tests/cpp (built by symx_amd/Makefile into
symx_amd/lib/api_tests) on the GPU: FD self-check of compiled derivatives, conditional
potentials, CompiledInLoop (against the global assembly and against closed-form triangle
areas with an activity mask), Newton on a polynomial, Compiled<double> ops."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "symx_amd", "lib", "api_tests")


def test_api_tests_binary_is_built():
    assert os.path.exists(EXE), "build with __graft_entry__.build()"


@pytest.mark.gpu
def test_cpp_api(gpu):
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    cases = ["stable_neo_hookean", "conditional", "polynomial_newton", "compiled_ops", "loop_triangle_area", "add_dof_overloads"]
    for c in cases:
        assert f"PASS {c}\n" in r.stdout, c
    assert "FAIL" not in r.stdout
