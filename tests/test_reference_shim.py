"""The reference-side binding (hybrid-aphi-solver_amd/shim) compiles against the
reference's own headers with the exact signatures it replaces (refill,
include/aphi_solver/assembly.hpp:197-199; solve_symmetric,
include/aphi_solver/factorization.hpp:206-210): a definition whose signature
drifted would declare a new overload instead and fail the -Werror check of
missing declarations."""
import os
import subprocess

import pytest

from conftest import ROOT

REFERENCE = "/root/reference/include"


def test_binding_compiles_against_the_reference_headers():
    if not os.path.isdir(REFERENCE):
        pytest.skip("the reference's headers are not in this container")
    src = os.path.join(ROOT, "hybrid-aphi-solver_amd", "shim", "aphi_gpu_backend.cpp")
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror", "-Wmissing-declarations",
           f"-I{REFERENCE}", f"-I{os.path.join(ROOT, 'include')}", f"-I{os.path.dirname(src)}", "-I/opt/rocm/include",
           "-D__HIP_PLATFORM_AMD__", src]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
