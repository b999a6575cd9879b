"""The C++ facade (include/mci_amd/MCIntegrator.hpp) running the checks of the reference's
test/ut2, ut3 and ut4 on the GPU library (tests/cpp/test_facade.cpp)."""
import os
import subprocess

import pytest

from mcitest import ROOT

CPP = os.path.join(ROOT, "tests", "cpp")


def test_facade_header_compiles_standalone():
    # the facade needs only the C-ABI header and the C++ standard library
    r = subprocess.run(["g++", "-std=c++14", "-Wall", "-Wextra", "-Werror", "-fsyntax-only", "-x", "c++",
                        "-I", os.path.join(ROOT, "include"), "-"],
                       input='#include "mci_amd/MCIntegrator.hpp"\nint main() { mci::MCI *p = nullptr; (void)p; }\n',
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_reference_unit_tests_through_facade():
    exe = os.path.join(CPP, "build", "test_facade")
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/build/test_facade is not built (run __graft_entry__.build())")
    print("running", exe, flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "checks passed" in r.stdout
