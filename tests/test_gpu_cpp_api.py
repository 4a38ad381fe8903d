"""The reference-shaped C++ API (include/edpp_primitives.hpp) driven by the
C++ program tests/cpp/test_primitives.cpp, whose checks follow the
reference's own drivers; it runs as a child process."""
import os
import subprocess

import pytest

from conftest import REPO

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1 << 20, 100_003])
def test_cpp_reference_api(torch_gpu, n):
    exe = os.path.join(REPO, "tests", "cpp", "test_primitives")
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/test_primitives is not built: run __graft_entry__.build()")
    r = subprocess.run([exe, str(n)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("OK: 0 failure(s)"), r.stdout
