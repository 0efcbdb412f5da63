"""The reference's apps/bfs.cpp written against the C++ mirror API
(include/graphit_amd.hpp), built by build() and run as a process: parents
valid against the oracle, in both the pure-push and the hybrid mode."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "easy_graphit_amd", "apps", "bfs_app")


@pytest.mark.parametrize("mode", [[], ["power", "0.05"]])
def test_cpp_bfs_app(gi, tmp_path, mode):
    if not os.path.exists(APP):
        pytest.skip("C++ app not built")
    out = tmp_path / "parent.bin"
    args = [APP, "14", "16", "7"] + (mode or ["push", "0"]) + [str(out)]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    parent = np.fromfile(out, np.int32)
    s, d = O.rmat_edges(14, 16, 27491095)
    host = O.build_csr(1 << 14, s, d, True)
    depth = O.bfs_depth(host, 7)
    assert O.bfs_parent_violations(host, 7, depth, parent) == 0
    assert f"reached {(depth >= 0).sum()}" in r.stdout
