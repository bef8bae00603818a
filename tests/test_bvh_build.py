"""Host checks of the BVH builder (both tree widths), compiled from the
library's own sources with the host compiler: every case builds, passes the
structural checks (node reachability, leaf runs, valid bits, every
dequantised child box containing its triangles, depth and stack bounds),
the 8-wide tree has fewer levels and fewer nodes than the 4-wide one, a
refit of unmoved triangles reproduces both node arrays bit
for bit, and a refit after moving every vertex passes the checks."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "shiro_amd", "csrc")


@pytest.fixture(scope="module")
def report(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path_factory.mktemp("bvh") / "bvh_check")
    cmd = ["g++", "-O2", "-std=c++17", "-pthread", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I" + CSRC,
           os.path.join(ROOT, "tests", "native", "bvh_check.cpp"), os.path.join(CSRC, "bvh_build.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    rows = []
    for line in out.stdout.splitlines():
        if line.startswith("case "):
            rows.append(dict(re.findall(r"(\w+)=(\S+)", line)))
    return out.returncode, rows, out.stdout


def test_too_deep_tree_is_built_again_with_capped_depth(report):
    """A depth bound the SAH tree exceeds does not fail the build: it is
    built again with its binary nodes at the bound made leaves, and passes
    every structural check at that bound."""
    rc, _, text = report
    capped = [dict(re.findall(r"(\w+)=(\S+)", l)) for l in text.splitlines() if l.startswith("capped ")]
    assert len(capped) == 2, text
    for r in capped:
        assert r["ok"] == "1" and int(r["depth"]) <= int(r["bound"]), r
        assert int(r["maxleaf"]) > 4, r


def test_every_case_builds_checks_and_refits(report):
    rc, rows, text = report
    assert rows, text
    assert rc == 0, text
    for r in rows:
        assert r["ok"] == "1" and r["refit_same"] == "1" and r["refit_moved"] == "1", r


def test_wide_tree_is_shallower(report):
    _, rows, _ = report
    for r in rows:
        n, n8 = int(r["nodes"]), int(r["nodes8"])
        assert n8 >= 1, r
        assert int(r["depth8"]) <= int(r["depth"]), r
        assert int(r["stack8"]) + 1 <= 96, r
        if int(r["n"]) >= 20000:
            assert n8 < n, r
            assert int(r["depth8"]) < int(r["depth"]), r
