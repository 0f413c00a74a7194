"""CPU builds of the host code under sanitizers: the oracle's C restatement
(multithreaded solvers included) and the solver's host-side C++
(copsandrobbers_amd/csrc/cr_host.cpp: parsing, ranks, team moves and the
multithreaded blob writer), with AddressSanitizer + UndefinedBehaviorSanitizer
and, separately, ThreadSanitizer. Each build runs a driver under
tests/sanitize/ on small reference matrices; any report fails the run."""
import os
import shutil
import subprocess

import pytest

from conftest import REPO, matrix

SAN = os.path.join(REPO, "tests", "sanitize")
ORACLE = os.path.join(REPO, "oracle")
CSRC = os.path.join(REPO, "copsandrobbers_amd", "csrc")
MODES = {
    "asan_ubsan": ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"],
    "tsan": ["-fsanitize=thread"],
}
# verify_asan_link_order=0: the environment may preload a library ahead of the ASan runtime
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
           UBSAN_OPTIONS="halt_on_error=1", TSAN_OPTIONS="halt_on_error=1")


def build(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        pytest.skip(f"sanitizer build unavailable: {p.stderr[-400:]}")


def run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True, env=ENV, timeout=600)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-4000:]
    assert "runtime error" not in p.stderr and "WARNING: ThreadSanitizer" not in p.stderr, p.stderr[-4000:]
    return p.stdout


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no gcc")
@pytest.mark.parametrize("mode", sorted(MODES))
def test_oracle_under_sanitizers(tmp_path, mode):
    exe = str(tmp_path / "oracle_driver")
    build(["gcc", "-O1", "-g", "-std=c11", "-pthread", *MODES[mode], f"-I{ORACLE}",
           os.path.join(SAN, "oracle_driver.c"), os.path.join(ORACLE, "cr_oracle.c"), "-o", exe])
    for graph, k in (("peterson", 3), ("grid9", 2), ("outerplanar8", 3)):
        assert "oracle driver ok" in run([exe, matrix(graph), str(k), "4"])


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
@pytest.mark.parametrize("mode", sorted(MODES))
def test_host_code_under_sanitizers(tmp_path, mode):
    exe = str(tmp_path / "host_driver")
    build(["g++", "-O1", "-g", "-std=c++17", "-pthread", *MODES[mode], f"-I{CSRC}",
           f"-I{os.path.join(REPO, 'include')}", os.path.join(SAN, "host_driver.cpp"),
           os.path.join(CSRC, "cr_host.cpp"), "-o", exe])
    for graph, k in (("peterson", 3), ("grid36", 2)):
        out = str(tmp_path / graph)
        os.makedirs(out)
        assert "host driver ok" in run([exe, matrix(graph), str(k), out])
