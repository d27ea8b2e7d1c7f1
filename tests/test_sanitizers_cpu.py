"""The library's host C++ under AddressSanitizer and UndefinedBehaviorSanitizer
on the CPU: the SMILES packer (pack.cpp, randomised byte strings through its
three entry points) and the bit-sliced row-group builder (bitgroups.cpp,
replayed against sorted-merge intersections).  Any sanitizer report ends the
run with a non-zero status and fails the test."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "siml_amd", "csrc")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
       "-static-libasan", "-g", "-O1"]
ENV = dict(os.environ,
           ASAN_OPTIONS="halt_on_error=1:detect_leaks=1:verify_asan_link_order=0",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


def _build(tmp_path_factory, name, sources):
    if shutil.which("g++") is None:
        pytest.skip("g++ is not available")
    exe = str(tmp_path_factory.mktemp("san") / name)
    res = subprocess.run(["g++", "-std=c++17", *SAN, "-I", CSRC, "-o", exe, *sources, "-lpthread"],
                         capture_output=True, text=True)
    if res.returncode != 0 and "sanitizer" in res.stderr.lower() and "cannot find" in res.stderr.lower():
        pytest.skip("the sanitizer runtimes are not installed")
    assert res.returncode == 0, res.stderr
    return exe


@pytest.fixture(scope="module")
def pack_fuzz(tmp_path_factory):
    return _build(tmp_path_factory, "pack_fuzz",
                  [os.path.join(HERE, "cpu_replay", "pack_fuzz.cpp"), os.path.join(CSRC, "pack.cpp")])


@pytest.fixture(scope="module")
def replay(tmp_path_factory):
    return _build(tmp_path_factory, "bitgroups_replay",
                  [os.path.join(HERE, "cpu_replay", "bitgroups_replay.cpp"), os.path.join(CSRC, "bitgroups.cpp")])


@pytest.mark.parametrize("seed,n", [(1, 3000), (2, 3000), (3, 1), (4, 0)])
def test_packer_under_sanitizers(pack_fuzz, seed, n):
    res = subprocess.run([pack_fuzz, str(seed), str(n)], capture_output=True, text=True, env=ENV, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    assert " 0 differ" in res.stdout


@pytest.mark.parametrize("seed,nrows,ncols,nonascii,big,plant", [
    (1, 300, 200, 0, 0, 0),
    (2, 700, 150, 2, 2, 0),
    (3, 64, 300, 0, 30, 0),
    (4, 1000, 60, 0, 0, 0),
    (5, 300, 200, 0, 0, 1),
])
def test_group_builder_under_sanitizers(replay, seed, nrows, ncols, nonascii, big, plant):
    res = subprocess.run([replay, str(seed), str(nrows), str(ncols), str(nonascii), str(big), str(plant)],
                         capture_output=True, text=True, env=ENV, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    assert " 0 differ" in res.stdout
