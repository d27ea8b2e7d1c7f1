"""Host side of the bit-sliced kernels on the CPU: the group builder, the
level keys and the slot words, replayed with the kernels' probe rule
(tests/cpu_replay/bitgroups_replay.cpp) against sorted-merge intersections."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "siml_amd", "csrc")


@pytest.fixture(scope="module")
def replay(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ is not available")
    exe = str(tmp_path_factory.mktemp("replay") / "bitgroups_replay")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, "-o", exe,
                    os.path.join(HERE, "cpu_replay", "bitgroups_replay.cpp"),
                    os.path.join(CSRC, "bitgroups.cpp"), "-lpthread"], check=True)
    return exe


@pytest.mark.parametrize("seed,nrows,ncols,nonascii,big,plant", [
    (1, 300, 200, 0, 0, 0),        # every row groups
    (2, 700, 150, 2, 2, 0),        # some non-ASCII ids and counts above 15 -> merge rows
    (3, 64, 300, 0, 30, 0),        # many repeated Lingos
    (4, 1000, 60, 0, 0, 0),        # several groups per chunk of the parallel builder
    (5, 300, 200, 0, 0, 1),        # keys sharing both slots: give-back and merge-row paths
    (6, 5000, 40, 0, 0, 1),
])
def test_group_tables_reproduce_intersections(replay, seed, nrows, ncols, nonascii, big, plant):
    res = subprocess.run([replay, str(seed), str(nrows), str(ncols), str(nonascii), str(big), str(plant)],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stdout + res.stderr
    assert " 0 differ" in res.stdout
