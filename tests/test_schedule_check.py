"""The host-side schedule check (wgpu-voxel-engine_amd/csrc/vx_schedule.h)
that guards the arrays the fill and emit launches index through: compiled
on its own with the host C++ compiler and run on permutations and on the
corruptions it must refuse (an index out of range, a duplicate, a negative
entry).  CPU only."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "wgpu-voxel-engine_amd", "csrc")

DRIVER = r"""
#include <cstdio>
#include <cstdlib>
#include "vx_schedule.h"
int main(int argc, char **argv) {
    // argv[1..]: the schedule; prints 1 if it is a permutation of 0..n-1
    const int n = argc - 1;
    int32_t *a = n ? (int32_t *)malloc(sizeof(int32_t) * n) : nullptr;
    for (int i = 0; i < n; i++) a[i] = (int32_t)atoi(argv[i + 1]);
    printf("%d\n", vx::is_schedule(a, n) ? 1 : 0);
    free(a);
    return 0;
}
"""


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    d = tmp_path_factory.mktemp("sched")
    src, exe = d / "drv.cpp", d / "drv"
    src.write_text(DRIVER)
    subprocess.run([cxx, "-std=c++17", "-O1", "-I", CSRC, str(src), "-o", str(exe)], check=True)

    def run(seq):
        out = subprocess.run([str(exe), *map(str, seq)], capture_output=True, text=True, check=True)
        return out.stdout.strip() == "1"
    return run


def test_permutations_are_accepted(checker):
    assert checker([])
    assert checker([0])
    assert checker(list(range(17)))
    assert checker(list(range(16, -1, -1)))
    assert checker([3, 0, 4, 1, 2])


def test_corruptions_are_refused(checker):
    assert not checker([1])              # out of range
    assert not checker([0, 2])           # out of range
    assert not checker([0, 0, 1])        # duplicate (and one index missing)
    assert not checker([0, -1, 1])       # negative
    assert not checker([2, 1, 3])        # off by one
