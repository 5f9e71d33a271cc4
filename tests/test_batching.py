"""Host-side row batching of the big-row contraction (csrc/batching.h),
compiled with g++ and checked on random row sizes: the batches cover the
rows in order, each holds at most the cap unless it is a single larger row,
each is maximal (the next row would not fit), and the reported largest
total is the largest batch's."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = tmp_path_factory.mktemp("batching") / "batching_check"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "pennsylvaniagpugraphclustering_amd",
                                                                             "csrc"),
                           os.path.join(ROOT, "tests", "batching_check.cpp"), "-o", str(exe)])
    return str(exe)


def run(exe, sizes, cap):
    inp = f"{len(sizes)} {cap} " + " ".join(str(int(x)) for x in sizes)
    out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    return [int(x) for x in out[0].split()], int(out[1])


@pytest.mark.parametrize("seed", range(20))
def test_batch_cuts(checker, seed):
    rng = np.random.default_rng(seed)
    m = int(rng.integers(1, 300))
    sizes = rng.integers(0, 5000, m) * (rng.random(m) < 0.9)
    cap = int(rng.choice([1, 100, 4096, 20000, 10 ** 9]))
    cuts, largest = run(checker, sizes, cap)
    assert cuts[0] == 0 and cuts[-1] == m
    assert all(a < b for a, b in zip(cuts, cuts[1:]))
    totals = []
    for a, b in zip(cuts, cuts[1:]):
        t = int(sizes[a:b].sum())
        totals.append(t)
        assert t <= cap or b - a == 1
        if b < m:
            assert t + int(sizes[b]) > cap  # maximal
    assert largest == max(totals)


def test_single_row_larger_than_cap(checker):
    cuts, largest = run(checker, [10, 5000, 10], 100)
    assert cuts == [0, 1, 2, 3] and largest == 5000
