"""GPU: the out-of-core fourn (Four_FS.rs) against the oracle's fourn/four1,
with device budgets small enough to force every streaming path (row chunks,
batched slabs, column bands, the 1-D four-step with its scratch)."""
import numpy as np
import pytest

import numrs_amd
import oracle
from conftest import rel_l2
from numrs_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,budget", [
    ((32, 64, 128), 128 << 10),     # column bands on both strided axes
    ((32, 64, 128), 2 << 20),       # whole slabs batched (axis 1), bands (axis 0)
    ((32, 64, 128), 64 << 20),      # fits: one upload, the in-memory fourn plan
    ((8, 16384), 1 << 20),          # last axis four-step rows in chunks, short strided axis
    ((4096, 4), 128 << 10),         # long strided axis in bands of 1 column
    ((2048, 16), 256 << 10),        # long strided axis in bands of 2 columns
    ((4, 65536), 2 << 20),          # a row and its workspace fill the budget: one slot
])
@pytest.mark.parametrize("isign", [1, -1])
def test_fourfs_nd(shape, budget, isign):
    x = synth.normal(90, 2 * int(np.prod(shape)))
    g = x.copy()
    numrs_amd.fourfs(g, list(shape), len(shape), isign, device_bytes=budget)
    assert rel_l2(g, oracle.fourn(x, list(shape), isign)) < 1e-12, (shape, budget)


@pytest.mark.parametrize("n,budget", [(1 << 16, 256 << 10), (1 << 20, 4 << 20), (1 << 12, 64 << 20)])
def test_fourfs_1d(n, budget):
    x = synth.uniform(91, 2 * n)
    g = x.copy()
    numrs_amd.fourfs(g, [n], 1, 1, device_bytes=budget)
    assert rel_l2(g, oracle.four1_nr(x, 1)) < 1e-12, (n, budget)


def test_fourfs_1d_f32():
    n = 1 << 18
    x = synth.uniform(95, 2 * n).astype(np.float32)
    g = x.copy()
    numrs_amd.fourfs(g, [n], 1, -1, device_bytes=512 << 10)
    assert rel_l2(g, oracle.four1_nr(x.astype(np.float64), -1)) < 2e-6


def test_fourfs_f32():
    shape = (16, 32, 64)
    x = synth.normal(92, 2 * int(np.prod(shape))).astype(np.float32)
    g = x.copy()
    numrs_amd.fourfs(g, list(shape), 3, 1, device_bytes=32 << 10)
    assert rel_l2(g, oracle.fourn(x.astype(np.float64), list(shape), 1)) < 2e-6


def test_fourfs_file_backed(tmp_path):
    # Four_FS.rs:416-425 (test_fourfs_basic, nn = [8, 8]) with data in the
    # temporary file, plus a 1-D transform larger than the budget (scratch in
    # the second file) and the round trip.
    fs = numrs_amd.FourFS.new(tmp_path).with_memory_mapping()
    x = synth.normal(93, 2 * 64)
    fs.write_data(x)
    fs.fourfs([8, 8], 2, 1)
    assert rel_l2(fs.read_data(), oracle.fourn(x, [8, 8], 1)) < 1e-12
    fs2 = numrs_amd.FourFS(tmp_path, device_bytes=128 << 10)
    n = 1 << 15
    y = synth.uniform(94, 2 * n)
    fs2.write_data(y)
    fs2.fourfs([n], 1, 1)
    assert rel_l2(fs2.read_data(), oracle.four1_nr(y, 1)) < 1e-12
    fs2.fourfs([n], 1, -1)
    assert rel_l2(fs2.read_data(), n * y) < 1e-12
