"""The device's random numbers against the oracle, bit for bit
(tests/gpu_checks/philox_check.hip, built by __graft_entry__.build()).

* philox(): the kernels' Philox4x32-10 on Random123's known-answer inputs,
  edge words and 2^20 random (counter, key) blocks equals the oracle's
  orc_philox4x32_10, itself pinned by the published known answers.
* rng(): the counter layout of every draw (the cell in word x; tag, colour
  and trial in word y; the sweep's low 32 bits in z; the replica and the
  sweep's bits 32-43 in w; the seed as the key) on 2^18 random argument
  tuples at the fields' full widths: the device's block equals the oracle's
  Philox of the counter built from that layout.
* colour_perm_from(): every colour order drawn from 2^18 random blocks is a
  permutation of the 8 colours, each colour falls on each pass position
  with frequency 1/8, and the 8! orders are equally likely (chi-square)."""
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "patchyparticles_amd", "bin", "philox_check")

KAT = [([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
       ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
       ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1])]


@pytest.fixture(scope="module")
def device_draws(tmp_path_factory):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built: run __graft_entry__.build()")
    out = str(tmp_path_factory.mktemp("philox") / "draws.u32")
    res = subprocess.run([EXE, out], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    w = np.fromfile(out, dtype="<u4")
    pos = 0
    sections = []
    for width in (10, 13, 5):
        n = int(w[pos])
        sections.append(w[pos + 1:pos + 1 + n * width].reshape(n, width))
        pos += 1 + n * width
    assert pos == len(w)
    return sections


@pytest.mark.gpu
def test_device_philox_equals_oracle(device_draws, oracle):
    a = device_draws[0]
    assert len(a) >= (1 << 20)
    for i, (ctr, key, want) in enumerate(KAT):
        assert list(a[i, :4]) == ctr and list(a[i, 4:6]) == key
        assert list(a[i, 6:]) == want
    want = oracle.philox_n(a[:, :4], a[:, 4:6])
    assert np.array_equal(a[:, 6:], want)


@pytest.mark.gpu
def test_device_draw_counter_layout(device_draws, oracle):
    b = device_draws[1].astype(np.uint64)
    seed_lo, seed_hi, rep, sw_lo, sw_hi, cell, tag, col, trial = (b[:, k] for k in range(9))
    assert (sw_hi > 0).any() and (seed_hi > 0).any() and (rep >= 1 << 19).any() and (trial >= 1 << 15).any()
    ctr = np.stack([cell, tag | (col << 8) | (trial << 16), sw_lo, (rep & 0xFFFFF) | (sw_hi << 20)], 1)
    key = np.stack([seed_lo, seed_hi], 1)
    want = oracle.philox_n(ctr.astype(np.uint32), key.astype(np.uint32))
    assert np.array_equal(device_draws[1][:, 9:], want)


@pytest.mark.gpu
def test_device_colour_orders_are_uniform_permutations(device_draws):
    perm = device_draws[2][:, 4].astype(np.int64)
    n = len(perm)
    nib = np.stack([(perm >> (4 * k)) & 0xF for k in range(8)], 1)  # colour of pass k
    assert (np.sort(nib, axis=1) == np.arange(8)).all()
    # each colour on each pass position: 64 cells, expected n / 8 each
    table = np.zeros((8, 8))
    for k in range(8):
        table[k] = np.bincount(nib[:, k], minlength=8)
    chi2 = ((table - n / 8) ** 2 / (n / 8)).sum()
    dof = 49
    assert chi2 < dof + 6 * math.sqrt(2 * dof), chi2
    # the 8! orders: Lehmer code of each, then a chi-square over all of them
    code = np.zeros(n, np.int64)
    for k in range(8):
        smaller_later = (nib[:, k + 1:] < nib[:, k:k + 1]).sum(axis=1)
        code = code * (8 - k) + smaller_later
    counts = np.bincount(code, minlength=math.factorial(8))
    exp = n / math.factorial(8)
    chi2 = ((counts - exp) ** 2 / exp).sum()
    dof = math.factorial(8) - 1
    assert abs(chi2 - dof) < 6 * math.sqrt(2 * dof), chi2
