"""TAPE3 line files (clblm_LineData.f90:52-70,99-256; RDLIN clblm_ODLAY.f90:2571-2798):
the writer's files read back through the library's reader (host code, no GPU), with
coupled lines, species broadening, short panels (padded to 250 entries with records
that are not lines), the second header record, the range selection and an end mark."""
import struct

import numpy as np
import pytest

from clblm_amd import workloads as W
from clblm_amd.tape3 import read_tape3, write_tape3


def _list():
    lines = W.make_lines(1200, 600.0, 700.0, seed=21, mols=(1, 2, 7))
    rng = np.random.default_rng(21)
    n = lines.n
    lines.brd_flg = (rng.uniform(size=(n, 7)) < 0.3).astype(np.int32)
    lines.brd_hw = rng.uniform(0.02, 0.1, (n, 7)).astype(np.float32)
    lines.brd_tmp = rng.uniform(0.5, 0.8, (n, 7)).astype(np.float32)
    lines.brd_shft = rng.uniform(-0.01, 0.01, (n, 7)).astype(np.float32)
    idx = np.arange(5, n, 9)
    return W.couple_lines(lines, idx, 1, rng.uniform(-0.05, 0.05, (idx.size, 4)), rng.uniform(-0.01, 0.01, (idx.size, 4)))


FIELDS = ("nu", "s", "a_air", "a_self", "e_lo", "tmpalf", "shift", "mol", "iflg", "brd_flg", "brd_hw", "brd_tmp",
          "brd_shft")


def test_round_trip_full_panels(tmp_path):
    lines = _list()
    p = tmp_path / "tape3"
    write_tape3(p, lines, negepp=True)
    got = read_tape3(p)
    n = lines.n
    assert got.n == ((n + 249) // 250) * 250
    for f in FIELDS:
        assert np.array_equal(getattr(got, f)[:n], getattr(lines, f)), f
    assert np.all(got.iflg[n:] == -1)


def test_short_panels_are_padded(tmp_path):
    lines = _list()
    sizes = []
    left = lines.n
    rng = np.random.default_rng(3)
    a = 0
    while a < lines.n:      # panels of 120-250 entries that never split a coupled pair
        k = int(rng.integers(120, 251))
        k = min(k, lines.n - a)
        if a + k < lines.n and lines.iflg[a + k - 1] in (1, 3):
            k -= 1
        sizes.append(k)
        a += k
    p = tmp_path / "tape3"
    write_tape3(p, lines, panel_sizes=sizes)
    got = read_tape3(p)
    assert got.n == 250 * len(sizes)
    a = 0
    for q, k in enumerate(sizes):
        rec = slice(250 * q, 250 * q + k)
        for f in FIELDS:
            assert np.array_equal(getattr(got, f)[rec], getattr(lines, f)[a:a + k]), (q, f)
        assert np.all(got.iflg[250 * q + k:250 * (q + 1)] == -1)
        a += k


def test_range_selection_and_end_mark(tmp_path):
    lines = _list()
    p = tmp_path / "tape3"
    write_tape3(p, lines)
    with open(p, "ab") as f:     # an end mark: a panel header whose fourth word is -99
        hdr = struct.pack("<ddii", 0.0, 0.0, 0, 0)
        hdr = hdr[:12] + struct.pack("<i", -99) + hdr[16:]
        f.write(struct.pack("<i", len(hdr)) + hdr + struct.pack("<i", len(hdr)))
        f.write(b"trailing bytes that are never read")
    full = read_tape3(p)
    assert full.n == ((lines.n + 249) // 250) * 250
    part = read_tape3(p, 640.0, 650.0)
    # the panels overlapping [640, 650] and the first one beyond it
    recs = [(r, lines.nu[250 * r:250 * (r + 1)][lines.iflg[250 * r:250 * (r + 1)] >= 0]) for r in range(full.n // 250)]
    keep = [r for r, v in recs if v.max() >= 640.0]
    out = []
    for r in keep:
        out.append(r)
        if recs[r][1].min() > 650.0:
            break
    assert part.n == 250 * len(out)
    assert np.array_equal(part.nu, full.nu[250 * out[0]:250 * (out[-1] + 1)])


def test_errors(tmp_path):
    from clblm_amd._lib import ClblmError
    with pytest.raises(ClblmError):
        read_tape3(tmp_path / "missing")
    p = tmp_path / "bad"
    p.write_bytes(b"\x10\x00\x00\x00short")
    with pytest.raises(ClblmError):
        read_tape3(p)
