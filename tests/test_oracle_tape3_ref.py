"""TAPE3 line files against the reference's own reader (oracle/_ref, CPU only).

`ref_tape3` (oracle/ref_bind.f90) reads a file with the reference's record types
INPUT_HEADER / INPUT_BLOCK (src/clblm_src/clblm_LineData.f90:52-70) and its record
reader BUFIN_sgl (src/common/util_linux_intel.f90:121-150), in the order RDLIN and
RDLNFL read (clblm_ODLAY.f90:2626-2658, clblm_ODLAY_LineF4.f90:1408-1418).  Files
written by clblm_amd.tape3 must come back field for field, through that reader and
through the library's own (clblm_tape3_read), so the writer, the library's reader and
the reference agree on the record layout: word order, widths, the panel header, the
negative-E'' second header record and short panels.  Skipped where the reference
sources are absent."""
import ctypes
import os

import numpy as np
import pytest

from clblm_amd import workloads as W
from clblm_amd.tape3 import read_tape3, write_tape3

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "libclblm_ref.so")
pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built (reference sources absent)")


def _ref_read(path, cap=64):
    lib = ctypes.CDLL(REF, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
    f32, f64, i32 = np.float32, np.float64, np.int32
    out = dict(nrec=np.zeros(cap, i32), nwds=np.zeros(cap, i32), vmin=np.zeros(cap, f64), vmax=np.zeros(cap, f64),
               nu=np.zeros((cap, 250), f64), s=np.zeros((cap, 250), f32), a_air=np.zeros((cap, 250), f32),
               e_lo=np.zeros((cap, 250), f32), mol=np.zeros((cap, 250), i32), a_self=np.zeros((cap, 250), f32),
               tmpalf=np.zeros((cap, 250), f32), shift=np.zeros((cap, 250), f32), iflg=np.zeros((cap, 250), i32),
               brd_flg=np.zeros((cap, 250, 7), i32), brd_dat=np.zeros((cap, 250, 21), f32),
               sdep=np.zeros((cap, 250), f32), hdr_i=np.zeros(5, i32), hdr_f=np.zeros(2, f32))
    npanel = ctypes.c_int()
    order = ("nrec", "nwds", "vmin", "vmax", "nu", "s", "a_air", "e_lo", "mol", "a_self", "tmpalf", "shift", "iflg",
             "brd_flg", "brd_dat", "sdep", "hdr_i", "hdr_f")
    lib.ref_tape3(str(path).encode() + b"\0", ctypes.c_int(cap), ctypes.byref(npanel),
                  *[ctypes.c_void_p(out[k].ctypes.data) for k in order])
    out["npanel"] = npanel.value
    return out


def _list():
    lines = W.make_lines(1300, 600.0, 700.0, seed=31, mols=(1, 2, 7))
    rng = np.random.default_rng(31)
    n = lines.n
    lines.brd_flg = (rng.uniform(size=(n, 7)) < 0.3).astype(np.int32)
    lines.brd_hw = rng.uniform(0.02, 0.1, (n, 7)).astype(np.float32)
    lines.brd_tmp = rng.uniform(0.5, 0.8, (n, 7)).astype(np.float32)
    lines.brd_shft = rng.uniform(-0.01, 0.01, (n, 7)).astype(np.float32)
    lines.shift = rng.uniform(-0.005, 0.005, n).astype(np.float32)
    lines.e_lo[::17] = -np.float32(12.5)          # negative E'' entries (the '^' header case)
    idx = np.arange(7, n, 11)
    cpl = W.couple_lines(lines, idx, np.where(np.arange(idx.size) % 3 == 0, 3, 1),
                         rng.uniform(-0.05, 0.05, (idx.size, 4)), rng.uniform(-0.01, 0.01, (idx.size, 4)))
    return cpl


def _panels(lines, sizes):
    a = 0
    for k in sizes:
        yield a, k
        a += k


@pytest.mark.parametrize("short", [False, True])
@pytest.mark.parametrize("negepp", [False, True])
def test_reference_reader_reads_the_writer(tmp_path, short, negepp):
    lines = _list()
    n = lines.n
    if short:
        rng = np.random.default_rng(5)
        sizes, a = [], 0
        while a < n:
            k = min(int(rng.integers(90, 251)), n - a)
            if a + k < n and lines.iflg[a + k - 1] in (1, 3):
                k -= 1          # a coupled line and its coefficient record share a panel
            sizes.append(k)
            a += k
    else:
        sizes = [250] * (n // 250) + ([n % 250] if n % 250 else [])
    p = tmp_path / "tape3"
    write_tape3(p, lines, panel_sizes=sizes, negepp=negepp)
    ref = _ref_read(p)
    assert ref["npanel"] == len(sizes)
    line = lines.iflg >= 0
    assert ref["hdr_i"][0] == np.unique(lines.mol[line] % 100).size      # LINMOL
    assert ref["hdr_i"][1] == int(line.sum())                          # LINCNT
    assert ref["hdr_i"][3] == (1 if negepp else 0)                    # second header record
    assert ref["hdr_i"][4] == 1                                       # isotope flag 'I'
    assert ref["hdr_f"][0] == np.float32(lines.nu[line].min())
    assert ref["hdr_f"][1] == np.float32(lines.nu[line].max())
    ours = read_tape3(p)
    assert ours.n == 250 * len(sizes)
    for q, (a, k) in enumerate(_panels(lines, sizes)):
        assert ref["nrec"][q] == k and ref["nwds"][q] == 39 * 250
        sl = slice(a, a + k)
        for f in ("nu", "s", "a_air", "e_lo", "mol", "a_self", "tmpalf", "shift", "iflg"):
            want = getattr(lines, f)[sl]
            assert np.array_equal(ref[f][q, :k], want), (q, f)
            assert np.array_equal(getattr(ours, f)[250 * q:250 * q + k], ref[f][q, :k]), (q, f)
        assert np.array_equal(ref["brd_flg"][q, :k], lines.brd_flg[sl])
        dat = ref["brd_dat"][q, :k].reshape(k, 7, 3)
        for j, f in enumerate(("brd_hw", "brd_tmp", "brd_shft")):
            assert np.array_equal(dat[:, :, j], getattr(lines, f)[sl]), (q, f)
            assert np.array_equal(getattr(ours, f)[250 * q:250 * q + k], dat[:, :, j]), (q, f)
        assert np.array_equal(getattr(ours, "brd_flg")[250 * q:250 * q + k], ref["brd_flg"][q, :k])
        # the panel bounds are those of the panel's lines
        pl = line[sl]
        if pl.any():
            assert ref["vmin"][q] == lines.nu[sl][pl].min() and ref["vmax"][q] == lines.nu[sl][pl].max()
        assert np.all(ours.iflg[250 * q + k:250 * (q + 1)] == -1)
