"""TAPE3 line files from Python: ``write_tape3`` writes a LineList in the layout LNFL
writes and the reference reads (header record, then per panel its header VMIN, VMAX,
NREC, NWDS and its 250-line block: src/clblm_src/clblm_LineData.f90:52-70,99-256,
src/clblm_src/clblm_ODLAY.f90:2571-2798); ``read_tape3`` reads one through the
library's reader (clblm_tape3_read in libclblm_hip.so, host code only).

Files are Fortran unformatted sequential: each record is framed by its byte length
as a little-endian 4-byte integer before and after it."""
from __future__ import annotations

import ctypes
import struct

import numpy as np

from . import _lib
from .workloads import LineList

NREC = 250
NWDS = 39 * NREC


def _record(f, payload: bytes):
    n = struct.pack("<i", len(payload))
    f.write(n)
    f.write(payload)
    f.write(n)


def _header(lines: LineList, negepp: bool) -> bytes:
    hlinid = [b"CLBLMHIP", b"LNFLFILE", b"        ", b"        ", b"        ", b"        ",
              b"       ^" if negepp else b"        ", b"        ", b"        ", b"I       "]
    mols = lines.mol[lines.iflg >= 0] % 100
    molcnt = np.bincount(mols, minlength=65)[1:65].astype(np.int32)
    b = b"".join(hlinid)
    b += b"".join(b"MOL%05d" % m for m in range(1, 65))            # BMOLID(64)
    b += molcnt.tobytes() + np.zeros(64, np.int32).tobytes() * 2    # MOLCNT, MCNTLC, MCNTNL
    b += np.zeros(64, np.float32).tobytes()                         # SUMSTR
    line = lines.iflg >= 0
    b += struct.pack("<iffiiiii", int((molcnt > 0).sum()), float(lines.nu[line].min()), float(lines.nu[line].max()),
                     int(line.sum()), int((lines.iflg > 0).sum()), 0, (lines.n + NREC - 1) // NREC, 0)
    b += b"HID1    HID2    "
    return b


def write_tape3(path, lines: LineList, panel_sizes=None, negepp=False):
    """Writes ``lines`` as a TAPE3 file.  Panels take the list's entries in order,
    250 at a time, or ``panel_sizes`` entries each (every panel at most 250; a coupled
    line and its coefficient record must share a panel).  ``negepp`` writes the second
    header record (negative-E'' counts) that a '^' in HLINID(7) announces."""
    n = lines.n
    sizes = list(panel_sizes) if panel_sizes is not None else [NREC] * ((n + NREC - 1) // NREC)
    if sum(sizes) < n or max(sizes) > NREC:
        raise ValueError("panel sizes must cover the list with panels of at most 250 entries")
    brd = lines.brd_flg is not None
    with open(path, "wb") as f:
        _record(f, _header(lines, negepp))
        if negepp:
            _record(f, np.zeros(64, np.int32).tobytes() * 2 + np.zeros(4096, np.float32).tobytes())
        a = 0
        for k in sizes:
            b = min(a + k, n)
            if b <= a:
                break
            sl = slice(a, b)
            cnt = b - a
            line = lines.iflg[sl] >= 0
            vv = lines.nu[sl][line] if line.any() else lines.nu[sl]
            _record(f, struct.pack("<ddii", float(vv.min()), float(vv.max()), cnt, NWDS))

            def pad(x, dt, width=1):
                out = np.zeros((NREC, width) if width > 1 else NREC, dt)
                out[:cnt] = np.asarray(x, dt).reshape(out[:cnt].shape)
                return out.tobytes()

            blk = pad(lines.nu[sl], np.float64) + pad(lines.s[sl], np.float32) + pad(lines.a_air[sl], np.float32)
            blk += pad(lines.e_lo[sl], np.float32) + pad(lines.mol[sl], np.int32) + pad(lines.a_self[sl], np.float32)
            blk += pad(lines.tmpalf[sl], np.float32) + pad(lines.shift[sl], np.float32) + pad(lines.iflg[sl], np.int32)
            if brd:
                blk += pad(lines.brd_flg[sl], np.int32, 7)
                dat = np.stack([lines.brd_hw[sl], lines.brd_tmp[sl], lines.brd_shft[sl]], axis=2).reshape(cnt, 21)
                blk += pad(dat, np.float32, 21)
            else:
                blk += np.zeros(NREC * 7, np.int32).tobytes() + np.zeros(NREC * 21, np.float32).tobytes()
            blk += np.zeros(NREC, np.float32).tobytes()                 # SPEED_DEP
            assert len(blk) == 4 * NWDS
            _record(f, blk)
            a = b


def read_tape3(path, vlo=1.0, vhi=0.0) -> LineList:
    """The records of the panels overlapping [vlo, vhi] (every panel when vlo > vhi)
    as clblm_tape3_read returns them: 250 entries per panel, short panels padded with
    entries that are not lines (IFLG -1)."""
    lib = _lib.load()
    n = ctypes.c_int64()
    bp = str(path).encode()
    _lib.check(lib.clblm_tape3_read(bp, vlo, vhi, ctypes.byref(n), *([None] * 13)))
    m = n.value
    arr = dict(nu=np.zeros(m), s=np.zeros(m, np.float32), a_air=np.zeros(m, np.float32),
               a_self=np.zeros(m, np.float32), e_lo=np.zeros(m, np.float32), tmpalf=np.zeros(m, np.float32),
               shift=np.zeros(m, np.float32), mol=np.zeros(m, np.int32), iflg=np.zeros(m, np.int32),
               brd_flg=np.zeros((m, 7), np.int32), brd_hw=np.zeros((m, 7), np.float32),
               brd_tmp=np.zeros((m, 7), np.float32), brd_shft=np.zeros((m, 7), np.float32))
    order = ("nu", "s", "a_air", "a_self", "e_lo", "tmpalf", "shift", "mol", "iflg", "brd_flg", "brd_hw", "brd_tmp",
             "brd_shft")
    _lib.check(lib.clblm_tape3_read(bp, vlo, vhi, ctypes.byref(n), *[arr[k].ctypes.data for k in order]))
    return LineList(**arr)
