"""The Fortran side: the ISO_C_BINDING module and the drop-in Module_ODLAY compile, the
drop-in against the reference's own Module_ConstParam/Module_Config; the Fortran
callers reproduce the Python/ctypes results on the GPU, the single-layer ODLAY and
the batched OD + radiance/Jacobian chain on device arrays."""
import os
import struct
import subprocess
import tempfile

import numpy as np
import pytest

from clblm_amd import build as B
from clblm_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFMOD = os.path.join(ROOT, "oracle", "_ref")


@pytest.mark.skipif(not os.path.exists(B.FLANG), reason="flang not installed")
def test_shim_compiles():
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run([B.FLANG, "-J", d, "-c", os.path.join(B.FDIR, "clblm_hip.f90"), "-o",
                            os.path.join(d, "a.o")], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


@pytest.mark.skipif(not os.path.exists(os.path.join(REFMOD, "module_config.mod")),
                    reason="reference modules not built (oracle/_ref)")
def test_dropin_module_compiles_against_reference_modules():
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run([B.FLANG, "-J", d, "-c", os.path.join(B.FDIR, "clblm_hip.f90"), "-o",
                            os.path.join(d, "a.o")], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        # ODLAY_struct needs Module_AtmPath/Module_DV, which USE the netCDF scene module
        # and do not build here; the non-struct form is compiled against the reference
        r = subprocess.run([B.FLANG, "-fdefault-real-8", "-cpp", "-DCLBLM_HIP_NO_STRUCT", "-I", d, "-I", REFMOD,
                            "-J", d, "-c",
                            os.path.join(B.FDIR, "module_odlay_hip.f90"), "-o", os.path.join(d, "b.o")],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_fortran_caller_matches_ctypes():
    exe = os.path.join(B.FBIN, "test_odlay_f")
    assert os.path.exists(exe), "build() compiles the Fortran caller"
    from clblm_amd.odlay import Context
    lines = W.profile_lines(8000, 500.0, 800.0, seed=1001)
    lay = W.atmosphere()[7]
    dv = W.calculate_dv(650.0, lay.pave, lay.tave, lay.w[1] / lay.wtot)
    mols = sorted(lay.w)
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<ii", lines.n, len(mols)))
            for a in (lines.nu, lines.s, lines.a_air, lines.a_self, lines.e_lo, lines.tmpalf, lines.shift,
                      lines.mol, lines.iflg):
                f.write(np.ascontiguousarray(a).tobytes())
            f.write(np.array(mols, np.int32).tobytes())
            f.write(np.array([lay.w[m] for m in mols], np.float64).tobytes())
            f.write(np.array([lay.pave, lay.tave, lay.wtot, 600.0, 700.0, dv, dv, 4.0, 0.0]).tobytes())
        r = subprocess.run([exe, fin, fout], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        raw = open(fout, "rb").read()
        n, dvu, jrad = struct.unpack("<idi", raw[:16])
        od_f = np.frombuffer(raw[16:16 + 8 * n], np.float64)
    ctx = Context(0)
    h = ctx.upload_lines(lines)
    od, dvp = ctx.odlay(h, lay, 600.0, 700.0, dv)
    ctx.close()
    assert dvu == dvp and jrad == ctx.jrad == 0
    np.testing.assert_allclose(od_f, od, rtol=1e-13, atol=0)


@pytest.mark.gpu
def test_fortran_batched_chain_matches_ctypes():
    """c_od_batch and c_rt_jac called from Fortran on device arrays (clblm_hip_dev_*):
    the radiance and dR/dtau of the Python route, bit for bit (same kernels, same
    inputs; the OD pass is run once per route and its atomic sums are reordered, so
    1e-13)."""
    import torch
    exe = os.path.join(B.FBIN, "test_jac_f")
    assert os.path.exists(exe), "build() compiles the Fortran caller"
    from clblm_amd.odlay import Context
    lines = W.profile_lines(6000, 500.0, 800.0, seed=1001)
    layers = W.atmosphere(6, 1.0)
    v1, v2 = 640.0, 652.0
    dvs = [W.calculate_dv(650.0, l.pave, l.tave, l.w[1] / l.wtot) for l in layers]
    dvout = min(dvs)
    t_lev = np.array([layers[0].tave + 2.0] + [0.5 * (a.tave + b.tave) for a, b in zip(layers[:-1], layers[1:])] +
                     [layers[-1].tave - 1.0])
    up, lo = np.full(6, 0.45), np.full(6, 0.55)
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<ii", lines.n, len(layers)))
            for a in (lines.nu, lines.s, lines.a_air, lines.a_self, lines.e_lo, lines.tmpalf, lines.shift,
                      lines.mol, lines.iflg):
                f.write(np.ascontiguousarray(a).tobytes())
            for l in layers:
                f.write(np.array([l.pave, l.tave, l.wtot] + [l.w.get(m, 0.0) for m in range(1, 8)]).tobytes())
            f.write(np.array([v1, v2, dvout], np.float64).tobytes())
            for a in (dvs, t_lev, up, lo):
                f.write(np.asarray(a, np.float64).tobytes())
            f.write(np.array([0.98, 295.0, 1.0], np.float64).tobytes())
            f.write(struct.pack("<i", 1))
        r = subprocess.run([exe, fin, fout], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        raw = open(fout, "rb").read()
        n = struct.unpack("<q", raw[:8])[0]
        vals = np.frombuffer(raw[8:], np.float64)
        rad_f, dtau_f = vals[:n], vals[n:n + n * len(layers)].reshape(len(layers), n)
    import ctypes
    ctx = Context(0)
    h = ctx.upload_lines(lines)
    pars = [Context.params(v1, v2, dv, dvout=dvout) for dv in dvs]
    ctx.jrad = 1
    od, off = ctx.od_batch(h, layers, pars)
    L = len(layers)
    e = lambda *s: torch.empty(*s, dtype=torch.float64, device="cuda")
    rad, dtau, dtlev, dts, de = e(n), e(L, n), e(L + 1, n), e(n), e(n)
    t = lambda a: torch.tensor(np.asarray(a, np.float64), device="cuda")
    tl, tv, tu, to = t([l.tave for l in layers]), t(t_lev), t(up), t(lo)
    p = lambda x: ctypes.c_void_p(x.data_ptr())
    from clblm_amd._lib import check
    check(ctx.lib.clblm_rt_jac(ctx.ptr, L, n, ctypes.c_double(v1), ctypes.c_double(dvout), p(od), p(tl), p(tv), p(tu),
                               p(to), ctypes.c_double(0.98), ctypes.c_double(295.0), 1, ctypes.c_double(1.0), p(rad),
                               p(dtau), p(dtlev), p(dts), p(de), None))
    torch.cuda.synchronize()
    r, dt_ = rad.cpu().numpy(), dtau.cpu().numpy()
    ctx.close()
    np.testing.assert_allclose(rad_f, r, rtol=1e-13, atol=0)
    np.testing.assert_allclose(dtau_f, dt_, rtol=1e-12, atol=1e-13 * np.abs(dt_).max())


DROPIN = os.path.join(REFMOD, "dropin_caller")


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(DROPIN), reason="drop-in caller not built (oracle/_ref)")
@pytest.mark.parametrize("cont,brd", [(0, 0), (1, 1)])
def test_dropin_odlay_called_as_the_reference_calls_it(cont, brd):
    """A Fortran program calling the reference's generic ODLAY with the reference's own
    CLBLM_OD_Ctrl and ioFiles (Module_Config, Module_ConstParam compiled from the
    reference's sources), against the drop-in Module_ODLAY: the line list comes from the
    TAPE3 file named by ioFiles%lineFile, speciesBroad and the continua (MT_CKD from a
    netCDF file at the reference's data path, CO2 and the others) follow the control
    record, and the optical depths equal the Python route's (same library, 1e-13)."""
    from scipy.io import netcdf_file
    from clblm_amd.continuum import load_mtckd
    from clblm_amd.odlay import Context
    from clblm_amd.tape3 import write_tape3
    lines = W.profile_lines(8000, 500.0, 800.0, seed=1001)
    rng = np.random.default_rng(3)
    lines.brd_flg = (rng.uniform(size=(lines.n, 7)) < 0.3).astype(np.int32)
    lines.brd_hw = rng.uniform(0.02, 0.1, (lines.n, 7)).astype(np.float32)
    lines.brd_tmp = rng.uniform(0.5, 0.8, (lines.n, 7)).astype(np.float32)
    lines.brd_shft = rng.uniform(-0.01, 0.01, (lines.n, 7)).astype(np.float32)
    # the synthetic O2 lines' self widths (5x air) make RDLIN's O2 air-width rescale
    # negative (clblm_ODLAY.f90:2728-2738); species weights on them would then give a
    # negative Lorentz width, which the library refuses: they keep air/self broadening
    lines.brd_flg[lines.mol % 100 == 7] = 0
    layers = W.atmosphere(60, 0.1)[::20]
    dvs = [W.calculate_dv(650.0, l.pave, l.tave, l.w[1] / l.wtot) for l in layers]
    table = load_mtckd(os.path.join(ROOT, "tests", "golden", "mtckd_h2o.npz"))
    with tempfile.TemporaryDirectory() as d:
        tape3 = os.path.join(d, "TAPE3")
        write_tape3(tape3, lines)
        os.makedirs(os.path.join(d, "clblm_data", "spectroscopy"))
        with netcdf_file(os.path.join(d, "clblm_data", "spectroscopy", "absco-ref_wv-mt-ckd.nc"), "w") as f:
            f.createDimension("wavenumbers", table["wavenumbers"].size)
            for k in ("wavenumbers", "self_absco_ref", "for_absco_ref", "self_texp"):
                f.createVariable(k, "d", ("wavenumbers",))[:] = table[k]
            for k in ("ref_temp", "ref_press"):
                f.createVariable(k, "d", ())[...] = float(table[k])
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(tape3.encode().ljust(256))
            f.write(struct.pack("<iii", len(layers), cont, brd))
            f.write(np.array([600.0, 700.0], np.float64).tobytes())
            for l, dv in zip(layers, dvs):
                f.write(np.array([l.pave, l.tave, l.wtot] + [l.w.get(m, 0.0) for m in range(1, 8)] + [dv],
                                 np.float64).tobytes())
        r = subprocess.run([DROPIN, fin, fout], capture_output=True, text=True, timeout=120, cwd=d)
        assert r.returncode == 0, r.stdout + r.stderr
        raw = open(fout, "rb").read()
        ctx = Context(0)
        h = ctx.load_tape3(tape3)
    if cont:
        ctx.mtckd_load(table)
    ctx.jrad = 1
    pos = 0
    for l, dv in zip(layers, dvs):
        n, dvo = struct.unpack("<id", raw[pos:pos + 12])
        pos += 12
        od_f = np.frombuffer(raw[pos:pos + 8 * n], np.float64)
        pos += 8 * n
        od, dvp = ctx.odlay(h, l, 600.0, 700.0, dv, mtckd=bool(cont), other_continua=bool(cont),
                            species_broad=bool(brd))
        assert n == od.size and dvo == dvp
        np.testing.assert_allclose(od_f, od, rtol=1e-13, atol=0)
    ctx.close()
