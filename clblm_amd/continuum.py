"""MT_CKD water-vapour continuum coefficients: the table mt_ckd_h2o_absco reads
(src/clblm_src/mt_ckd_h2o_module.f90:27-103, via read_module.f90's getData) from
absco-ref_wv-mt-ckd.nc, a netCDF-3 file: wavenumbers, self_absco_ref,
for_absco_ref, self_texp and the reference pressure and temperature.  The arrays
go to the library with Context.mtckd_load; the OD path then adds the continuum
itself (CLBLM_OD_MTCKD)."""
import numpy as np

FIELDS = ("wavenumbers", "self_absco_ref", "for_absco_ref", "self_texp", "ref_temp", "ref_press")


def load_mtckd(path):
    """The coefficient table from the reference's netCDF file, or from an .npz
    written from it; returns a dict of float64 arrays (scalars for ref_temp and
    ref_press) plus the file's title and copyright strings."""
    path = str(path)
    if path.endswith(".npz"):
        with np.load(path, allow_pickle=False) as z:
            return {k: z[k] for k in z.files}
    from scipy.io import netcdf_file
    with netcdf_file(path, "r", mmap=False) as f:
        out = {k: np.asarray(f.variables[k].data, dtype=np.float64).copy() for k in FIELDS}
        for k in ("title", "copyright"):
            out[k] = np.array(f._attributes.get(k, b"").decode("utf-8", "replace").strip())
    return out
