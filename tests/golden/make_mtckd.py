"""Writes tests/golden/mtckd_h2o.npz: the MT_CKD water-vapour continuum coefficients
the reference reads at run time (clblm_data/spectroscopy/absco-ref_wv-mt-ckd.nc,
mt_ckd_h2o_module.f90:16), as data for the continuum tests on machines without the
reference.  The file's title and copyright notice travel with the arrays, as the
notice asks.  Run here, where the reference's data directory is present."""
import os
import sys

import numpy as np

SRC = "/root/reference/clblm_data/spectroscopy/absco-ref_wv-mt-ckd.nc"
HERE = os.path.dirname(os.path.abspath(__file__))


def main(src=SRC):
    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    from clblm_amd.continuum import load_mtckd
    t = load_mtckd(src)
    np.savez(os.path.join(HERE, "mtckd_h2o.npz"), **t)


if __name__ == "__main__":
    main(*sys.argv[1:])
