"""Builds libclblm_hip.so in-tree with hipcc for gfx950."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRCS = [os.path.join(HERE, "csrc", f) for f in ("clblm_od.hip", "clblm_rt.hip", "clblm_scan.hip", "clblm_tape3.cpp", "clblm_setdv.cpp",
                                                        "clblm_netcdf3.cpp")]
OUT = os.path.join(HERE, "libclblm_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         # FP64 parity: every expression rounds as written, no FMA contraction
         "-ffp-contract=off",
         # no packed FP32 instructions (v_pk_*_f32)
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


# hipFFT for the instrument scan; PyTorch-ROCm loads a copy with the same soname
# (libhipfft.so.0) at import, so both resolve to one library and one HIP runtime
LIBS = ["-L/opt/rocm/lib", "-lhipfft"]


def build_hip(verbose=False):
    cmd = [HIPCC] + FLAGS + ["-o", OUT] + SRCS + LIBS
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    if verbose:
        print(" ".join(cmd))
    return OUT


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FDIR = os.path.join(HERE, "fortran")
FBIN = os.path.join(FDIR, "_bin")


def build_fortran():
    """The ISO_C_BINDING module and the Fortran test callers (skipped without flang)."""
    if not os.path.exists(FLANG):
        return None
    os.makedirs(FBIN, exist_ok=True)
    steps = [
        [FLANG, "-O2", "-J", FBIN, "-c", os.path.join(FDIR, "clblm_hip.f90"), "-o", os.path.join(FBIN, "clblm_hip.o")],
        [FLANG, "-O2", "-J", FBIN, "-o", os.path.join(FBIN, "test_odlay_f"), os.path.join(FDIR, "test_odlay_f.f90"),
         os.path.join(FBIN, "clblm_hip.o"), "-L" + HERE, "-lclblm_hip", "-Wl,-rpath," + HERE],
        [FLANG, "-O2", "-J", FBIN, "-o", os.path.join(FBIN, "test_jac_f"), os.path.join(FDIR, "test_jac_f.f90"),
         os.path.join(FBIN, "clblm_hip.o"), "-L" + HERE, "-lclblm_hip", "-Wl,-rpath," + HERE],
    ]
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("flang failed:\n" + r.stdout + r.stderr)
    return os.path.join(FBIN, "test_odlay_f")


if __name__ == "__main__":
    print(build_hip(verbose=True))
