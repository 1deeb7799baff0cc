"""Build recipe for the CPU oracle (test infrastructure only).

Compiles oracle/hcore_oracle.c with gcc into oracle/_build/liboracle.so,
linked to the LAPACK/BLAS bundled in this image's scipy wheel
(scipy-openblas, Fortran symbols prefixed "scipy_").  The reference's own
sources are NOT compiled: they need <cblas.h>/<lapacke.h>, which the image
does not ship.
"""
import glob
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "liboracle.so")


def openblas_path():
    import scipy

    libs = glob.glob(os.path.join(os.path.dirname(os.path.dirname(scipy.__file__)),
                                  "scipy.libs", "libscipy_openblas*.so"))
    if not libs:
        raise RuntimeError("scipy-openblas not found next to the scipy package")
    return libs[0]


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    blas = openblas_path()
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-std=gnu11",
           os.path.join(HERE, "hcore_oracle.c"), "-o", LIB + ".tmp",
           blas, "-Wl,-rpath," + os.path.dirname(blas), "-lm"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
