"""Builds the example user-PDE library of the registration tests:
tests/user_pde/_build/libuser_pdes.so from user_pdes.hip, compiled for
gfx950 against include/exahype_amd_pde.h and linked to the library (which
must be built first)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build", "libuser_pdes.so")


def build(verbose=False):
    from exahype_amd import build as hipbuild
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    libdir = os.path.dirname(hipbuild.LIB)
    rel = os.path.relpath(libdir, os.path.dirname(OUT))
    cmd = [hipbuild.hipcc()] + hipbuild.HIPCC_FLAGS + [
        "-shared", "-I", os.path.join(ROOT, "include"), "-I", hipbuild.CSRC, os.path.join(HERE, "user_pdes.hip"),
        "-L", libdir, "-lexahype_amd", "-Wl,-rpath,$ORIGIN/" + rel, "-o", OUT + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
