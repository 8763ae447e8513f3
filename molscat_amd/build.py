"""Build the HIP library (libmcprop.so) for gfx950 in-tree with hipcc."""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmcprop.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 instructions stay off
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build_lib(verbose: bool = False) -> str:
    src = os.path.join(CSRC, "mcp_api.hip")
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", tmp, src]
    if verbose:
        print(" ".join(cmd))
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


FORTRAN = os.path.join(HERE, "fortran")
FORTRAN_BUILD = os.path.join(FORTRAN, "build")
FORTRAN_HOST = os.path.join(FORTRAN_BUILD, "fortran_host")


def amdflang() -> str:
    for cand in (os.environ.get("FC"), "/opt/rocm/bin/amdflang", shutil.which("amdflang")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("amdflang not found")


def build_fortran(verbose: bool = False) -> str:
    """Compile the ISO_C_BINDING module (fortran/mcprop.f90) and the drop-in
    propagators (fortran/mcprop_dropin.f90) with amdflang, and link them with
    the test host program (tests/native/fortran_host.f90, which supplies the
    host-side module `potential`, POTENL, SCAPOT, PERTRB and DRCALC) against
    libmcprop.so.  Returns the path of the linked host program."""
    fc = amdflang()
    os.makedirs(FORTRAN_BUILD, exist_ok=True)
    host_src = os.path.join(os.path.dirname(HERE), "tests", "native", "fortran_host.f90")
    flags = ["-O2", "-fPIC", "-J", FORTRAN_BUILD]
    objs = []
    for src in (os.path.join(FORTRAN, "mcprop.f90"), host_src, os.path.join(FORTRAN, "mcprop_dropin.f90")):
        obj = os.path.join(FORTRAN_BUILD, os.path.splitext(os.path.basename(src))[0] + ".o")
        cmd = [fc] + flags + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd))
        subprocess.check_call(cmd)
        objs.append(obj)
    tmp = FORTRAN_HOST + ".tmp"
    cmd = [fc, "-O2"] + objs + ["-o", tmp, "-L", HERE, "-lmcprop", "-Wl,-rpath,$ORIGIN/../.."]
    if verbose:
        print(" ".join(cmd))
    subprocess.check_call(cmd)
    os.replace(tmp, FORTRAN_HOST)
    return FORTRAN_HOST


if __name__ == "__main__":
    print(build_lib(verbose=True))
    print(build_fortran(verbose=True))
