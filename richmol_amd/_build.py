"""Compile the HIP library for gfx950 in-tree (hipcc, no JIT cache)."""
import os
import subprocess

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_NAME = "librichmol_amd.so"
LIB_PATH = os.path.join(PKG_DIR, LIB_NAME)
SOURCES = [os.path.join(PKG_DIR, "csrc", "rm_tdse.hip"),
           os.path.join(PKG_DIR, "csrc", "rm_subspace.hip"),
           os.path.join(PKG_DIR, "csrc", "rm_coop.hip"),
           os.path.join(PKG_DIR, "csrc", "rm_kron.hip"),
           os.path.join(PKG_DIR, "csrc", "rm_expokit.hip"),
           os.path.join(PKG_DIR, "csrc", "rm_rotdens.hip")]

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    "--offload-arch=gfx950",
    # no packed-FP32 instructions in any generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(REPO_DIR, "include"),
]
# dense products of the Kronecker-block operator
LINK_FLAGS = ["-L/opt/rocm/lib", "-lrocblas"]


def hipcc():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return os.path.join(rocm, "bin", "hipcc")


def build(verbose=False):
    """Compile SOURCES into LIB_PATH (always recompiles)."""
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", tmp, *SOURCES, *LINK_FLAGS]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stderr[-4000:]}")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
