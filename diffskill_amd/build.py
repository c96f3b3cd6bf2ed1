"""In-tree build of libdiffskill_mpm.so for gfx950 (hipcc)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = [os.path.join(HERE, "csrc", "mpm_engine.hip"), os.path.join(HERE, "csrc", "pointcloud.hip")]
OUT = os.path.join(HERE, "libdiffskill_mpm.so")
ARCH = os.environ.get("DSK_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops: no v_pk_{add,mul,fma}_f32 generation.
# -munsafe-fp-atomics: hardware float atomics (global_atomic_add_f32) instead of CAS loops.
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=" + ARCH, "-munsafe-fp-atomics",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose=False):
    cmd = ["hipcc"] + FLAGS + ["-o", OUT + ".tmp"] + SRC
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libdiffskill_mpm.so")
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
