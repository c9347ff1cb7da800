"""Build libwos_hip.so for gfx950 in-tree (tetsuhiro-wos_amd/libwos_hip.so).

Every translation unit is compiled with hipcc --offload-arch=gfx950. Device
code keeps hipcc's default FMA contraction except where a kernel disables
it locally (#pragma clang fp contract(off) around the FP64 phase arguments);
host code is compiled with -ffp-contract=off so transfer-function scalars
computed on the host round exactly like the reference's amd64 build.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libwos_hip.so")
OBJ = os.path.join(HERE, "_build")
SOURCES = ["wos_hip.hip", "wos_launch_f64.hip", "wos_launch_f32.hip", "wos_launch8_f64.hip", "wos_launch8_f32.hip",
           "wos_mixed_f64.hip", "wos_mixed_f32.hip", "wos_persist.hip"]
# The persistent transfer kernel runs three task bodies in one loop: without
# machine-code loop-invariant motion the constants they materialise stay
# inside the tasks instead of occupying registers across the whole loop.
EXTRA = {"wos_persist.hip": ["-mllvm", "-disable-machine-licm"]}
ARCH = os.environ.get("WOS_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Xarch_host", "-ffp-contract=off",
    "-I", os.path.join(HERE, "..", "include"),
]


def _compile(src):
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, *EXTRA.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{err}")
    return obj


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


def build_oracle(verbose=True):
    oracle_dir = os.path.join(HERE, "..", "oracle")
    r = subprocess.run(["make", "-s", "-C", oracle_dir, "-B", "liboracle.so"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print("built oracle/liboracle.so")


if __name__ == "__main__":
    build()
    if "--no-oracle" not in sys.argv:
        build_oracle()
