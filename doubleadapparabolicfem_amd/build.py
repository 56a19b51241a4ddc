"""Build the HIP library libapf.so in-tree for gfx950.

    python -m doubleadapparabolicfem_amd.build

Every HIP translation unit is compiled with packed FP32 instructions disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libapf.so")
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("APF_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-result"]
HIP_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics"]

SOURCES = ["apf_tables.cpp", "apf_space.cpp", "apf_mg.cpp", "apf_marking.cpp", "apf_mesh1d.cpp", "apf_gpu.hip", "apf_estimator.hip", "apf_g.hip",
           "apf_mg.hip", "apf_csr.hip", "apf_general.hip"]


def _compile(src):
    obj = os.path.join(BUILD, src.replace(".", "_") + ".o")
    cmd = [HIPCC] + COMMON + ["-I", CSRC]
    if src.endswith(".hip"):
        cmd += ["-x", "hip"] + HIP_FLAGS
    cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if r.stderr.strip():
        sys.stderr.write(r.stderr)
    return obj


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    build()
