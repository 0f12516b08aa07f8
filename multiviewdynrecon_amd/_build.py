"""Build recipe for the native libraries (called by __graft_entry__.build()).

    libmvdr_solver.so  HIP kernels + C ABI, hipcc for gfx950
    libmvdr_scene.so   synthetic scene generator, g++
    oracle/libmvdr_oracle.so   CPU oracle used by tests / the CPU baseline, g++

All outputs land in-tree next to their sources.  Each call recompiles.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MVDR_OFFLOAD_ARCH", "gfx950")
# packed FP32 instructions disabled for device code (v_pk_*_f32 forms)
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wall", "-Wno-unused-function", "-I", INCLUDE]
HIP_SOURCES = ["kernels_index.hip", "kernels_solve.hip", "kernels_pcg.hip", "kernels_mf.hip", "crossview.hip",
               "warpfield.hip", "render.hip", "rigid.hip",
               "solver.cpp", "tuning.cpp"]


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError("build failed: " + " ".join(cmd))
    return proc.stdout


def build_solver(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []

    def compile_one(src):
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        _run([HIPCC, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj] + HIP_FLAGS)
        return obj

    with ThreadPoolExecutor(max_workers=len(HIP_SOURCES)) as ex:
        objs = list(ex.map(compile_one, HIP_SOURCES))
    out = os.path.join(HERE, "libmvdr_solver.so")
    tmp = out + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build_scene(verbose=False):
    out = os.path.join(HERE, "libmvdr_scene.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", INCLUDE, "-o", tmp,
          os.path.join(CSRC, "scene", "scene.cpp")])
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build_oracle(verbose=False):
    odir = os.path.join(ROOT, "oracle")
    out = os.path.join(odir, "libmvdr_oracle.so")
    tmp = out + ".tmp"
    # no FMA contraction: selection thresholds round like the device's
    _run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-shared", "-fPIC", "-I", INCLUDE, "-o", tmp,
          os.path.join(odir, "mvdr_oracle.cpp")])
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build_all(verbose=False):
    with ThreadPoolExecutor(max_workers=3) as ex:
        futs = [ex.submit(f, verbose) for f in (build_solver, build_scene, build_oracle)]
        return [f.result() for f in futs]


if __name__ == "__main__":
    build_all(verbose=True)
