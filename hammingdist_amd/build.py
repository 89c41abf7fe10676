"""Builds the native parts of hammingdist_amd in-tree.

  libhd_engine.so      HIP engine (csrc/*.hip) for gfx950, C ABI of
                       include/hammingdist_amd.h
  hammingdist*.so      C++ host layer + pybind11 module `hammingdist`
                       (csrc/host/*.cc),
                       linked against libhd_engine.so

Run ``python -m hammingdist_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(CSRC, "host")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
ARCH = os.environ.get("HD_OFFLOAD_ARCH", "gfx950")

ENGINE_SOURCES = ["engine.hip", "batch.hip", "serve.hip"]
HOST_SOURCES = ["fasta.cc", "dataset.cc", "loaders.cc", "writer.cc", "synth.cc", "module.cc"]

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
# roctx ranges of the host-side phases (rocprofv3 --marker-trace)
ROCTX_LINK = [f"-L{ROCM}/lib", "-lrocprofiler-sdk-roctx", f"-Wl,-rpath,{ROCM}/lib"]

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    # No packed-FP32 instructions anywhere in this library.
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wall",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        line for line in proc.stdout.splitlines() if "packed-fp32-ops" not in line
    )
    if proc.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")


def engine_sources() -> list[str]:
    return [os.path.join(CSRC, s) for s in ENGINE_SOURCES if os.path.exists(os.path.join(CSRC, s))]


def build_engine() -> str:
    os.makedirs(BUILD, exist_ok=True)
    hipcc = _hipcc()
    objs = []
    jobs = []
    for src in engine_sources():
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        objs.append(obj)
        jobs.append([hipcc, *HIPCC_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", src, "-o", obj])
    with ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        list(ex.map(_run, jobs))
    out = os.path.join(HERE, "libhd_engine.so")
    tmp = out + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
          "-Wl,-soname,libhd_engine.so", *ROCTX_LINK])
    os.replace(tmp, out)
    return out


def build_probe() -> str:
    """hd_valu_probe: the VALU issue-rate probe bench.py runs for the
    roofline ceiling of the all-pairs kernels."""
    out = os.path.join(HERE, "hd_valu_probe")
    tmp = out + ".tmp"
    _run([_hipcc(), *HIPCC_FLAGS, os.path.join(CSRC, "valu_probe.hip"), "-o", tmp])
    os.replace(tmp, out)
    return out


def _module_name() -> str:
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return "hammingdist" + suffix


def build_host() -> str:
    import pybind11

    os.makedirs(BUILD, exist_ok=True)
    py_inc = sysconfig.get_paths()["include"]
    flags = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-pthread", f"-I{INCLUDE}", f"-I{HOST}",
             f"-I{ROCM}/include",
             f"-I{pybind11.get_include()}", f"-I{py_inc}", "-fvisibility=hidden"]
    # HD_HOST_SANITIZE=address,undefined builds the host layer (g++ only; the
    # HIP engine is not affected) with those sanitizers for a checking run
    # (python loaded with LD_PRELOAD of the matching libasan).
    sanitize = os.environ.get("HD_HOST_SANITIZE", "")
    san = [f"-fsanitize={sanitize}", "-fno-omit-frame-pointer", "-g"] if sanitize else []
    flags += san
    objs, jobs = [], []
    for name in HOST_SOURCES:
        src = os.path.join(HOST, name)
        obj = os.path.join(BUILD, name + ".o")
        objs.append(obj)
        jobs.append(["g++", *flags, "-c", src, "-o", obj])
    with ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        list(ex.map(_run, jobs))
    out = os.path.join(HERE, _module_name())
    tmp = out + ".tmp"
    _run(["g++", "-shared", "-pthread", *san, "-o", tmp, *objs, f"-L{HERE}", "-lhd_engine",
          "-Wl,-rpath,$ORIGIN", *ROCTX_LINK])
    os.replace(tmp, out)
    # the module's former name, from older builds of this tree
    for name in os.listdir(HERE):
        if name.startswith("_hammingdist") and name.endswith(".so"):
            os.remove(os.path.join(HERE, name))
    return out


def build() -> None:
    build_engine()
    build_host()
    build_probe()


if __name__ == "__main__":
    build()
    print("built", os.path.join(HERE, "libhd_engine.so"), os.path.join(HERE, _module_name()),
          file=sys.stderr)
