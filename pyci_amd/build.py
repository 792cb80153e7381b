"""In-tree build of the native parts of pyci_amd.

  libpyci_engine.so   HIP kernels + C ABI (include/pyci_engine.h), hipcc for gfx950
  _pyci*.so           pybind11 host module (wave functions, integrals, operator facade),
                      g++, linked against libpyci_engine.so with an $ORIGIN rpath

Run `python -m pyci_amd.build` (or __graft_entry__.build()). Objects go to
pyci_amd/_build/ (git-ignored); the two shared libraries land in pyci_amd/.
"""
import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ENGINE_SRC = os.path.join(HERE, "csrc", "engine")
BINDING_SRC = os.path.join(HERE, "csrc", "binding")
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PYCI_AMD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
    "-Wall", "-Wno-unused-result", "-Wno-unused-variable",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def engine_sources():
    return sorted(os.path.join(ENGINE_SRC, f) for f in os.listdir(ENGINE_SRC) if f.endswith(".hip"))


def build_engine(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    objs = []

    def compile_one(src):
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", src, "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, engine_sources()))
    out = os.path.join(HERE, "libpyci_engine.so")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, out)
    return out


def build_binding():
    import pybind11

    srcs = sorted(os.path.join(BINDING_SRC, f) for f in os.listdir(BINDING_SRC) if f.endswith(".cpp"))
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(HERE, "_pyci" + suffix)
    tmp = out + ".tmp"
    cmd = [
        "g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden",
        "-I", pybind11.get_include(), "-I", sysconfig.get_paths()["include"],
        "-I", os.path.join(ROOT, "include"), "-I", BINDING_SRC,
        *srcs, "-o", tmp,
        "-L", HERE, "-lpyci_engine", "-Wl,-rpath,$ORIGIN",
    ]
    _run(cmd)
    os.replace(tmp, out)
    return out


def build(verbose=True):
    e = build_engine()
    b = build_binding()
    if verbose:
        print("built", e)
        print("built", b)
    return e, b


if __name__ == "__main__":
    build()
    sys.exit(0)
