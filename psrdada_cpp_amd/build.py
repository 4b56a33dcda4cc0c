"""Build recipe for the native library ``libfbfuse_hip.so`` (gfx950).

Every HIP/C++ source under ``csrc/`` is compiled with hipcc for gfx950 into
``_build/`` (git-ignored) and linked into ``psrdada_cpp_amd/libfbfuse_hip.so``,
in-tree, so that the built library travels with the source tree.  The C++
operator layer (``cpp/``) and its test driver are built by ``build_cpp``.

Run as ``python -m psrdada_cpp_amd.build`` or through ``__graft_entry__.build()``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
CPP = os.path.join(PKG, "cpp")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libfbfuse_hip.so")
CPP_LIB = os.path.join(PKG, "libfbfuse_cpp.so")
BIN = os.path.join(PKG, "bin")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are disabled for all device code.  MFMA results
# go to ordinary VGPRs (gfx950 reads them directly), which saves one
# v_accvgpr_read per accumulator value in the detection loops.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wall", "-Wno-unused-result",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    # hipcc's host pass warns that it does not know the packed-fp32 feature.
    noise = "'-packed-fp32-ops' is not a recognized feature"
    errs = "\n".join(l for l in res.stderr.splitlines() if noise not in l)
    if verbose and errs.strip():
        print(errs, file=sys.stderr)
    return res


def _sources(folder, exts):
    return sorted(os.path.join(folder, f) for f in os.listdir(folder) if f.endswith(exts))


def build_lib(verbose=False, jobs=8):
    """Compile csrc/*.hip and csrc/*.cpp for gfx950 and link libfbfuse_hip.so."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = _sources(CSRC, (".hip", ".cpp"))
    objs = []
    cmds = []
    for s in srcs:
        o = os.path.join(BUILD, os.path.basename(s) + ".o")
        objs.append(o)
        lang = ["-x", "hip"]
        cmds.append([HIPCC, *HIP_FLAGS, *lang, "-I", os.path.join(ROOT, "include"), "-c", s, "-o", o])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs,
          "-Wl,-soname,libfbfuse_hip.so"], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_cpp(verbose=False, jobs=8):
    """Compile the C++ operator layer (cpp/src) into libfbfuse_cpp.so and the
    C++ test driver into tests/cpp/fbfuse_cpp_tests."""
    if not os.path.isdir(CPP):
        return None
    os.makedirs(BUILD, exist_ok=True)
    inc = ["-I", os.path.join(CPP, "include"), "-I", os.path.join(ROOT, "include")]
    cxx = ["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter", "-pthread"]
    srcs = _sources(os.path.join(CPP, "src"), (".cpp",))
    objs = []
    cmds = []
    for s in srcs:
        o = os.path.join(BUILD, "cpp_" + os.path.basename(s) + ".o")
        objs.append(o)
        cmds.append([*cxx, *inc, "-c", s, "-o", o])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = CPP_LIB + ".tmp"
    _run([*cxx, "-shared", "-o", tmp, *objs, "-L", PKG, "-lfbfuse_hip",
          "-Wl,-rpath,$ORIGIN", "-Wl,-soname,libfbfuse_cpp.so"], verbose)
    os.replace(tmp, CPP_LIB)
    test_dir = os.path.join(ROOT, "tests", "cpp")
    test_srcs = _sources(test_dir, (".cpp",)) if os.path.isdir(test_dir) else []
    outs = []
    for s in test_srcs:
        exe = os.path.join(test_dir, os.path.splitext(os.path.basename(s))[0])
        tmpx = exe + ".tmp"
        _run([*cxx, *inc, "-I", os.path.join(ROOT, "oracle"), s, "-o", tmpx,
              "-L", PKG, "-lfbfuse_cpp", "-lfbfuse_hip",
              "-L", os.path.join(ROOT, "oracle"), "-lfbf_oracle",
              f"-Wl,-rpath,{PKG}", f"-Wl,-rpath,{os.path.join(ROOT, 'oracle')}",
              "-Wl,-rpath,$ORIGIN/../../psrdada_cpp_amd", "-Wl,-rpath,$ORIGIN/../../oracle"], verbose)
        os.replace(tmpx, exe)
        outs.append(exe)
    # Command-line tools over the C++ layer (cpp/tools/<name>.cpp ->
    # bin/fbfuse_<name>), e.g. the Pipeline benchmark that bench.py runs.
    tool_dir = os.path.join(CPP, "tools")
    if os.path.isdir(tool_dir):
        os.makedirs(BIN, exist_ok=True)
        for s in _sources(tool_dir, (".cpp",)):
            exe = os.path.join(BIN, "fbfuse_" + os.path.splitext(os.path.basename(s))[0])
            tmpx = exe + ".tmp"
            _run([*cxx, *inc, s, "-o", tmpx, "-L", PKG, "-lfbfuse_cpp", "-lfbfuse_hip",
                  "-Wl,-rpath,$ORIGIN/.."], verbose)
            os.replace(tmpx, exe)
            outs.append(exe)
    return CPP_LIB, outs


def build_oracle(verbose=False):
    """Build the CPU oracle (test infrastructure) with its Makefile."""
    _run(["make", "-B", "-C", os.path.join(ROOT, "oracle")], verbose)
    return os.path.join(ROOT, "oracle", "libfbf_oracle.so")


def build_all(verbose=False):
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    build_oracle(verbose)
    build_lib(verbose)
    build_cpp(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print("built", LIB)
