"""Build recipe of the native libraries (in-tree, no JIT cache).

  python -m emfa_amd.build            -> emfa_amd/libemfa_hip.so, emfa_amd/libemfa_core.so, emfa_amd/bin/emfa

libemfa_hip.so   gfx950 kernels + the C ABI of include/emfa_hip.h
libemfa_core.so  C++ host: config loader, mesh readers, solver classes
emfa             launcher (reads the CONFIG.yml path from stdin)
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIP_DIR = os.path.join(HERE, "csrc", "hip")
HOST_DIR = os.path.join(HERE, "csrc", "host")
OBJ_DIR = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 VALU ops stay disabled on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-command-line-argument",
]
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", "-Wno-sign-compare",
              "-I" + os.path.join(ROOT, "include"), "-I" + HOST_DIR]

LIB_HIP = os.path.join(HERE, "libemfa_hip.so")
LIB_CORE = os.path.join(HERE, "libemfa_core.so")
BIN_DIR = os.path.join(HERE, "bin")
EXE = os.path.join(BIN_DIR, "emfa")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise RuntimeError("build step failed: " + os.path.basename(cmd[-1]))
    return r.stdout


def _sources(d, exts):
    if not os.path.isdir(d):
        return []
    return sorted(os.path.join(d, f) for f in os.listdir(d) if os.path.splitext(f)[1] in exts)


def build_hip(jobs=8):
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = _sources(HIP_DIR, {".hip", ".cpp"})
    objs = []
    cmds = []
    for s in srcs:
        o = os.path.join(OBJ_DIR, "hip_" + os.path.basename(s) + ".o")
        objs.append(o)
        lang = ["-x", "hip"] if s.endswith(".hip") else []
        cmds.append([HIPCC] + HIP_FLAGS + lang + ["-c", "-o", o, s])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", LIB_HIP] + objs)
    return LIB_HIP


def build_core(jobs=8):
    srcs = _sources(HOST_DIR, {".cpp"})
    if not srcs:
        return None
    os.makedirs(OBJ_DIR, exist_ok=True)
    lib_srcs = [s for s in srcs if os.path.basename(s) != "launcher.cpp"]
    objs, cmds = [], []
    for s in lib_srcs:
        o = os.path.join(OBJ_DIR, "core_" + os.path.basename(s) + ".o")
        objs.append(o)
        cmds.append(["g++"] + HOST_FLAGS + ["-c", "-o", o, s])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    _run(["g++", "-shared", "-o", LIB_CORE] + objs + ["-L" + HERE, "-lemfa_hip", "-Wl,-rpath,$ORIGIN"])
    launcher = os.path.join(HOST_DIR, "launcher.cpp")
    if os.path.exists(launcher):
        os.makedirs(BIN_DIR, exist_ok=True)
        _run(["g++"] + HOST_FLAGS + ["-o", EXE, launcher, "-L" + HERE, "-lemfa_core", "-lemfa_hip",
                                     "-Wl,-rpath,$ORIGIN/.."])
    return LIB_CORE


def build(jobs=8):
    build_hip(jobs)
    build_core(jobs)


if __name__ == "__main__":
    build()
    print("built", LIB_HIP, LIB_CORE if os.path.exists(LIB_CORE) else "")
