"""Builds libfilsim_amd.so in-tree with hipcc for gfx950.

    python -m filsim_amd.build
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libfilsim_amd.so")
OBJ = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Werror=format", "-Wno-unused-function",
         "-Wno-unused-variable"]


def sources():
    return sorted(f for f in os.listdir(CSRC)
                  if f.endswith(".hip") or f.endswith(".cpp"))


def _compile(src):
    obj = os.path.join(OBJ, src + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        for src, obj, r in ex.map(_compile, sources()):
            msgs = [l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l]
            if r.returncode != 0:
                sys.stderr.write("\n".join(msgs) + "\n")
                raise RuntimeError("hipcc failed on %s" % src)
            if verbose and msgs:
                sys.stderr.write("\n".join(msgs) + "\n")
            objs.append(obj)
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    # the C++ driver that mirrors main.cpp through include/filsim_shim.hpp
    drv = os.path.join(HERE, "filsim_driver")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", drv,
                    os.path.join(HERE, "host", "driver.cpp"),
                    "-L" + HERE, "-lfilsim_amd", "-Wl,-rpath,$ORIGIN"],
                   check=True)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
