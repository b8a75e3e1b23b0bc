"""Build the native libraries in-tree.

``libabcmb_hip.so``  - HIP kernels + C ABI, gfx950 only (hipcc).
``libabcmb_host.so`` - host background producer (g++ with OpenMP).
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=fast", "-munsafe-fp-atomics",
    f"-I{os.path.join(ROOT, 'include')}",
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_host():
    out = os.path.join(HERE, "libabcmb_host.so")
    _run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-fopenmp",
          f"-I{os.path.join(ROOT, 'include')}",
          os.path.join(CSRC, "host_background.cpp"), "-o", out])
    return out


def build_hip():
    """Each .hip source to an object in parallel (the ODE kernels dominate the
    compile time), then one link."""
    from concurrent.futures import ThreadPoolExecutor
    out = os.path.join(HERE, "libabcmb_hip.so")
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    objdir = os.path.join(ROOT, "_build", "hip")
    os.makedirs(objdir, exist_ok=True)
    cflags = [f for f in HIP_FLAGS if f != "-shared"]
    objs = [os.path.join(objdir, os.path.basename(s)[:-4] + ".o") for s in srcs]
    with ThreadPoolExecutor(max(1, min(len(srcs), os.cpu_count() or 1))) as ex:
        for f in [ex.submit(_run, [HIPCC] + cflags + ["-c", s, "-o", o]) for s, o in zip(srcs, objs)]:
            f.result()
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", tmp])
    os.replace(tmp, out)
    return out


def build_fluid_variant(directory: str, host: bool = True, hip: bool = True):
    """Kernel and host libraries for one model's user fluids: the same sources
    compiled against the generated ``fluid_dev.h`` / ``fluid_host.h`` in
    ``directory`` (abcmb_amd/fluidgen.py), linked into that directory."""
    from concurrent.futures import ThreadPoolExecutor
    directory = os.path.abspath(directory)
    dev_h = os.path.join(directory, "fluid_dev.h")
    host_h = os.path.join(directory, "fluid_host.h")
    if not (os.path.exists(dev_h) and os.path.exists(host_h)):
        raise FileNotFoundError(f"{directory} has no generated fluid headers")
    if host:
        tmp = os.path.join(directory, "libabcmb_host.so.tmp")
        _run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-fopenmp", f"-I{os.path.join(ROOT, 'include')}",
              f'-DABCMB_FLUID_HOST="{host_h}"', os.path.join(CSRC, "host_background.cpp"), "-o", tmp])
        os.replace(tmp, os.path.join(directory, "libabcmb_host.so"))
    if not hip:
        return None
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    objdir = os.path.join(ROOT, "_build", "fluids", os.path.basename(directory))
    os.makedirs(objdir, exist_ok=True)
    cflags = [f for f in HIP_FLAGS if f != "-shared"] + [f'-DABCMB_FLUID_DEV="{dev_h}"']
    objs = [os.path.join(objdir, os.path.basename(s)[:-4] + ".o") for s in srcs]
    with ThreadPoolExecutor(max(1, min(len(srcs), os.cpu_count() or 1))) as ex:
        for f in [ex.submit(_run, [HIPCC] + cflags + ["-c", s, "-o", o]) for s, o in zip(srcs, objs)]:
            f.result()
    out = os.path.join(directory, "libabcmb_hip.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_all():
    build_host()
    build_hip()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    {"host": build_host, "hip": build_hip, "all": build_all}[what]()
