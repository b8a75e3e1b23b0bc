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


HOST_SRCS = ("host_background.cpp", "hyrex.cpp")


def sources(kind: str):
    """The files a library is compiled from: "host" (the producer, its headers and
    the HyRex tables) or "hip" (every kernel source and header)."""
    inc = [os.path.join(ROOT, "include", "abcmb.h")]
    host_headers = ("hyrex.h", "host_dual.h", "fluid_none_host.h")
    if kind == "host":
        return ([os.path.join(CSRC, f) for f in HOST_SRCS + host_headers] + inc +
                sorted(glob.glob(os.path.join(HERE, "data", "hyrex", "*.dat"))))
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")) +
                  [h for h in glob.glob(os.path.join(CSRC, "*.h")) if os.path.basename(h) not in host_headers]) + inc
GEN = os.path.join(ROOT, "_build", "gen")


def hyrex_tables():
    """The HyRec tables (abcmb_amd/data/hyrex/*.dat) as a C++ include for hyrex.cpp,
    written to _build/gen/hyrex_tables.inc."""
    import numpy as np
    d = os.path.join(HERE, "data", "hyrex")
    alpha = np.loadtxt(os.path.join(d, "Alpha_inf.dat"))
    r = np.loadtxt(os.path.join(d, "R_inf.dat"))
    sw = np.loadtxt(os.path.join(d, "fit_swift.dat"))
    if alpha.shape != (4000, 4) or r.shape != (100,) or sw.ndim != 2 or sw.shape[1] != 5:
        raise ValueError("unexpected HyRex table shapes")
    fmt = lambda a: ",".join(repr(float(v)) for v in np.asarray(a).ravel())
    os.makedirs(GEN, exist_ok=True)
    src = (f"// generated by abcmb_amd/build.py from abcmb_amd/data/hyrex/*.dat\n"
           f"const double HX_ALPHA[4000][4] = {{{fmt(alpha)}}};\n"
           f"const double HX_R[100] = {{{fmt(r)}}};\n"
           f"constexpr int HX_SWIFT_N = {sw.shape[0]};\n"
           f"const double HX_SWIFT[{sw.shape[0]}][5] = {{{fmt(sw)}}};\n")
    path = os.path.join(GEN, "hyrex_tables.inc")
    if not os.path.exists(path) or open(path).read() != src:
        with open(path, "w") as fh:
            fh.write(src)
    return path


def _host_cmd(out, extra=()):
    hyrex_tables()
    return (["g++", "-O3", "-fno-math-errno", "-fPIC", "-shared", "-std=c++17", "-fopenmp", f"-I{os.path.join(ROOT, 'include')}",
             f"-I{GEN}"] + list(extra) + [os.path.join(CSRC, f) for f in HOST_SRCS] + ["-o", out])


def build_host():
    out = os.path.join(HERE, "libabcmb_host.so")
    _run(_host_cmd(out + ".tmp"))
    os.replace(out + ".tmp", out)
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
        _run(_host_cmd(tmp, [f'-DABCMB_FLUID_HOST="{host_h}"']))
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
