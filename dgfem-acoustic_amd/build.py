"""Build recipe for the dgalerkin MI355X libraries (run by __graft_entry__.build()).

Products, all in this directory:
  libdgk.so      HIP kernels + C ABI (hipcc, gfx950)
  libdgkhost.so  MSH / config / mesh generators (g++)
  dgalerkin      the command-line driver (links both)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are disabled in all device code.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
DEV_SOURCES = ["csrc/dgk_api.cpp", "csrc/setup.hip", "csrc/stage.hip"]
HOST_SOURCES = ["host/host_capi.cpp", "host/meshio.cpp", "host/config.cpp", "host/meshgen.cpp"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=HERE, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
    if out.strip():
        print(out, flush=True)
    if r.returncode:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")


def build_device():
    _run([HIPCC] + HIP_FLAGS + ["-shared", "-o", "libdgk.so"] + DEV_SOURCES)


def build_host():
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-shared", "-o", "libdgkhost.so"] + HOST_SOURCES)


def build_cli():
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-o", "dgalerkin", "host/dgalerkin.cpp", "-L.", "-ldgk", "-ldgkhost",
          "-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"])


def build_all():
    build_host()
    build_device()
    build_cli()


if __name__ == "__main__":
    targets = sys.argv[1:] or ["all"]
    for t in targets:
        {"all": build_all, "device": build_device, "host": build_host, "cli": build_cli}[t]()
