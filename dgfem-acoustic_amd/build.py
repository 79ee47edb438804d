"""Build recipe for the dgalerkin MI355X libraries (run by __graft_entry__.build()).

Products, all in this directory:
  libdgk.so      HIP kernels + C ABI (hipcc, gfx950): DG solver (dgk_*) and surrogate (dgs_*)
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
DEV_SOURCES = ["csrc/dgk_api.cpp", "csrc/setup.hip", "csrc/stage.hip", "csrc/surrogate.hip"]
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
    """Compile the device sources in parallel, then link libdgk.so."""
    objs, procs = [], []
    os.makedirs(os.path.join(HERE, "_obj"), exist_ok=True)
    for src in DEV_SOURCES:
        obj = os.path.join("_obj", os.path.basename(src) + ".o")
        cmd = [HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj]
        print("+", " ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, cwd=HERE, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                            text=True)))
        objs.append(obj)
    failed = []
    for cmd, p in procs:
        out, _ = p.communicate()
        out = "\n".join(l for l in out.splitlines()
                        if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
        if out.strip():
            print(out, flush=True)
        if p.returncode:
            failed.append(" ".join(cmd))
    if failed:
        raise RuntimeError("build step failed: " + "; ".join(failed))
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", "libdgk.so"] + objs)


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
