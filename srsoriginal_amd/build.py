"""Build recipe for the native parts of srsoriginal_amd (run from the repo root or anywhere).

    python -m srsoriginal_amd.build          # or __graft_entry__.build()

Produces, in-tree:
  srsoriginal_amd/lib/libsrs_hip.so   gfx950 HIP kernels + the C ABI of include/srs_hip.h
  srsoriginal_amd/lib/_sr*.so         the host engine (OSCARSSR-compatible C++ class) and
                                      its Python binding; links libsrs_hip.so
Every call recompiles.  hipcc cross-compiles for gfx950 without a GPU.
"""
import os
import subprocess
import sys
import sysconfig
import time

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
HIP_DIR = os.path.join(PKG, "csrc", "hip")
HOST_DIR = os.path.join(PKG, "csrc", "host")
LIB_DIR = os.path.join(PKG, "lib")
BUILD_DIR = os.path.join(PKG, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 (v_pk_*_f32) is disabled for all device code.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wno-unused-command-line-argument"]
# The FP64 kernels rely on the written operation order: no FMA contraction there.
PER_FILE = {
    "srs_fp64.hip": ["-ffp-contract=off"],
    "srs_mixed.hip": ["-ffp-contract=off"],
    "srs_batch.hip": ["-ffp-contract=off"],
    "srs_adjoint.hip": ["-ffp-contract=off"],
    "srs_spectrum.hip": ["-ffp-contract=off"],
    "srs_abi.hip": [],
}


def _run_all(cmds):
    procs = [(c, subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)) for c in cmds]
    failed = []
    for c, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((c, out))
        elif out.strip():
            # hipcc's host pass does not know the packed-fp32 feature; that warning is harmless
            lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l and "is not a recognized feature" not in l]
            if lines:
                print("\n".join(lines), file=sys.stderr)
    if failed:
        for c, out in failed:
            print(" ".join(c), file=sys.stderr)
            print(out, file=sys.stderr)
        raise RuntimeError("native build failed")


def build_hip():
    os.makedirs(BUILD_DIR, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    srcs = [f for f in sorted(os.listdir(HIP_DIR)) if f.endswith(".hip")]
    objs, cmds = [], []
    for f in srcs:
        o = os.path.join(BUILD_DIR, f + ".o")
        objs.append(o)
        cmds.append([HIPCC] + HIP_FLAGS + PER_FILE.get(f, []) + ["-c", os.path.join(HIP_DIR, f), "-o", o])
    _run_all(cmds)
    out = os.path.join(LIB_DIR, "libsrs_hip.so")
    _run_all([[HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", out] + objs])
    return out


def build_host():
    """Host engine + Python binding (g++, pybind11), linked to libsrs_hip.so by rpath."""
    import pybind11

    os.makedirs(LIB_DIR, exist_ok=True)
    srcs = [os.path.join(HOST_DIR, f) for f in sorted(os.listdir(HOST_DIR)) if f.endswith(".cc")]
    if not srcs:
        return None
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(LIB_DIR, "_sr" + suffix)
    # The reference's compile flags (setup.py:33) keep host FP64 arithmetic (the RK4
    # trajectory) bit-identical to the reference: no -march, no FMA contraction.
    cmd = ["g++", "-std=c++17", "-O3", "-fPIC", "-pthread", "-shared", "-ffp-contract=off", "-fvisibility=hidden",
           "-I" + os.path.join(ROOT, "include"), "-I" + HOST_DIR, "-I" + pybind11.get_include(),
           "-I" + sysconfig.get_paths()["include"], "-o", out] + srcs + \
          ["-L" + LIB_DIR, "-lsrs_hip", "-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"]
    _run_all([cmd])
    return out


def build():
    t0 = time.time()
    lib = build_hip()
    ext = build_host()
    print("built %s%s in %.1fs" % (os.path.relpath(lib, ROOT), (", " + os.path.relpath(ext, ROOT)) if ext else "",
                                    time.time() - t0))


if __name__ == "__main__":
    build()
