"""Build recipe for the gfx950 engine library and the FoMo C++ facade.

Produces, in-tree:
  fomo_amd/libfomo_hip.so  HIP kernels + C ABI (include/fomo_hip.h); fomo_device.hip and the sixteen
                           k_nn_los instantiation units of fomo_nnlos_part.hip (eight lambda chunk
                           groups x two index kinds) compile in parallel
  fomo_amd/libFoMo.so      FoMo.h C++ API (FoMoObject & co.) on top of libfomo_hip.so

Run `python -m fomo_amd.build` or call build_all(). hipcc cross-compiles without a GPU.
"""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "fomo_amd")
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")
BUILD = os.path.join(ROOT, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Device flags: gfx950 only; no FMA contraction (bit-exact FP64 geometry); packed FP32 off.
HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-I" + INC,
]
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-I" + INC, "-pthread"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


NNLOS_GROUPS = 8  # lambda chunk groups of fomo_nnlos_part.hip (k_nn_los instantiations), per index kind


def _run_all(cmds):
    """Run compile commands in parallel; raise if any fails."""
    procs = []
    for cmd in cmds:
        print("+", " ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd)))
    bad = [cmd for cmd, p in procs if p.wait() != 0]
    if bad:
        raise subprocess.CalledProcessError(1, bad[0])


def build_engine():
    os.makedirs(BUILD, exist_ok=True)
    dev_o = os.path.join(BUILD, "fomo_device.o")
    host_o = os.path.join(BUILD, "fomo_host.o")
    parts = [(g, r) for g in range(NNLOS_GROUPS) for r in (0, 1)]
    part_o = [os.path.join(BUILD, f"fomo_nnlos_part{g}_{r}.o") for g, r in parts]
    cmds = [[HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, "fomo_device.hip"), "-o", dev_o],
            ["g++"] + HOST_FLAGS + ["-c", os.path.join(CSRC, "fomo_host.cpp"), "-o", host_o]]
    cmds += [[HIPCC] + HIP_FLAGS + [f"-DFOMO_NNLOS_GROUP={g}", f"-DFOMO_NNLOS_RECT={r}", "-c",
                                    os.path.join(CSRC, "fomo_nnlos_part.hip"), "-o", o]
             for (g, r), o in zip(parts, part_o)]
    _run_all(cmds)
    out = os.path.join(PKG, "libfomo_hip.so")
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", out, dev_o, host_o] + part_o + ["-pthread"])
    return out


def build_facade():
    srcs = [os.path.join(CSRC, f) for f in ("fomo_object.cpp", "fomo_io.cpp")]
    srcs = [s for s in srcs if os.path.exists(s)]
    if not srcs:
        return None
    out = os.path.join(PKG, "libFoMo.so")
    _run(["g++"] + HOST_FLAGS + ["-shared", "-o", out] + srcs +
         ["-L" + PKG, "-lfomo_hip", "-Wl,-rpath,$ORIGIN", "-lz"])
    return out


def build_all():
    build_engine()
    build_facade()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
