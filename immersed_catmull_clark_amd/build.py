"""Builds libicc.so (HIP kernels for gfx950 + C ABI) in-tree with hipcc.

Usage: python -m immersed_catmull_clark_amd.build   (or build() from Python)
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libicc.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-munsafe-fp-atomics",
    # packed FP32 (v_pk_*_f32) stays off: the scalar forms are used throughout
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
HOST_FLAGS = ["-x", "c++", "-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__", "-Wno-unused-result",
              "-Wno-unused-value",
              "-I/opt/rocm/include"]

SOURCES = [
    # (file, extra flags)
    ("k_apply.hip", []),
    ("k_vec.hip", []),
    ("k_misc.hip", ["-ffp-contract=off"]),
    ("k_immersed.hip", ["-ffp-contract=off"]),
    ("k_faces.hip", []),
    ("k_topopt.hip", ["-ffp-contract=off"]),
    ("k_wq.hip", []),
    ("k_split.hip", []),
    ("k_noalloc.hip", []),
    ("k_surface.hip", ["-ffp-contract=off"]),
    ("setup.cpp", None),
    ("wq.cpp", None),
    ("api.cpp", None),
    ("api_ext.cpp", None),
    ("api_imm.cpp", None),
    ("api_imm_topopt.cpp", None),
    ("api_wq.cpp", None),
    ("api_surface.cpp", None),
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + out)
    return out


def build(verbose=False, jobs=None):
    """Compiles every source (in parallel, at most `jobs` at once) and links
    libicc.so and the topopt_3d driver."""
    from concurrent.futures import ThreadPoolExecutor

    os.makedirs(BUILD, exist_ok=True)
    cmds = []
    objs = []
    for name, extra in SOURCES:
        src = os.path.join(CSRC, name)
        if not os.path.exists(src):
            continue
        obj = os.path.join(BUILD, name + ".o")
        if extra is None:
            cmd = [HIPCC] + HOST_FLAGS + ["-c", src, "-o", obj]
        else:
            cmd = [HIPCC] + HIP_FLAGS + extra + ["-c", src, "-o", obj]
        cmds.append(cmd)
        objs.append(obj)
    jobs = jobs or min(8, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    # the topopt_3d command-line driver (links libicc.so next to it)
    exe = os.path.join(PKG, "topopt_3d")
    _run(["g++", "-O2", "-std=c++17", os.path.join(CSRC, "topopt_3d.cpp"), "-o", exe + ".tmp", "-L" + PKG, "-licc",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(exe + ".tmp", exe)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
