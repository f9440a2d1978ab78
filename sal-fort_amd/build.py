"""Build libsalfort.so (HIP kernels for gfx950 + the C ABI) in-tree, and the Fortran drop-in modules.

    python sal-fort_amd/build.py            # library + Fortran modules
    python sal-fort_amd/build.py --lib      # library only

All HIP code is compiled with -ffp-contract=off (fused multiply-adds are written explicitly) and
without packed FP32 instructions.
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
FSRC = os.path.join(HERE, "fortran")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsalfort.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
FC = os.path.join(ROCM, "bin", "amdflang")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def build_lib(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for f in sources():
        o = os.path.join(BUILD, f.rsplit(".", 1)[0] + ".o")
        objs.append(o)
        lang = ["-x", "hip"] if f.endswith(".hip") else []
        cmds.append([HIPCC] + HIP_FLAGS + lang + ["-c", os.path.join(CSRC, f), "-o", o])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, cmds):
            if out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-pthread"])
    os.replace(tmp, LIB)
    return LIB


def build_fortran():
    """Compile the Fortran drop-in modules (tree_module, bli_module, cube_sphere_module,
    fast_sum_module, mom6_fast_sum_module) into libsalfort_f.so + .mod files under _build/fortran."""
    if not os.path.isdir(FSRC) or not os.path.exists(FC):
        return None
    order = ["salfort_capi.f90", "cube_sphere_module.f90", "bli_module.f90", "tree_module.f90",
             "fast_sum_module.f90", "mom6_fast_sum_module.f90"]
    srcs = [f for f in order if os.path.exists(os.path.join(FSRC, f))]
    if not srcs:
        return None
    moddir = os.path.join(BUILD, "fortran")
    os.makedirs(moddir, exist_ok=True)
    objs = []
    for f in srcs:
        o = os.path.join(moddir, f.replace(".f90", ".o"))
        _run([FC, "-O2", "-fPIC", "-module-dir", moddir, "-c", os.path.join(FSRC, f), "-o", o])
        objs.append(o)
    out = os.path.join(HERE, "libsalfort_f.so")
    _run([FC, "-shared", "-o", out] + objs + ["-L" + HERE, "-lsalfort", "-Wl,-rpath," + HERE])
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="HIP library only")
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args()
    print(build_lib(a.j))
    if not a.lib:
        f = build_fortran()
        if f:
            print(f)


if __name__ == "__main__":
    sys.exit(main())
