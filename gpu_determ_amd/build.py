"""Builds the native libraries in-tree with hipcc (gfx950) and g++.

    python -m gpu_determ_amd.build            # libgdt.so, libgpu_determ.so, determ
    python -m gpu_determ_amd.build --oracle   # also oracle/liboracle.so

Outputs (git-ignored, shipped with the tree):
  gpu_determ_amd/libgdt.so          C ABI + HIP kernels (include/gdt.h)
  gpu_determ_amd/libgpu_determ.so   C++ drop-in of the reference host API
  gpu_determ_amd/determ             the reference's driver program (gpu.conf ->
                                    results.out) on top of the drop-in
"""
import os
import subprocess
import sys
import concurrent.futures as cf

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-I" + os.path.join(ROOT, "include"),
]
# Setup kernels select integers from floating-point comparisons: no contraction.
EXTRA = {"setup_kernels.hip": ["-ffp-contract=off"]}
HIP_SOURCES = ["setup_kernels.hip", "collide_kernels.hip", "korobov_kernels.hip", "gdt_abi.hip", "batch_abi.hip", "mfma_gemm.hip", "matrix_mfma.hip", "mfma_abi.hip", "surrogate.hip", "transport.hip", "cell_tiles.hip", "noalloc.hip"]
COMPAT_SOURCES = ["options.cpp", "velocity_grid.cpp", "compat.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise RuntimeError("build step failed: " + os.path.basename(cmd[-1]))
    return r.stdout


def build_gdt():
    os.makedirs(BUILD, exist_ok=True)
    objs, jobs = [], []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [HIPCC] + HIP_FLAGS + EXTRA.get(src, []) + ["-c", "-o", obj, os.path.join(CSRC, src)]
        jobs.append(cmd)
    with cf.ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        list(ex.map(_run, jobs))
    out = os.path.join(PKG, "libgdt.so")
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", out] + objs + ["-lpthread"])
    return out


def build_compat():
    comp = os.path.join(CSRC, "compat")
    if not os.path.isdir(comp):
        return None
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "include", "gpu_determ")]
    srcs = [os.path.join(comp, s) for s in COMPAT_SOURCES if os.path.exists(os.path.join(comp, s))]
    lib = os.path.join(PKG, "libgpu_determ.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", lib] + inc + srcs +
         ["-L" + PKG, "-lgdt", "-Wl,-rpath,$ORIGIN"])
    main = os.path.join(comp, "determ_main.cpp")
    exe = os.path.join(PKG, "determ")
    if os.path.exists(main):
        _run(["g++", "-O2", "-std=c++17", "-o", exe] + inc + [main, "-L" + PKG, "-lgpu_determ", "-lgdt",
                                                              "-Wl,-rpath,$ORIGIN"])
    return lib


def build_oracle():
    _run(["make", "-B", "-C", os.path.join(ROOT, "oracle")])


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    build_gdt()
    build_compat()
    if "--oracle" in argv:
        build_oracle()


if __name__ == "__main__":
    main()
