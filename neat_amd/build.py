"""Builds libneat_amd.so (HIP kernels + C ABI, gfx950) and the libtorch host extension in-tree.

    python -m neat_amd.build            # both
    python -m neat_amd.build --abi      # only the C-ABI library
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
KDIR = os.path.join(HERE, "csrc", "kernels")
HDIR = os.path.join(HERE, "csrc", "host")
OBJ = os.path.join(ROOT, "_build", "obj")
ABI_LIB = os.path.join(HERE, "libneat_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -ffp-contract=off: the exact paths (ray sampling, trilinear weights) must round like the oracle; hot loops use
#   explicit fmaf where a fused multiply-add is wanted.
# -packed-fp32-ops disabled: packed FP32 forms are not trusted on this hardware pool.
HIP_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-ffp-contract=off", "-munsafe-fp-atomics",
    "-fhip-fp32-correctly-rounded-divide-sqrt", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-Wno-unused-result",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{KDIR}",
]

# per-file additions: the relu BF16 backward keeps matrix-core results in VGPRs (its products feed vector ALU code)
FILE_FLAGS = {"projector_relu.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def _source_id(kind):
    sys.path.insert(0, HERE)
    try:
        import buildinfo
    finally:
        sys.path.pop(0)
    return buildinfo.source_id(kind)


def build_abi(jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(f for f in os.listdir(KDIR) if f.endswith(".hip"))
    objs = []
    bid = f"-DNEAT_BUILD_ID=id{_source_id('abi')}"

    def compile_one(f):
        o = os.path.join(OBJ, f[:-4] + ".o")
        _run([HIPCC, *HIP_FLAGS, *FILE_FLAGS.get(f, []), bid, "-c", os.path.join(KDIR, f), "-o", o])
        return o

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = ABI_LIB + ".tmp"
    # the soname lets a module that needs libneat_amd.so reuse the copy the Python loader already mapped
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-soname,libneat_amd.so", *objs, "-o", tmp])
    os.replace(tmp, ABI_LIB)
    return ABI_LIB


def build_host(jobs=8):
    """libtorch C++ module layer (the reference's HyperTreeBase / IndirectGridSample3D / FCBlock / geometry API)."""
    import glob
    import torch.utils.cpp_extension as ce
    srcs = sorted(glob.glob(os.path.join(HDIR, "*.cpp")))
    if not srcs:
        return None
    os.environ.setdefault("MAX_JOBS", str(jobs))
    build_dir = os.path.join(ROOT, "_build", "host")
    os.makedirs(build_dir, exist_ok=True)
    ce.load(
        name="_neat_host",
        sources=srcs,
        extra_include_paths=[os.path.join(ROOT, "include"), HDIR, "/opt/rocm/include"],
        extra_cflags=["-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__=1", "-Wno-unused-variable",
                      f"-DNEAT_HOST_BUILD_ID=id{_source_id('host')}"],
        extra_ldflags=[f"-L{HERE}", "-lneat_amd", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{HERE}",
                       "-L" + os.path.join(os.path.dirname(ce.__file__), "..", "..", "lib"), "-lc10_hip"],
        build_directory=build_dir,
        verbose=False,
        # built and dlopened only: importing it here as a top-level module would register its pybind11 types once
        # more when the package imports its in-tree copy (neat_amd._neat_host) in the same process
        is_python_module=False,
    )
    # copy the built module next to the package so it travels with the tree and is loaded from there
    import shutil
    so = os.path.join(build_dir, "_neat_host.so")
    dst = os.path.join(HERE, "_neat_host.so")
    shutil.copyfile(so, dst + ".tmp")
    os.replace(dst + ".tmp", dst)
    return dst


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--abi", action="store_true")
    ap.add_argument("--host", action="store_true")
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args(argv)
    both = not (a.abi or a.host)
    if a.abi or both:
        print("built", build_abi(a.j))
    if a.host or both:
        print("built", build_host(a.j))


if __name__ == "__main__":
    sys.exit(main())
