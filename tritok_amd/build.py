"""Builds libtritok_amd.so (HIP for gfx950 + host C++) in-tree, and the CPU
oracle used by the tests.

    python tritok_amd/build.py          # library + oracle (+ reference if present)
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtritok_amd.so")
OBJDIR = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["-O3", "-fPIC", "-std=c++17", "-I" + os.path.join(REPO, "include"), "-I" + CSRC,
          "-Wall", "-Wno-unused-function"]
# device code: gfx950 only; packed FP32 instructions disabled
HIPFLAGS = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-ffp-contract=fast", "-munsafe-fp-atomics"]
# sources whose products and sums must round separately, as the reference's
# (pivot searches and thresholds compare values with the CPU restatement)
NO_CONTRACT = {"tt_compress.hip"}
# sources whose products and sums fuse only within one expression, so that
# the same expressions compile to the same roundings in every kernel that
# holds them (the least-squares kernels: the persistent and the kernel-by-
# kernel forms must give the same bits)
CONTRACT_ON = {"tt_lsqr.hip"}
# sources whose matrix-core loops keep their accumulators in VGPRs (without
# the option the compiler copies loop-carried accumulators between VGPRs and
# AGPRs around every product chain)
MFMA_VGPR = {"tt_integrals.hip", "tt_dense.hip", "tt_mhsd.hip"}


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _stale(obj, srcs):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(s) > t for s in srcs)


def build_library(force=True, jobs=8):
    os.makedirs(OBJDIR, exist_ok=True)
    headers = glob.glob(os.path.join(CSRC, "*.h")) + glob.glob(os.path.join(REPO, "include", "*.h"))
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.cpp")) + glob.glob(os.path.join(CSRC, "*.hip")))
    procs, objs = [], []
    for s in srcs:
        o = os.path.join(OBJDIR, os.path.basename(s) + ".o")
        objs.append(o)
        if not force and not _stale(o, [s] + headers):
            continue
        if s.endswith(".hip"):
            flags = HIPFLAGS
            if os.path.basename(s) in NO_CONTRACT:
                flags = [f for f in HIPFLAGS if not f.startswith("-ffp-contract")] + ["-ffp-contract=off"]
            if os.path.basename(s) in CONTRACT_ON:
                flags = [f for f in HIPFLAGS if not f.startswith("-ffp-contract")] + ["-ffp-contract=on"]
            if os.path.basename(s) in MFMA_VGPR:
                flags = flags + ["-mllvm", "-amdgpu-mfma-vgpr-form"]
            cmd = [HIPCC, "-x", "hip"] + COMMON + flags + ["-c", s, "-o", o]
        else:
            cmd = [HIPCC, "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"] + COMMON + ["-c", s, "-o", o]
        print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        if len(procs) >= jobs:
            for p in procs:
                if p.wait() != 0:
                    raise RuntimeError("compilation failed")
            procs = []
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("compilation failed")
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-lamdhip64", "-lpthread", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    return LIB


FORTRAN = os.path.join(HERE, "fortran")
DRIVER = os.path.join(FORTRAN, "tt_fortran_driver")
FC = os.environ.get("FC_AMD", "/opt/rocm/bin/amdflang")


def build_fortran_driver():
    """The Fortran host example (fortran/tt_fortran_driver.f90 over the
    ISO_C_BINDING module), linked to the in-tree library; skipped without a
    Fortran compiler."""
    if not os.path.exists(FC):
        print("no Fortran compiler at", FC, "- driver not built", flush=True)
        return None
    os.makedirs(OBJDIR, exist_ok=True)
    mod = os.path.join(OBJDIR, "tritok_amd_bind.o")
    _run([FC, "-O2", "-c", os.path.join(FORTRAN, "tritok_amd_bind.f90"), "-o", mod, "-module-dir", OBJDIR])
    _run([FC, "-O2", os.path.join(FORTRAN, "tt_fortran_driver.f90"), mod, "-I", OBJDIR, "-o", DRIVER + ".tmp",
          "-L" + HERE, "-ltritok_amd", "-Wl,-rpath,$ORIGIN/.."])
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


def build_oracle():
    _run(["make", "-C", os.path.join(REPO, "oracle"), "-B"])


def build_reference():
    """Compiles the reference binary into oracle/_ref when its sources exist."""
    if os.path.isdir("/root/reference/src"):
        _run(["make", "-f", os.path.join(REPO, "oracle", "Makefile.ref"), "-C", REPO, "-j8"])


def main():
    build_library(force="--incremental" not in sys.argv)
    build_fortran_driver()
    build_oracle()
    if "--no-ref" not in sys.argv:
        build_reference()


if __name__ == "__main__":
    main()
