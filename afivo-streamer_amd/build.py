"""Build the native libraries in-tree.

* libafgpu.so  -- HIP for gfx950 (hipcc), the product.
* libafhost.so -- host-side tree / initial-condition / table stand-in (g++).
* oracle/liboracle_afcpu.so -- the CPU oracle (test infrastructure).
* oracle/_ref/ -- the reference's own routines and the fixture drivers
  (test infrastructure), where the reference checkout is present.
* tests/fortran/_build/drive_step -- a Fortran driver of libafgpu through
  include/m_afgpu.f90 (test infrastructure), where amdflang is present.

All floating-point code is built with -ffp-contract=off so that no fused
multiply-adds are formed: the FP64 device path then evaluates every cell
value with the same operations, in the same order, as the CPU path.
"""
import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
INC = os.path.join(REPO_DIR, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("AFGPU_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=" + ARCH,
    "-ffp-contract=off",
    # no packed FP32 instructions: some of their forms have given wrong
    # results on this hardware now and then, so the compiler must not emit them
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd, cwd=None):
    print("+ " + " ".join(cmd), flush=True)
    subprocess.run(cmd, cwd=cwd, check=True)


def build_gpu():
    out = os.path.join(PKG_DIR, "libafgpu.so")
    srcs = [os.path.join(PKG_DIR, "csrc", "afgpu.hip"),
            os.path.join(PKG_DIR, "csrc", "coarse_factors.cpp")]
    tmp = out + ".tmp"
    _run([HIPCC] + HIP_FLAGS + ["-I" + INC, "-I" + os.path.join(PKG_DIR, "csrc")] + srcs + ["-o", tmp])
    os.replace(tmp, out)
    return out


def build_host():
    out = os.path.join(PKG_DIR, "libafhost.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
          os.path.join(PKG_DIR, "host", "afhost.cpp"), "-o", tmp])
    os.replace(tmp, out)
    return out


def build_oracle():
    _run(["make", "-B", "-C", os.path.join(REPO_DIR, "oracle")])
    return os.path.join(REPO_DIR, "oracle", "liboracle_afcpu.so")


REFERENCE = os.environ.get("AFGPU_REFERENCE", "/root/reference")
FLANG = "/opt/rocm/bin/amdflang"


def build_ref():
    """The reference's own afivo and lookup-table routines and the fixture
    drivers that call them (oracle/ref/Makefile, outputs in oracle/_ref),
    where the reference checkout and amdflang are present; the tests that
    use them are skipped elsewhere."""
    if not (os.path.isdir(REFERENCE) and os.path.exists(FLANG)):
        print("reference checkout or amdflang not present: oracle/_ref not built", flush=True)
        return None
    _run(["make", "-C", os.path.join(REPO_DIR, "oracle", "ref"), "REF=" + REFERENCE, "all"])
    return os.path.join(REPO_DIR, "oracle", "_ref")


def build_fortran():
    """The Fortran driver of the C-ABI (tests/fortran/drive_step.f90 over the
    interface module include/m_afgpu.f90), linked against libafgpu.so; the
    GPU test that runs it is skipped where amdflang is absent."""
    if not os.path.exists(FLANG):
        print("amdflang not present: the Fortran driver is not built", flush=True)
        return None
    out_dir = os.path.join(REPO_DIR, "tests", "fortran", "_build")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "drive_step")
    _run([FLANG, "-O2", "-module-dir", out_dir, os.path.join(INC, "m_afgpu.f90"),
          os.path.join(REPO_DIR, "tests", "fortran", "drive_step.f90"), "-L" + PKG_DIR, "-lafgpu",
          "-Wl,-rpath,$ORIGIN/../../../afivo-streamer_amd", "-o", out + ".tmp"], cwd=out_dir)
    os.replace(out + ".tmp", out)
    return out


def build_all():
    build_host()
    build_oracle()
    build_ref()
    build_gpu()
    build_fortran()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "gpu": build_gpu, "host": build_host, "oracle": build_oracle, "ref": build_ref,
         "fortran": build_fortran}[w]()
