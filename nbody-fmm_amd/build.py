"""Builds libnbfmm.so (HIP, gfx950) in-tree and the CPU oracle libraries.

    python nbody-fmm_amd/build.py            # library + oracle (+ reference build if present)
    python nbody-fmm_amd/build.py --lib-only # only libnbfmm.so

hipcc cross-compiles for gfx950 without a GPU.  Objects go to
nbody-fmm_amd/_build/, the library to nbody-fmm_amd/libnbfmm.so.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libnbfmm.so")
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("NBFMM_ARCH", "gfx950")

# -packed-fp32-ops: the compiler emits no packed FP32 instructions (v_pk_*_f32),
# some forms of which have given wrong results on gfx950 without an error.
FLAGS = [
    "-O3", "-std=c++20", f"--offload-arch={ARCH}", "-fPIC", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", f"-I{os.path.join(ROOT, 'include')}",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build_lib(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    objs = []
    cmds = []
    for f in srcs:
        o = os.path.join(BUILD, f[:-4] + ".o")
        objs.append(o)
        cmds.append([HIPCC, *FLAGS, "-c", os.path.join(CSRC, f), "-o", o])
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    tmp = OUT + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, OUT)
    return OUT


HOST_LIB = os.path.join(HERE, "libnbfmm_host.so")
DROPIN_TEST = os.path.join(BUILD, "test_dropin")
FRAMES_TEST = os.path.join(BUILD, "test_frames")


def build_host():
    """The C++20 drop-in classes (Tensor, Particle, FMM, NaiveUpdate,
    SyplecticEuler, RungeKutta4) over the C ABI, and the C++ test program."""
    os.makedirs(BUILD, exist_ok=True)
    inc = os.path.join(HERE, "host", "include")
    srcs = [os.path.join(HERE, "host", "src", f) for f in sorted(os.listdir(os.path.join(HERE, "host", "src")))
            if f.endswith(".cpp")]
    rpath = "-Wl,-rpath,$ORIGIN"
    _run(["g++", "-std=c++20", "-O2", "-fPIC", "-shared", f"-I{inc}", "-o", HOST_LIB, *srcs,
          f"-L{HERE}", "-lnbfmm", rpath])
    for exe, src in ((DROPIN_TEST, "test_dropin.cpp"), (FRAMES_TEST, "test_frames.cpp")):
        _run(["g++", "-std=c++20", "-O2", f"-I{inc}", "-o", exe, os.path.join(ROOT, "tests", "cpp", src),
              f"-L{HERE}", "-lnbfmm_host", "-lnbfmm", f"-Wl,-rpath,{HERE}"])
    return HOST_LIB


def build_oracle():
    od = os.path.join(ROOT, "oracle")
    _run(["make", "-B", "-C", od, "liborc.so"])
    ref_src = "/root/reference/src"
    if os.path.isdir(ref_src):
        _run(["make", "-C", od, "-j8", "ref"])


if __name__ == "__main__":
    build_lib()
    build_host()
    if "--lib-only" not in sys.argv:
        build_oracle()
    print("built", OUT)
