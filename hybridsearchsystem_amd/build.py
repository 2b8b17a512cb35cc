"""Builds the in-tree shared library ``hybridsearchsystem_amd/libfrk.so`` for gfx950, and the search
driver ``hybridsearchsystem_amd/fr_search`` linked against it.

Every ``csrc/*.hip`` file (device kernels and their C-ABI wrappers) and every ``csrc/**/*.cpp`` file
outside ``csrc/driver`` (the C++ host engine and the handle API) is compiled with hipcc in parallel into
``_build/``, then linked. The driver (``csrc/driver/*.cpp``) is compiled against the reference's include
paths (``include/ref_paths``) and links ``libfrk.so`` with an rpath to its own directory. Run
``python -m hybridsearchsystem_amd.build`` (or ``__graft_entry__.build()``).
"""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfrk.so")
DRIVER = os.path.join(HERE, "fr_search")
DRIVER_SRC = os.path.join(CSRC, "driver")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", "-I" + os.path.join(ROOT, "include"),
          "-I" + CSRC]
# Packed FP32 ops are kept out of generated code (device-only flag; the host pass ignores it).
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]
# The scorers keep their MFMA accumulators in ordinary VGPRs: the epilogue reads every accumulator
# once per doc block, and from AGPRs each read would be an extra copy instruction.
PER_FILE = {name: ["-mllvm", "-amdgpu-mfma-vgpr-form"]
            for name in ("frk_score_f32.hip", "frk_score_bf16.hip", "frk_score_fp8.hip", "frk_score_f16x2.hip")}


def _sources():
    hip = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    cpp = sorted(p for p in glob.glob(os.path.join(CSRC, "**", "*.cpp"), recursive=True)
                 if not p.startswith(DRIVER_SRC + os.sep))
    return hip, cpp


def _obj(src):
    rel = os.path.relpath(src, CSRC).replace(os.sep, "_")
    return os.path.join(BUILD, rel + ".o")


def _compile(src, is_hip):
    obj = _obj(src)
    # host-only C++ still goes through hipcc (HIP runtime headers); the arch flag keeps it from
    # compiling for a default device target
    cmd = [HIPCC] + COMMON + (["-x", "hip"] + DEVICE if is_hip else ["--offload-arch=" + ARCH])
    cmd += PER_FILE.get(os.path.basename(src), []) + ["-c", src, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in p.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    return src, p.returncode, err


def build(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    hip, cpp = _sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    failed = []
    with cf.ThreadPoolExecutor(jobs) as ex:
        futs = [ex.submit(_compile, s, True) for s in hip] + [ex.submit(_compile, s, False) for s in cpp]
        for f in cf.as_completed(futs):
            src, rc, err = f.result()
            if rc != 0:
                failed.append((src, err))
            elif verbose and err:
                print(err, file=sys.stderr)
    if failed:
        msg = "\n\n".join("%s:\n%s" % (s, e) for s, e in failed)
        raise RuntimeError("hipcc failed:\n" + msg)
    objs = [_obj(s) for s in hip + cpp]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-lpthread"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed:\n" + p.stderr)
    os.replace(tmp, LIB)
    _build_driver()
    return LIB


def _build_driver():
    """fr_search: the driver sources against the reference's include paths, linked to libfrk.so."""
    srcs = sorted(glob.glob(os.path.join(DRIVER_SRC, "*.cpp")))
    if not srcs:
        return
    tmp = DRIVER + ".tmp"
    cmd = [HIPCC, "-O2", "-std=c++17", "-Wall", "--offload-arch=" + ARCH, "-I" + os.path.join(ROOT, "include", "ref_paths"),
           "-I" + os.path.join(ROOT, "include"), "-o", tmp] + srcs + [
        "-L" + HERE, "-lfrk", "-Wl,-rpath,$ORIGIN", "-lpthread"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("fr_search build failed:\n" + p.stderr)
    os.replace(tmp, DRIVER)


if __name__ == "__main__":
    print(build(verbose=True))
