"""Builds the gfx950 HIP library and the host-only generator library in-tree.

    python -m comer2_amd.build        (or comer2_amd.build.build())
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

HIP_SOURCES = ["c2_phase1.hip", "c2_phase2.hip", "c2_map.hip", "c2_sdp.hip", "c2_lib.cpp", "c2_pipe.cpp",
               "c2_prep.cpp", "c2_fields.cpp", "c2_batchproc.cpp", "c2_stream.cpp",
               "c2_noalloc.hip", "c2_dlist.hip", "c2_dpipe.cpp", "c2_cvs.cpp", "c2_proread.cpp", "c2_format.cpp"]
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]
OBJDIR = os.path.join(HERE, "_build")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd[:3]) + " ...")
    return r


def build_synth():
    out = os.path.join(HERE, "libc2synth.so")
    _run(["g++", "-O2", "-fopenmp", "-fPIC", "-shared", "-ffp-contract=off", "-o", out,
          os.path.join(CSRC, "c2_synth.cpp")])
    return out


def build_hip():
    """one object per source, compiled in parallel, then linked"""
    from concurrent.futures import ThreadPoolExecutor
    out = os.path.join(HERE, "libcomer2_hip.so")
    tmp = out + ".tmp"
    os.makedirs(OBJDIR, exist_ok=True)
    objs = [os.path.join(OBJDIR, os.path.splitext(s)[0] + ".o") for s in HIP_SOURCES]
    jobs = [[HIPCC] + HIP_FLAGS + ["-c", "-o", o, os.path.join(CSRC, s)] for s, o in zip(HIP_SOURCES, objs)]
    with ThreadPoolExecutor(min(8, len(jobs))) as ex:
        list(ex.map(_run, jobs))
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, out)
    return out


def build():
    return build_synth(), build_hip()


if __name__ == "__main__":
    print(build())
