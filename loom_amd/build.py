"""Build libloomgpu.so (HIP, gfx950) in-tree.

    python -m loom_amd.build         # or __graft_entry__.build()

Steps: embed kb.h as a string for the run-time DAG compiler (kb_src.inc),
then one hipcc invocation over the HIP and C++ sources.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libloomgpu.so")
SOURCES = ["kernels.hip", "noalloc.hip", "prover.cpp", "verifier.cpp", "dagjit.cpp", "server.cpp", "capi.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("LOOMGPU_ARCH", "gfx950")


def embed_kb():
    src = open(os.path.join(CSRC, "kb.h")).read()
    out = os.path.join(CSRC, "kb_src.inc")
    text = 'R"LGKB(' + src + ')LGKB"\n'
    if not os.path.exists(out) or open(out).read() != text:
        with open(out, "w") as f:
            f.write(text)


def build(verbose=False):
    embed_kb()
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wno-unused-command-line-argument",
           "-I" + CSRC, "-o", LIB + ".tmp"]
    for s in SOURCES:
        p = os.path.join(CSRC, s)
        if s.endswith(".cpp"):
            cmd += ["-x", "hip", p]
        else:
            cmd += ["-x", "hip", p]
    cmd += ["-lhiprtc", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed (%d)" % r.returncode)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
