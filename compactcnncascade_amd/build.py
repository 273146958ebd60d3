"""Builds libccnn.so (HIP kernels for gfx950 + C++ host side + C ABI) in-tree.

    python -m compactcnncascade_amd.build
    python -m compactcnncascade_amd.build --phase-stamps    # libccnn_phases.so (profiling)

The shipped model files are embedded into the library as byte arrays, as the
reference embeds them (serialized_models.h), so the library has no run-time file
dependency.
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "libccnn.so")
# the profiling build with the check and stage-1 kernels' phase timestamps compiled in
PHASES_LIB = os.path.join(PKG, "libccnn_phases.so")
MODELS = ["cnn4face1_cntk.bin", "cnn4face2_cntk.bin", "cnn4face3_cntk.bin", "cnn4landmarks_cntk.bin"]

SOURCES = ["ccnn_kernels.hip", "ccnn_noalloc.hip", "ccnn_skipnet.hip", "ccnn_detector.cpp", "ccnn_batch.cpp",
           "ccnn_api.cpp", "ccnn_stages.cpp", "ccnn_convert.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CCNN_OFFLOAD_ARCH", "gfx950")


def gen_models(path):
    lines = ["// generated from compactcnncascade_amd/models by build.py", "#include <cstddef>",
             "namespace ccnn {"]
    for i, name in enumerate(MODELS):
        with open(os.path.join(PKG, "models", name), "rb") as f:
            data = f.read()
        body = ",".join(str(b) for b in data)
        lines.append("static const unsigned char k_model%d[%d] = {%s};" % (i, len(data), body))
    lines.append("void embedded_model(int stage, const unsigned char** bytes, size_t* n)")
    lines.append("{")
    for i in range(len(MODELS)):
        lines.append("    if (stage == %d) { *bytes = k_model%d; *n = sizeof(k_model%d); return; }" % (i, i, i))
    lines.append("    *bytes = nullptr; *n = 0;")
    lines.append("}")
    lines.append("}  // namespace ccnn")
    text = "\n".join(lines) + "\n"
    old = open(path).read() if os.path.exists(path) else None
    if old != text:
        with open(path, "w") as f:
            f.write(text)


def build(verbose=True, phase_stamps=False):
    """Compiles libccnn.so for gfx950 (with phase_stamps: libccnn_phases.so, the same
    library with the kernels' phase timestamps, for ccnn_phase_stamps)."""
    os.makedirs(BUILD, exist_ok=True)
    models_cpp = os.path.join(BUILD, "ccnn_models.cpp")
    gen_models(models_cpp)
    srcs = [os.path.join(CSRC, s) for s in SOURCES] + [models_cpp]
    lib = PHASES_LIB if phase_stamps else LIB
    tmp = lib + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
           # exact path: no a*b+c contraction anywhere (the reference rounds each op)
           "-ffp-contract=off",
           # packed FP32 instructions stay off on this target
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wno-unused-command-line-argument",
           # every symbol the library references must resolve at link time
           "-Wl,-z,defs",
           "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
           "-o", tmp] + (["-DCCNN_PHASE_STAMPS"] if phase_stamps else []) + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip() and verbose:
        print(out)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed (%d):\n%s" % (r.returncode, out))
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    build(phase_stamps="--phase-stamps" in sys.argv[1:])
    sys.exit(0)
