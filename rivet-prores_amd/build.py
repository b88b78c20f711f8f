"""Builds the ProRes slice engine for gfx950.

- csrc/*_kernels.hip -> code objects (hipcc --genco --offload-arch=gfx950),
  embedded into the shared library with .incbin and loaded at run time with
  hipModuleLoadData;
- csrc/*.cpp (host runtime and serving loop, plain C++ over the HIP runtime API) ->
  prores_amd/libprores_hip.so, the C ABI of include/prores_hip.h.

Every kernel is compiled with -ffp-contract=off (the reference's f32
transforms multiply then add, never fused) and without packed-FP32
instructions.
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
PKG = os.path.join(HERE, "prores_amd")
LIB = os.path.join(PKG, "libprores_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = "gfx950"

KERNEL_FLAGS = [
    "--genco", f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-ffp-contract=off",
    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "--no-gpu-bundle-output",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-pthread", "-Wall", "-Wextra", "-Wno-unused-parameter",
              "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include"]

MODULES = [("decode", "decode_kernels.hip"), ("encode", "encode_kernels.hip")]
HOST_SOURCES = ["host_common.cpp", "decode_host.cpp", "encode_host.cpp", "serve_host.cpp"]
# Host code with a device variable of its own (the allocation-free entry
# point's scratch array): compiled by hipcc as one HIP source.
HIP_HOST_SOURCES = ["noalloc.hip"]
HIP_HOST_FLAGS = ["-c", "-fPIC", f"--offload-arch={ARCH}", "-O2", "-std=c++17", "-ffp-contract=off",
                  "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ({r.returncode})")
    return r


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    lines = ["\t.section .rodata"]
    for name, src in MODULES:
        co = os.path.join(BUILD, f"{name}.hsaco")
        _run([f"{ROCM}/bin/hipcc", *KERNEL_FLAGS, "-I", CSRC, "-o", co, os.path.join(CSRC, src)], verbose)
        lines += ["\t.p2align 8", f"\t.globl prores_{name}_co", f"prores_{name}_co:",
                  f'\t.incbin "{co}"', "\t.byte 0"]
    lines.append('\t.section .note.GNU-stack,"",@progbits')
    embed = os.path.join(BUILD, "embed.S")
    with open(embed, "w") as f:
        f.write("\n".join(lines) + "\n")
    eo = os.path.join(BUILD, "embed.o")
    _run(["gcc", "-c", "-o", eo, embed], verbose)
    objs.append(eo)
    for src in HOST_SOURCES:
        o = os.path.join(BUILD, src.replace(".cpp", ".o"))
        _run(["g++", *HOST_FLAGS, "-c", "-o", o, os.path.join(CSRC, src)], verbose)
        objs.append(o)
    for src in HIP_HOST_SOURCES:
        o = os.path.join(BUILD, src.replace(".hip", ".o"))
        _run([f"{ROCM}/bin/hipcc", *HIP_HOST_FLAGS, "-I", CSRC, "-o", o, os.path.join(CSRC, src)], verbose)
        objs.append(o)
    tmp = LIB + ".tmp"
    _run(["g++", "-shared", "-pthread", "-o", tmp, *objs, f"-L{ROCM}/lib", "-lamdhip64", f"-Wl,-rpath,{ROCM}/lib"],
         verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
