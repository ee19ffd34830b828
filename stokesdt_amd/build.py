"""Builds stokesdt_amd/libstokesdt_amd.so for gfx950 with hipcc (in-tree) and the
command-line BD driver stokesdt_amd/stokesdt_bd and the calc_diff tool linked
against it."""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libstokesdt_amd.so")
PKG = HERE
EXE = os.path.join(HERE, "stokesdt_bd")
CALC_DIFF = os.path.join(HERE, "calc_diff")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SDT_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = [
    ("spme_kernels.hip", []),
    ("sort_kernels.hip", []),
    # exact FP64 pair decisions: no contraction into fused multiply-adds
    ("pair_kernels.hip", ["-ffp-contract=off"]),
    ("real_kernels.hip", []),
    ("block_kernels.hip", []),
    ("mob_gpu.hip", []),
    ("simbox.hip", []),
    ("diffusion.hip", []),
    ("brwn_random.hip", []),
    ("plain_fp64.hip", []),
    ("ewald_mob.hip", []),
    ("tune.hip", []),
    ("ensemble.hip", []),
    ("serve.hip", []),
    ("capi.hip", []),
]
HOST_SOURCES = [("host_linalg.cpp", []), ("host_util.cpp", []), ("molecule_io.cpp", [])]

HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-result",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in r.stdout.splitlines()
                    if "packed-fp32-ops" not in line)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build(verbose=False, jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    cmds = []
    objs = []
    for src, extra in HIP_SOURCES:
        o = os.path.join(OBJ, src + ".o")
        objs.append(o)
        cmds.append([HIPCC] + HIP_FLAGS + extra + ["-c", os.path.join(CSRC, src), "-o", o])
    for src, extra in HOST_SOURCES:
        o = os.path.join(OBJ, src + ".o")
        objs.append(o)
        cmds.append(["g++", "-O3", "-fPIC", "-std=c++17"] + extra +
                    ["-c", os.path.join(CSRC, src), "-o", o])
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs +
         ["-L/opt/rocm/lib", "-lrocblas", "-lrocsolver", "-lhipfft", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    # command-line BD driver over the C ABI
    exe_tmp = EXE + ".tmp"
    _run(["g++", "-O2", "-std=c++17", os.path.join(CSRC, "bd_main.cpp"), "-o", exe_tmp,
          "-L" + PKG, "-lstokesdt_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(exe_tmp, EXE)
    # calc_diff (tools/calc_diff) over the C ABI
    exe_tmp = CALC_DIFF + ".tmp"
    _run(["g++", "-O2", "-std=c++17", os.path.join(CSRC, "calc_diff_main.cpp"), "-o", exe_tmp,
          "-L" + PKG, "-lstokesdt_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(exe_tmp, CALC_DIFF)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
