"""Builds libpargemslr_amd.so in-tree: host setup (g++) + HIP kernels (hipcc, gfx950).

Usage: python -m pargemslr_amd.build  (or pargemslr_amd.build.build())
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpargemslr_amd.so")

HOST_SRC = ["host/partition.cpp", "host/rcm_ilut.cpp", "host/dense.cpp", "host/gemslr_setup.cpp",
            "host/gen.cpp", "host/trsv_pack.cpp", "host/capi_host.cpp", "host/capi_gen.cpp",
            "host/mm_io.cpp", "host/op_store.cpp",
            "host/inputs.cpp", "host/hypre_shim.cpp"]
HIP_SRC = ["device/dev_op.hip", "device/server.hip"]

# host-only checks built a second time with AddressSanitizer and UBSan (g++,
# no HIP code): the setup, the level-set packing and its replay, the thread pool
SAN_SRC = ["host/partition.cpp", "host/rcm_ilut.cpp", "host/dense.cpp", "host/gemslr_setup.cpp", "host/gen.cpp",
           "host/trsv_pack.cpp"]
SAN_FLAGS = ["-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
             "-fno-sanitize-recover=undefined", "-pthread"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PGS_OFFLOAD_ARCH", "gfx950")
HOST_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", "-pthread"]
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for s in HOST_SRC:
        o = os.path.join(OBJ, s.replace("/", "_") + ".o")
        jobs.append(["g++"] + HOST_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
    for s in HIP_SRC:
        o = os.path.join(OBJ, s.replace("/", "_") + ".o")
        jobs.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
    nlib = len(jobs)
    root = os.path.dirname(HERE)
    san_dir = os.path.join(OBJ, "san")
    os.makedirs(san_dir, exist_ok=True)
    san_objs = []
    for s in SAN_SRC + ["../../tests/cpp/pack_check.cpp"]:
        o = os.path.join(san_dir, os.path.basename(s) + ".o")
        san_objs.append(o)
        jobs.append(["g++"] + SAN_FLAGS + ["-I", os.path.join(root, "include"), "-c", os.path.join(CSRC, s), "-o", o])
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    objs = [j[-1] for j in jobs[:nlib]]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-pthread"])
    os.replace(tmp, LIB)
    cpp = os.path.join(root, "tests", "cpp")
    _run(["g++"] + SAN_FLAGS + san_objs + ["-o", os.path.join(cpp, "pack_check_san")])
    _run(["g++"] + SAN_FLAGS + [os.path.join(cpp, "parallel_for_check.cpp"), "-o",
                                os.path.join(cpp, "parallel_for_check_san")])
    # C++ driver on the class surface (include/pargemslr_amd.hpp), run by the tests
    for name in ("driver_laplacian", "driver_gen", "driver_surface", "pack_check"):
        drv_src = os.path.join(root, "tests", "cpp", name + ".cpp")
        if os.path.exists(drv_src):
            drv = os.path.join(root, "tests", "cpp", name)
            _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(root, "include"), drv_src, "-o", drv,
                  "-L", HERE, "-lpargemslr_amd", "-Wl,-rpath,$ORIGIN/../../pargemslr_amd"])
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
