"""Build the native libraries in-tree.

* ``yetanotherslam_amd/libyas.so``        HIP kernels + C ABI (hipcc, gfx950)
* ``yetanotherslam_amd/libyas_synth.so``  synthetic RGB-D generator (g++)
* ``oracle/_build/liboracle.so``          scalar C++ oracle (g++, test infrastructure)

Usage: python -m yetanotherslam_amd.build [--no-oracle]
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(REPO, "_build", "obj")

HIP_SOURCES = ["yas_map.hip", "yas_mapops.hip", "yas_batch.hip", "yas_raycast.hip", "yas_tracker.hip",
               "yas_mesh.hip", "yas_pipeline.hip", "yas_server.hip"]

# -ffp-contract=off: one rounding per written operation (bit-exact parity
# with the scalar oracle); packed-fp32 instructions are disabled.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-mcode-object-version=5", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build_hip(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    hipcc = _hipcc()

    def one(src):
        obj = os.path.join(OBJ, src.replace(".hip", ".o"))
        _run([hipcc] + HIPCC_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=min(8, len(HIP_SOURCES))) as ex:
        objs = list(ex.map(one, HIP_SOURCES))
    out = os.path.join(HERE, "libyas.so")
    tmp = out + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build_synth(verbose=False):
    out = os.path.join(HERE, "libyas_synth.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-o", tmp, os.path.join(CSRC, "synth.cpp")])
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build_oracle(verbose=False):
    """liboracle.so (single thread: the checker) and liboracle_omp.so (the
    same source with OpenMP: the all-cores CPU baseline's timing only)."""
    odir = os.path.join(REPO, "oracle", "_build")
    os.makedirs(odir, exist_ok=True)
    outs = []
    for name, extra in (("liboracle.so", []), ("liboracle_omp.so", ["-fopenmp"])):
        out = os.path.join(odir, name)
        tmp = out + ".tmp"
        _run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared"] + extra +
             ["-o", tmp, os.path.join(REPO, "oracle", "oracle.cpp")])
        os.replace(tmp, out)
        outs.append(out)
        if verbose:
            print("built", out)
    return outs[0]


def build_facade_replay(verbose=False):
    """tests/cpp/facade_replay: the C++ facade driven like FullSystem's Tracking."""
    out = os.path.join(REPO, "tests", "cpp", "facade_replay")
    _run([_hipcc(), "--offload-arch=gfx950", "-O1", "-std=c++17", os.path.join(REPO, "tests", "cpp", "facade_replay.cpp"),
          "-L" + HERE, "-lyas", "-Wl,-rpath," + HERE, "-o", out])
    if verbose:
        print("built", out)
    return out


def build_all(oracle=True, verbose=False):
    with ThreadPoolExecutor(max_workers=3) as ex:
        futs = [ex.submit(build_hip, verbose), ex.submit(build_synth, verbose)]
        if oracle:
            futs.append(ex.submit(build_oracle, verbose))
        outs = [f.result() for f in futs]
    outs.append(build_facade_replay(verbose))
    return outs


if __name__ == "__main__":
    build_all(oracle="--no-oracle" not in sys.argv, verbose=True)
