"""Builds libvahana_hip.so (gfx950) in-tree from csrc/ with hipcc.

Usage: python vahana.jl_amd/build.py [-j N]
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libvahana_hip.so")
BUILD = os.path.join(HERE, "_build")
ARCH = os.environ.get("VH_OFFLOAD_ARCH", "gfx950")

SOURCES = ["vh_prims.hip", "vh_engine.hip", "vh_query.hip", "vh_bodies.hip", "vh_ensemble.hip",
           "vh_graphgen.hip", "vh_userbody.hip", "vh_graph.cpp"]

# -ffp-contract=off: no fused multiply-add, the model bodies round every
# product and sum like the reference's IEEE arithmetic; packed FP32 ops are
# disabled for the device code.
HIPFLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-result", "-Wno-unused-function"]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = ["hipcc"] + HIPFLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
               "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile {src} failed:\n{r.stdout}\n{r.stderr}")
    return obj


EXAMPLES = os.path.join(HERE, "examples")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")


def build_example_bodies():
    """Code objects of the run-time bodies in examples/ (vh_register_body)."""
    out = []
    for f in sorted(os.listdir(EXAMPLES)):
        if not f.endswith(".hip"):
            continue
        co = os.path.join(EXAMPLES, f[:-4] + ".co")
        cmd = (["hipcc", "--genco"] + [x for x in HIPFLAGS if x not in ("-fPIC",)] +
               ["-I", INCLUDE, os.path.join(EXAMPLES, f), "-o", co])
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile {f} failed:\n{r.stdout}\n{r.stderr}")
        out.append(co)
    return out


def build(jobs=None, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4)
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = ["hipcc", f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    build_example_bodies()
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    j = None
    if "-j" in sys.argv:
        j = int(sys.argv[sys.argv.index("-j") + 1])
    build(j, verbose=True)
