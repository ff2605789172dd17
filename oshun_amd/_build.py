"""Build recipe of liboshun_hip.so (HIP for gfx950, compiled in-tree).

The library is split into translation units (csrc/*.hip): oshun_hip.hip holds
the C ABI and the launch code, each k_*.hip the explicit instantiations of one
family of kernel templates.  They compile in parallel into _build/ and link
into oshun_amd/liboshun_hip.so.  build() always recompiles every unit;
`python -m oshun_amd._build --incremental` recompiles only the units whose
source or included headers changed (dependency files from -MMD).
"""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "liboshun_hip.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # keep the reference's rounding sequence (no a*b+c contraction) except
    # where a kernel turns contraction on for a block of its own
    "-ffp-contract=off",
    # no packed-FP32 instructions: some of their forms have given occasional
    # wrong results on MI355X without any error
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")))


def _obj(src):
    return os.path.join(OBJ, os.path.basename(src)[:-4] + ".o")


def _stale(src):
    o = _obj(src)
    d = o[:-2] + ".d"
    if not (os.path.exists(o) and os.path.exists(d)):
        return True
    t = os.path.getmtime(o)
    with open(d) as fh:
        deps = fh.read().replace("\\\n", " ").split(":", 1)[-1].split()
    return any(not os.path.exists(p) or os.path.getmtime(p) > t for p in deps + [src])


def _compile(hipcc, src, verbose):
    o = _obj(src)
    cmd = [hipcc, *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-MMD", "-MF", o[:-2] + ".d",
           "-c", src, "-o", o + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {os.path.basename(src)}:\n{r.stderr[-6000:]}")
    os.replace(o + ".tmp", o)
    return src


def build(verbose=False, incremental=False, jobs=None):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJ, exist_ok=True)
    srcs = _sources()
    todo = [s for s in srcs if not incremental or _stale(s)]
    jobs = jobs or min(len(todo) or 1, os.cpu_count() or 1, 16)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for s in ex.map(lambda s: _compile(hipcc, s, verbose), todo):
            if verbose:
                print("compiled", os.path.basename(s), flush=True)
    if todo or not os.path.exists(OUT):
        cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,--no-undefined",
               *[_obj(s) for s in srcs], "-o", OUT + ".tmp"]
        subprocess.run(cmd, check=True)
        os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True, incremental="--incremental" in sys.argv)
