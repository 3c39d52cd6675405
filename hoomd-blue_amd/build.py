"""Build libhoomd_amd.so (gfx950) in-tree with hipcc.

    python hoomd-blue_amd/build.py [--jobs N] [--incremental]

Each .hip source under csrc/ is compiled to an object in parallel, then linked into
hoomd-blue_amd/hoomd_amd/libhoomd_amd.so, next to the Python package that loads it.
"""

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "hoomd_amd", "libhoomd_amd.so")
ARCH = os.environ.get("HM_OFFLOAD_ARCH", "gfx950")

SOURCES = ["kernels.hip", "sort.hip", "capi.hip", "engine.hip", "mlp.hip", "batch.hip", "pair_ref.hip"]

CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # no packed FP32 instructions (v_pk_*_f32): some of their forms have given silently wrong
    # results on these MI355X machines; hipcc would otherwise form them from float2-like code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{os.path.join(ROOT, 'include')}",
    f"-I{CSRC}",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    raise RuntimeError("hipcc not found")


def _filter(out):
    lines = [ln for ln in out.splitlines()
             if "packed-fp32-ops" not in ln and "not a recognized feature" not in ln and ln.strip()]
    return "\n".join(lines)


def _includes(path, seen):
    """Quoted includes of path found in csrc/ or include/, recursively."""
    import re
    for name in re.findall(r'^\s*#\s*include\s+"([^"]+)"', open(path).read(), re.M):
        for d in (os.path.dirname(path), CSRC, os.path.join(ROOT, "include")):
            q = os.path.join(d, name)
            if os.path.exists(q) and q not in seen:
                seen.add(q)
                _includes(q, seen)
                break
    return seen


def _up_to_date(src, obj):
    """The object is newer than its source and every header it includes (transitively)."""
    if not os.path.exists(obj):
        return False
    t = os.path.getmtime(obj)
    path = os.path.join(CSRC, src)
    deps = [path] + sorted(_includes(path, set()))
    return all(os.path.getmtime(p) < t for p in deps)


def compile_one(src, incremental=False):
    os.makedirs(OBJ, exist_ok=True)
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    if incremental and _up_to_date(src, obj):
        return obj
    cmd = [hipcc()] + CFLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{_filter(p.stdout + p.stderr)}")
    return obj


# Out-of-library evaluator plugins (each its own shared library, as a user's plugin would be)
PLUGINS = [os.path.join(HERE, "plugins", "pair_example", "pair_example.hip")]


def build_plugin(src, incremental=False):
    out = os.path.join(os.path.dirname(src), "lib" + os.path.basename(src)[:-4] + ".so")
    deps = [src] + sorted(_includes(src, set()))
    if incremental and os.path.exists(out) and \
            all(os.path.getmtime(p) < os.path.getmtime(out) for p in deps):
        return out
    cmd = [hipcc()] + CFLAGS + ["-shared", src, "-o", out + ".tmp"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"plugin build failed: {src}\n{_filter(p.stdout + p.stderr)}")
    os.replace(out + ".tmp", out)
    return out


def build(jobs=4, verbose=False, incremental=False):
    """Compile every source (or, with incremental, only those older than their inputs) and
    link the library; then build the evaluator plugins."""
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: compile_one(s, incremental), SOURCES))
        plugins = list(ex.map(lambda s: build_plugin(s, incremental), PLUGINS))
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", OUT + ".tmp"] + objs
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"link failed\n{_filter(p.stdout + p.stderr)}")
    os.replace(OUT + ".tmp", OUT)
    if verbose:
        print("built", OUT, *plugins)
    return OUT


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("--incremental", action="store_true",
                    help="recompile only sources older than their object files")
    a = ap.parse_args()
    build(a.jobs, verbose=True, incremental=a.incremental)


if __name__ == "__main__":
    sys.exit(main())
