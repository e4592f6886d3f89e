"""Build ``niq/libniq.so`` (gfx950) with hipcc, in-tree.

Usage: python neural-implicit-queries_amd/build.py
"""
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "niq", "libniq.so")
SOURCES = ["api_model.hip", "api_classify.hip", "api_classify_bf16.hip", "api_classify_rows.hip", "api_tree.hip", "api_rays.hip", "api_rays_rows.hip", "api_cp_mc.hip",
           "api_serve.hip", "api_frustum.hip", "api_cast_noalloc.hip",
           "api_isect.hip", "api_sample.hip", "api_render.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # elementwise range rules are evaluated exactly as written (no FMA
    # contraction), like the reference's separate multiply and add
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def _clean(text):
    return "\n".join(l for l in text.splitlines() if "packed-fp32-ops" not in l)


def build(verbose=True):
    """Compile each translation unit in parallel, then link the shared library."""
    t0 = time.time()
    obj_dir = os.path.join(HERE, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    procs = []
    for src in SOURCES:
        obj = os.path.join(obj_dir, src.replace(".hip", ".o"))
        cmd = [HIPCC] + [f for f in FLAGS if f != "-shared"] + ["-c", "-o", obj, os.path.join(CSRC, src)]
        procs.append((src, obj, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
    objs = []
    for src, obj, p in procs:
        _, err = p.communicate()
        if p.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src} ({p.returncode}):\n{_clean(err)}")
        objs.append(obj)
    tmp = OUT + ".tmp"
    res = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", tmp] + objs,
                         capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed ({res.returncode}):\n{_clean(res.stderr)}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT} in {time.time() - t0:.1f}s", flush=True)
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
