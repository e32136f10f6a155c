"""Builds libbpl_amd.so (all HIP sources of csrc/) for gfx950 with hipcc, in-tree.

Usage: python bp-layers_amd/build.py [-j N] [--only FILE ...]
The library lands in bp-layers_amd/bplayers_amd/_lib/libbpl_amd.so.  ``--only`` recompiles the
named csrc files and links them with the objects of the last full build.
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "bplayers_amd", "_lib")
OBJ_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libbpl_amd.so")
ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops: no v_pk_*_f32 in the generated code.
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-fno-gpu-rdc",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-lambda-capture"]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


# The LBP kernels compare and reduce floats that are never NaN (labels past C carry +inf):
# told there are no NaNs, the compiler folds DPP moves into v_min_f32_dpp / v_max_f32_dpp
# instead of canonicalising every operand first.
FILE_FLAGS = {"lbp_fwd.hip": ["-fno-honor-nans"]}


def compile_one(src):
    obj = os.path.join(OBJ_DIR, os.path.basename(src) + ".o")
    cmd = [HIPCC] + FLAGS + FILE_FLAGS.get(os.path.basename(src), []) + ["-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r


def build(jobs=8, verbose=False, only=None):
    os.makedirs(OUT_DIR, exist_ok=True)
    os.makedirs(OBJ_DIR, exist_ok=True)
    todo = sources()
    if only:
        names = {os.path.basename(f) for f in only}
        unknown = names - {os.path.basename(f) for f in todo}
        if unknown:
            raise ValueError(f"not csrc sources: {sorted(unknown)}")
        # sources that include a named file (lbp_bwd_pair<L>.hip includes lbp_bwd.hip) rebuild too
        for f in todo:
            with open(f) as fh:
                text = fh.read()
            if any(f'#include "{n}"' in text for n in names):
                names.add(os.path.basename(f))
        todo = [f for f in todo if os.path.basename(f) in names]
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(compile_one, todo))
    failed = False
    for src, obj, r in results:
        if r.returncode != 0 or verbose:
            sys.stderr.write(r.stdout + r.stderr)
        if r.returncode != 0:
            failed = True
            sys.stderr.write(f"FAILED: {src}\n")
    if failed:
        raise RuntimeError("libbpl_amd build failed")
    objs = [os.path.join(OBJ_DIR, os.path.basename(f) + ".o") for f in sources()]
    missing = [o for o in objs if not os.path.exists(o)]
    if missing:
        raise RuntimeError(f"objects missing (run a full build first): {missing}")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("libbpl_amd link failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("-v", action="store_true")
    ap.add_argument("--only", nargs="+", metavar="FILE", help="recompile only these csrc files")
    args = ap.parse_args()
    print(build(args.j, args.v, args.only))
