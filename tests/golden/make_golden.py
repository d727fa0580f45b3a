"""Regenerate tests/golden/ref_host.json from the reference's own host code, and
tests/golden/assimp_math.json from the Assimp inline math the reference vendors.

Builds ``oracle/_ref`` (``make -C oracle ref``: the reference's rng.cpp, aliastable.h,
colour.cpp, lalgebra.cpp, scene/camera.cpp, io/imagewrite.cpp and util/meansq/meansq.cpp,
compiled where they lie in the reference checkout), runs the harness, and runs the
reference ``meansq`` tool on pairs of the harness's ``.hdr`` outputs.

Only numbers are committed (inputs and outputs); no reference source is copied.
Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import os
import re
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
OUT = os.path.join(ROOT, "tests", "golden", "ref_host.json")
OUT_AI = os.path.join(ROOT, "tests", "golden", "assimp_math.json")


def main():
    subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "ref"])
    ref = os.path.join(ROOT, "oracle", "_ref")
    with tempfile.TemporaryDirectory() as tmp:
        raw = subprocess.check_output([os.path.join(ref, "refharness"), tmp], text=True)
        data = json.loads(raw)
        for case in data["hdr"]:
            a = case.pop("path")
            b = a.replace(".hdr", "_b.hdr")
            with open(b, "rb") as f:
                case["twin_file_hex"] = f.read().hex()
            out = subprocess.check_output([os.path.join(ref, "meansq"), a, b], text=True)
            m = re.search(r"MSE:\s*([0-9eE+\-.]+)", out)
            if not m:
                sys.exit("meansq printed no MSE: " + out)
            case["meansq_twin_vs_file"] = m.group(1)
            out = subprocess.check_output([os.path.join(ref, "meansq"), a, a], text=True)
            case["meansq_self"] = re.search(r"MSE:\s*([0-9eE+\-.]+)", out).group(1)
    data["_source"] = ("reference host code compiled by oracle/Makefile `ref` "
                       "(src/rng.cpp, src/aliastable.h, src/colour.cpp, src/lalgebra.cpp, "
                       "src/scene/camera.cpp, src/io/imagewrite.cpp, util/meansq/meansq.cpp)")
    with open(OUT, "w") as f:
        json.dump(data, f, indent=0, sort_keys=True)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")
    ai = json.loads(subprocess.check_output([os.path.join(ref, "aiharness")], text=True))
    ai["_source"] = ("oracle/assimp_harness.cpp compiled by oracle/Makefile `ref` against the reference's "
                     "include/assimp/{vector3,matrix3x3,matrix4x4,quaternion}.inl (Assimp 5.4.3); "
                     "floats are bit patterns")
    with open(OUT_AI, "w") as f:
        json.dump(ai, f, indent=0, sort_keys=True)
    print("wrote", OUT_AI, os.path.getsize(OUT_AI), "bytes")


if __name__ == "__main__":
    main()
