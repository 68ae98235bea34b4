"""Writes tests/golden/meshes.npz from the reference's mesh data files
(applications/particles_gpu/meshes/hopper.obj, meshes/torusknot1.obj): the
rigid boundaries the particle applications load.  The arrays are the parsed
vertices (float64) and triangles (uint32, 0-based, polygons fanned).

Run from the repository root:  python tests/golden/make_meshes.py
"""
from pathlib import Path
import sys

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from fullc0ntact_amd.meshes import load_obj  # noqa: E402

REF = Path("/root/reference")
SOURCES = {
    "hopper": REF / "applications/particles_gpu/meshes/hopper.obj",
    "torusknot1": REF / "meshes/torusknot1.obj",
}


def main():
    out = {}
    for name, path in SOURCES.items():
        v, f = load_obj(path)
        out[f"{name}_vertices"] = v
        out[f"{name}_triangles"] = f
    np.savez_compressed(Path(__file__).with_name("meshes.npz"), **out)


if __name__ == "__main__":
    main()
