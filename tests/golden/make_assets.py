"""Copies the reference's scene assets used as test inputs into
tests/golden/assets/, in the reference's own layout so the relative paths
inside the scene files still resolve:

* assets/bezier/{teapot,teacup,teaspoon,longwave}: the `.be` patch tables;
* assets/scenes/*.dl: the scene description files that load them
  (Teaset01, Teaset02, teaset, Test32 … Test152);
* assets/surfaces/*.se: the B-spline and Gregory element surfaces
  (quad16, tube, ball0, dice16, cube16).

These are data files (control points, scene descriptions), not source.
assets.json records each file's byte size and SHA-256, and for the models the
patch and point counts of their headers.

Run once where the reference checkout is at /root/reference:
    python tests/golden/make_assets.py
"""
from __future__ import annotations

import hashlib
import json
import os

SRC = "/root/reference/assets"
HERE = os.path.dirname(os.path.abspath(__file__))
MODELS = ("teapot", "teacup", "teaspoon", "longwave")
SURFACES = ("quad16.se", "tube.se", "ball0.se", "dice16.se", "cube16.se")
SCENES = ("Teaset01.dl", "Teaset02.dl", "teaset.dl", "Test32.dl", "Test56.dl", "Test80.dl", "Test104.dl",
          "Test128.dl", "Test152.dl")


def header_counts(text: bytes) -> tuple[int, int]:
    """Patch count (first number) and point count (the number after the
    16·npatches indices)."""
    tokens = text.replace(b",", b" ").split()
    npatch = int(tokens[0])
    return npatch, int(tokens[1 + 16 * npatch])


def copy(rel: str) -> bytes:
    with open(os.path.join(SRC, rel), "rb") as f:
        data = f.read()
    dst = os.path.join(HERE, "assets", rel)
    os.makedirs(os.path.dirname(dst), exist_ok=True)
    with open(dst, "wb") as f:
        f.write(data)
    return data


def main():
    meta = {}
    for n in MODELS:
        data = copy(f"bezier/{n}")
        npatch, npoints = header_counts(data)
        meta[f"bezier/{n}"] = {"bytes": len(data), "sha256": hashlib.sha256(data).hexdigest(),
                               "patches": npatch, "points": npoints}
    for n in SCENES:
        data = copy(f"scenes/{n}")
        meta[f"scenes/{n}"] = {"bytes": len(data), "sha256": hashlib.sha256(data).hexdigest()}
    for n in SURFACES:
        data = copy(f"surfaces/{n}")
        meta[f"surfaces/{n}"] = {"bytes": len(data), "sha256": hashlib.sha256(data).hexdigest()}
    with open(os.path.join(HERE, "assets", "assets.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
        f.write("\n")
    print(json.dumps(meta, indent=1))


if __name__ == "__main__":
    main()
