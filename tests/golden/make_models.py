"""Copies the reference's four Bézier models (assets/bezier/teapot, teacup,
teaspoon, longwave: data files, the patch tables the reference's `.dl` scenes
load) into tests/golden/models/<name>.be and records their byte size, SHA-256
and the patch/point counts of their headers in models.json.

Run once where the reference checkout is at /root/reference:
    python tests/golden/make_models.py
"""
from __future__ import annotations

import hashlib
import json
import os

SRC = "/root/reference/assets/bezier"
HERE = os.path.dirname(os.path.abspath(__file__))
NAMES = ("teapot", "teacup", "teaspoon", "longwave")


def header_counts(text: bytes) -> tuple[int, int]:
    """Patch count (first number) and point count (the number after the
    16·npatches indices)."""
    tokens = text.replace(b",", b" ").split()
    npatch = int(tokens[0])
    return npatch, int(tokens[1 + 16 * npatch])


def main():
    out = os.path.join(HERE, "models")
    os.makedirs(out, exist_ok=True)
    meta = {}
    for n in NAMES:
        with open(os.path.join(SRC, n), "rb") as f:
            data = f.read()
        with open(os.path.join(out, n + ".be"), "wb") as f:
            f.write(data)
        npatch, npoints = header_counts(data)
        meta[n] = {"bytes": len(data), "sha256": hashlib.sha256(data).hexdigest(),
                   "patches": npatch, "points": npoints, "source": f"assets/bezier/{n}"}
    with open(os.path.join(out, "models.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
        f.write("\n")
    print(json.dumps(meta, indent=1))


if __name__ == "__main__":
    main()
