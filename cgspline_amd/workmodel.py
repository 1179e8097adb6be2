"""Algorithmic work model of the ray–patch path (bytes and FP32 operations).

The work of a launch is fixed by event counts of the deterministic algorithm
(node visits, hash probes, leaf visits, patch tests, clip iterations by kind,
onHit evaluations, rays), not by hardware counters.  The GPU path executes the
same per-ray algorithm bit for bit, so the counts the CPU oracle reports for
a workload are the GPU launch's counts.

Bytes per unit (the layout the kernels read; every fetch counted as if it
missed in cache — the "algorithmic" traffic):

* inner node visit: 64 B (two child boxes + child offsets)
* hash probe on ascent: 8 B (displacement word + table word)
* leaf visit: 8 B (element id + EMPTY terminator)
* patch test: 256 B (16 pre-gathered float4 control points)
* object visit (top-level leaf): 160 B (world→local transform + material +
  surface record)
* per pixel written: 4 B (packed frame), +12 B if the float colour is written

FP32 operations per unit (add/sub/mul/min/max/compare = 1, div/sqrt = 1),
counted from the kernels' source:

* inner node visit: 2 slab tests × 3 axes × (2 sub + 2 mul + 2 select +
  2 compare) + 6 = 54
* patch projection: 2 cross products and versors (≈ 40) + 16 × (3 sub + 2
  dot3) = 248
* clip iteration (hull + slice): tolerance 14 + line 14 + distance patch 48
  + envelope 24 + hull walk ≈ 152 + widen/delta 7 + position update 4 +
  4 curves × 2 coordinates × 12 lerps × 3 = 288 → 551
* subdivision iteration: hull part 259 + two half patches 576 → 835
* prune iteration: hull part 259
* converged region: tolerance 14 + two triangle tests ≈ 40 → 54
* onHit evaluation: power-basis weights 62 + 5 point combinations 140 +
  projection/length/dot ≈ 20 → 222
"""
from __future__ import annotations

B_NODE, B_PROBE, B_LEAF, B_PATCH, B_OBJECT, B_PIXEL, B_COLOR = 64, 8, 8, 256, 160, 4, 12

F_NODE, F_PROJ, F_CLIP, F_SUBDIV, F_PRUNE, F_CONVERGED, F_ONHIT = 54, 248, 551, 835, 259, 54, 222

# MI355X published peaks.  The FP32 vector peak counts packed (two-wide)
# FP32 instructions; without them (this library is built with packed FP32
# disabled) a CU issues 64 FP32 FMA lanes per clock: 256 CUs x 2.4 GHz x 128.
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP32_PEAK_NONPACKED_TFLOPS = 78.6


# The events a launch of the library counts (cgs_work_counts, a
# CGS_WORK_COUNTS=1 build), in the ABI's order.
GPU_EVENTS = ("cullNodes", "cullElems", "nodes", "hashProbes", "leaves", "patchTests", "clips", "splits", "prunes",
              "converged", "onHits", "objects")

# Per event of the library's own traversal: a cull-tree node or element box
# test is one slab test (3 axes × (2 sub + 2 mul + 2 select + 2 compare) + 3
# compares = 27 operations) reading 32 B (+ 8 B of links for a node); an
# object-local ray costs the world→local transform (28), its projection (3),
# the direction's transform and normalisation (15 + 10) and the clipping
# frame (2 cross products and versors, ≈ 40): ≈ 100.
F_CULL, F_OBJECT = 27, 100
B_CULL_NODE, B_CULL_ELEM = 40, 32


def flops_of_counts(c: dict) -> float:
    """FP32 operations of a launch from the library's own event counts."""
    return (F_CULL * (c["cullNodes"] + c["cullElems"]) + F_NODE * c["nodes"] + F_PROJ * c["patchTests"]
            + F_CLIP * c["clips"] + F_SUBDIV * c["splits"] + F_PRUNE * c["prunes"]
            + F_CONVERGED * c["converged"] + F_ONHIT * c["onHits"] + F_OBJECT * c["objects"])


def bytes_of_counts(c: dict, pixels: int, color: bool = False) -> float:
    """Algorithmic bytes of a launch from the library's own event counts."""
    return (B_CULL_NODE * c["cullNodes"] + B_CULL_ELEM * c["cullElems"] + B_NODE * c["nodes"]
            + B_PROBE * c["hashProbes"] + B_LEAF * c["leaves"] + B_PATCH * c["patchTests"]
            + B_OBJECT * c["objects"] + pixels * (B_PIXEL + (B_COLOR if color else 0)))


def bytes_of(work: dict, pixels: int, objects_visited: int = 0, color: bool = False) -> float:
    """Algorithmic bytes of a render/trace launch from oracle work counts."""
    return (B_NODE * work["nodes"] + B_PROBE * work["hashProbes"] + B_LEAF * work["leaves"]
            + B_PATCH * work["patchTests"] + B_OBJECT * objects_visited
            + pixels * (B_PIXEL + (B_COLOR if color else 0)))


def flops_of(work: dict) -> float:
    """FP32 operations of the reference algorithm on a launch's rays, from
    the oracle's work counts (the reference's traversal at both levels)."""
    plain_clips = work["clipIters"] - work["subdivs"] - work["prunes"]
    return (F_NODE * work["nodes"] + F_PROJ * work["patchTests"] + F_CLIP * plain_clips
            + F_SUBDIV * work["subdivs"] + F_PRUNE * work["prunes"] + F_CONVERGED * work["converged"]
            + F_ONHIT * work["onHits"])
