"""Launch-shape autotuner.

A workload (one scene or a batch) is rendered once per candidate launch
shape, and the fastest is kept and cached under a signature of the
workload (pixels, samples, primitives).  The shape covers the persistent
traversal and shading grids (blocks per compute unit), the live-path count
at which the fused tail takes over, whether passes are captured in a
hipGraph, which tree (binary, four-wide, or the flat leaf array of a small
scene) the render kernels walk, and the queue sizes: the most paths per
pass and the paths a wave claims per work-pool atomic (candidates add them
as "pass_paths" / "claim_paths"; 0 is the library's choice).
Shapes change speed only, never results (tests check the images bit for
bit), so the tuner can be run at any time.
"""
import itertools
import json
import os
import time

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "aurora_amd", "autotune.json")

CANDIDATES = {
    "extend_blocks_per_cu": [4, 8, 16],
    "shade_blocks_per_cu": [8],
    "tail_paths": [1 << 20, (1 << 32) - 1],
    "graph": [False, True],
    "tree": [0, 1, 2],
}


def tree_of(shape):
    """The render tree of a shape: its "tree" (0 binary, 1 four-wide, 2 flat),
    or its "wide" flag."""
    return int(shape["tree"]) if "tree" in shape else (1 if shape.get("wide", False) else 0)


def apply(ctx, shape):
    """Put a launch shape on the context (keys absent from it keep their
    defaults)."""
    ctx.set_launch(shape["extend_blocks_per_cu"], shape["shade_blocks_per_cu"], shape["tail_paths"])
    if hasattr(ctx, "set_render_tree"):
        ctx.set_render_tree(tree_of(shape))
    if hasattr(ctx, "set_queue"):
        ctx.set_queue(shape.get("pass_paths", 0), shape.get("claim_paths", 0))


def signature(items):
    """Workload signature from (scene, camera, settings) triples."""
    px = sum(c.image_width * c.image_height for _, c, _ in items)
    spp = sum(c.image_width * c.image_height * rs.spp for _, c, rs in items)
    prims = sum(sc.arrays.n_prims for sc, _, _ in items)
    return f"n{len(items)}_px{px}_paths{spp}_prims{prims}"


def tune(ctx, run, sig, cache_path=DEFAULT_CACHE, candidates=None, repeats=2, log=None):
    """run(graph) renders the workload once.  Returns (best shape, table of
    timings); the best shape is left applied to ctx."""
    cache = {}
    if cache_path and os.path.exists(cache_path):
        try:
            with open(cache_path) as f:
                cache = json.load(f)
        except (OSError, ValueError):
            cache = {}
    if sig in cache:
        best = cache[sig]["best"]
        apply(ctx, best)
        return best, cache[sig]["table"]
    cand = candidates or CANDIDATES
    keys = list(cand)
    table = []
    for values in itertools.product(*(cand[k] for k in keys)):
        shape = dict(zip(keys, values))
        apply(ctx, shape)
        run(shape["graph"])  # warm-up (allocations, graph capture paths)
        ts = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            run(shape["graph"])
            ts.append(time.perf_counter() - t0)
        shape["seconds"] = min(ts)
        table.append(shape)
        if log:
            log(f"autotune {sig}: {shape}")
    best = min(table, key=lambda s: s["seconds"])
    apply(ctx, best)
    if cache_path:
        cache[sig] = {"best": best, "table": table}
        os.makedirs(os.path.dirname(cache_path), exist_ok=True)
        with open(cache_path, "w") as f:
            json.dump(cache, f, indent=1)
    return best, table


QUEUE_CANDIDATES = {"pass_paths": [0, 1 << 25, 1 << 26], "claim_paths": [0, 32, 256]}


def refine(ctx, run, sig, best, candidates=None, cache_path=DEFAULT_CACHE, repeats=2, log=None):
    """Second stage: with the launch shape `best` fixed, time the queue sizes
    (most paths per pass, paths per work-pool claim; the library's choice,
    0, among them) and keep the fastest.  Returns (best shape, table)."""
    cand = {k: [v] for k, v in best.items() if k != "seconds"}
    cand.update(candidates or QUEUE_CANDIDATES)
    return tune(ctx, run, sig + "_queue", cache_path=cache_path, candidates=cand, repeats=repeats, log=log)
