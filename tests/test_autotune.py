"""Autotuner logic on the CPU: a fake context records the launch shapes,
a fake workload's cost depends on the shape; the tuner must pick the
cheapest, leave it applied, and reuse the cached choice without timing."""
import json

from aurora_amd import autotune


class FakeCtx:
    def __init__(self):
        self.shape = None
        self.tree = None

    def set_launch(self, e, s, t):
        self.shape = (e, s, t)

    def set_render_tree(self, tree):
        self.tree = tree


def test_picks_cheapest_and_caches(tmp_path, monkeypatch):
    ctx = FakeCtx()
    calls = []
    clock = [0.0]

    def cost(shape, graph, tree):
        e, s, t = shape
        return (1.0 + abs(e - 8) + (0.5 if t != 1 << 20 else 0.0) + (0.0 if graph else 0.25) +
                {0: 0.1, 1: 0.2, 2: 0.0}[tree])

    def run(graph):
        calls.append((ctx.shape, graph))
        clock[0] += cost(ctx.shape, graph, ctx.tree)

    monkeypatch.setattr(autotune.time, "perf_counter", lambda: clock[0])
    cache = tmp_path / "tune.json"
    best, table = autotune.tune(ctx, run, "sigA", cache_path=str(cache))
    assert (best["extend_blocks_per_cu"], best["tail_paths"], best["graph"], best["tree"]) == (8, 1 << 20, True, 2)
    assert ctx.shape == (8, 8, 1 << 20) and ctx.tree == 2
    n_cand = 3 * 1 * 2 * 2 * 3
    assert len(table) == n_cand and len(calls) == n_cand * 3  # warm-up + 2 repeats each
    saved = json.loads(cache.read_text())
    assert saved["sigA"]["best"]["graph"] is True

    ctx2 = FakeCtx()
    calls.clear()
    best2, _ = autotune.tune(ctx2, run, "sigA", cache_path=str(cache))
    assert not calls and ctx2.shape == (8, 8, 1 << 20) and ctx2.tree == 2 and best2 == best


def test_tree_of_reads_old_and_new_shapes():
    assert autotune.tree_of({"tree": 2}) == 2
    assert autotune.tree_of({"wide": True}) == 1 and autotune.tree_of({"wide": False}) == 0 and autotune.tree_of({}) == 0


def test_signature_separates_workloads():
    class Cam:
        def __init__(self, w):
            self.image_width = self.image_height = w

    class RS:
        def __init__(self, spp):
            self.spp = spp

    class Arr:
        def __init__(self, n):
            self.n_prims = n

    class Sc:
        def __init__(self, n):
            self.arrays = Arr(n)

    a = autotune.signature([(Sc(10), Cam(64), RS(4))])
    b = autotune.signature([(Sc(10), Cam(64), RS(8))])
    c = autotune.signature([(Sc(11), Cam(64), RS(4))])
    assert len({a, b, c}) == 3


def test_refine_sweeps_queue_sizes_with_the_launch_shape_fixed(monkeypatch):
    class Ctx(FakeCtx):
        def __init__(self):
            super().__init__()
            self.queue = None

        def set_queue(self, pass_paths=0, claim_paths=0):
            self.queue = (pass_paths, claim_paths)

    ctx = Ctx()
    clock = [0.0]
    seen = []

    def run(graph):
        seen.append((ctx.shape, ctx.tree, ctx.queue, graph))
        p, c = ctx.queue
        clock[0] += 1.0 + (0.0 if p == 1 << 25 else 0.3) + (0.0 if c == 256 else 0.1)

    monkeypatch.setattr(autotune.time, "perf_counter", lambda: clock[0])
    best = {"extend_blocks_per_cu": 16, "shade_blocks_per_cu": 8, "tail_paths": 1 << 20, "graph": True, "tree": 1,
            "seconds": 5.0}
    best2, table = autotune.refine(ctx, run, "sigQ", best, cache_path=None)
    assert (best2["pass_paths"], best2["claim_paths"]) == (1 << 25, 256)
    assert ctx.queue == (1 << 25, 256) and ctx.shape == (16, 8, 1 << 20) and ctx.tree == 1
    assert len(table) == 9 and all(s[0] == (16, 8, 1 << 20) and s[1] == 1 and s[3] for s in seen)
    # the library's choice (0, 0) is among the candidates
    assert any(t["pass_paths"] == 0 and t["claim_paths"] == 0 for t in table)
