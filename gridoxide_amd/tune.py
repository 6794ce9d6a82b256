"""Launch-shape autotuner: picks the tile width (scenarios per wavefront) and the LDS solve-window
rows by timing short engine runs of the heaviest requests of a workload."""
from __future__ import annotations

import time
from typing import Iterable, List, Sequence, Tuple

from .solver import Engine, Mode, Topology

# 2-scenario tiles are left out: they pay off only while a launch has fewer tiles than SIMDs, which
# the multi-request engine's launches exceed.  Tile 8 runs pair tiles (two scenarios per lane) whose
# FP32 refactor runs as quad tiles (four per lane).
DEFAULT_CANDIDATES = ((4, 0), (8, 0), (4, 256), (8, 1024))


def _cost(topo: Topology, n_scen: int) -> float:
    s = topo.sizes()
    return float(s["n"]) * max(1.0, n_scen / 4.0)


def tune_launch(requests: Sequence[Tuple[Topology, object, object, object, object, object]], mode: Mode = Mode.MIXED,
                candidates: Iterable[Tuple[int, int]] = DEFAULT_CANDIDATES, steps: int = 2, heaviest: int = 3,
                log=None) -> Tuple[int, int, List[Tuple[int, int, float]]]:
    """requests: (topology, p, q, vm0, va0, outage) tuples.  Times ``steps`` Newton iterations of
    the ``heaviest`` requests under each (tile, ring_rows) candidate (after one untimed run that
    builds the schedules) and returns the fastest shape and every measurement."""
    ranked = sorted(requests, key=lambda r: _cost(r[0], r[1].shape[0]), reverse=True)[:heaviest]
    results = []
    for tile, ring in candidates:
        for topo, *_ in ranked:
            topo.set_launch(tile, ring)
        best = None
        for rep in range(2):
            eng = Engine(tile=tile)
            t0 = time.perf_counter()
            ids = [eng.submit(topo, p, q, vm0, va0, max_iter=steps, mode=mode, outage=o)
                   for topo, p, q, vm0, va0, o in ranked]
            eng.run()
            for rid in ids:
                eng.collect(rid)
            dt = time.perf_counter() - t0
            eng.close()
            if rep == 1:
                best = dt
        results.append((tile, ring, best))
        if log:
            log(f"tune: tile {tile} ring {ring or 'default'}: {best * 1e3:.1f} ms for {steps} steps")
    tile, ring, _ = min(results, key=lambda r: r[2])
    return tile, ring, results
