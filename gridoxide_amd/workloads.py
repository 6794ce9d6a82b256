"""Synthetic multi-topology workloads: the uneven batch (topologies of log-uniform size, half of
their scenarios N-1) and the serving request stream (Poisson arrivals over a pool of grids).

Grids come from :func:`gridoxide_amd.grid.generate_grid`, scenarios from the generator's own load
distribution (:func:`gridoxide_amd.grid.scenario_loads`)."""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from .grid import Network, generate_grid, scenario_loads


@dataclass
class Job:
    """One topology with its scenarios (outage[s]: branch out of service in scenario s, or -1)."""
    net: Network
    p: np.ndarray
    q: np.ndarray
    vm0: np.ndarray
    va0: np.ndarray
    outage: Optional[np.ndarray] = None


def log_uniform_int(rng, lo: int, hi: int) -> int:
    return int(round(float(np.exp(rng.uniform(np.log(lo), np.log(hi))))))


def bridges(net: Network) -> np.ndarray:
    """Bridge flag of every branch (lines, then transformers): removing a bridge splits the grid."""
    f = np.r_[net.line_from, net.tr_from].astype(np.int64)
    t = np.r_[net.line_to, net.tr_to].astype(np.int64)
    n = net.n_bus
    adj = [[] for _ in range(n)]
    for e in range(f.size):
        if f[e] != t[e]:
            adj[f[e]].append((t[e], e))
            adj[t[e]].append((f[e], e))
    disc = [-1] * n
    low = [0] * n
    out = np.zeros(f.size, bool)
    timer = 0
    for root in range(n):
        if disc[root] >= 0:
            continue
        disc[root] = low[root] = timer
        timer += 1
        stack = [(root, -1, 0)]
        while stack:
            v, pe, i = stack[-1]
            if i < len(adj[v]):
                stack[-1] = (v, pe, i + 1)
                u, e = adj[v][i]
                if e == pe:
                    continue
                if disc[u] < 0:
                    disc[u] = low[u] = timer
                    timer += 1
                    stack.append((u, e, 0))
                else:
                    low[v] = min(low[v], disc[u])
            else:
                stack.pop()
                if stack:
                    parent = stack[-1][0]
                    low[parent] = min(low[parent], low[v])
                    if low[v] > disc[parent]:
                        out[pe] = True
    return out


def uneven_batch(n_topo: int = 64, min_bus: int = 100, max_bus: int = 20000, min_scen: int = 16,
                 max_scen: int = 1024, seed: int = 404, n1_fraction: float = 0.5) -> List[Job]:
    """Topologies with log-uniform bus and scenario counts; ``n1_fraction`` of each topology's
    scenarios take one random non-bridge branch out of service (the Y-bus pattern is kept, so the
    symbolic factorization is shared).  Flat start."""
    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(n_topo):
        nb = log_uniform_int(rng, min_bus, max_bus)
        S = log_uniform_int(rng, min_scen, max_scen)
        net = generate_grid(nb, max(nb // 2, 1), seed=seed * 1000 + k)
        p, q = scenario_loads(net, S, seed=seed * 1000 + k)
        outage = np.full(S, -1, np.int32)
        n1 = int(round(S * n1_fraction))
        cand = np.nonzero(~bridges(net))[0]
        if n1 and cand.size:
            pick = rng.permutation(S)[:n1]
            outage[pick] = rng.choice(cand, n1)
        jobs.append(Job(net, p, q, np.repeat(net.vm[None], S, 0), np.repeat(net.va[None], S, 0), outage))
    return jobs


@dataclass
class Request:
    arrival: float  # seconds after the start of the stream
    grid: int       # index into the pool
    p: np.ndarray
    q: np.ndarray


def serving_pool(n_grids: int = 32, min_bus: int = 200, max_bus: int = 50000, seed: int = 505) -> List[Network]:
    """Grids with log-uniform bus counts."""
    rng = np.random.default_rng(seed)
    pool = []
    for k in range(n_grids):
        nb = log_uniform_int(rng, min_bus, max_bus)
        pool.append(generate_grid(nb, max(nb // 2, 1), seed=seed * 1000 + k))
    return pool


def poisson_requests(pool: List[Network], rate_hz: float, n_req: int, min_scen: int = 1, max_scen: int = 256,
                     seed: int = 606) -> List[Request]:
    """Arrivals with exponential gaps at ``rate_hz``; each request picks a grid uniformly and a
    uniform scenario count in [min_scen, max_scen]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(n_req):
        t += rng.exponential(1.0 / rate_hz)
        g = int(rng.integers(len(pool)))
        S = int(rng.integers(min_scen, max_scen + 1))
        p, q = scenario_loads(pool[g], S, seed=seed * 100000 + k)
        out.append(Request(t, g, p, q))
    return out
