"""Benchmark workloads of the D-core path, as concrete seeded inputs.

Each function returns host arrays built by the library's own generators
(dcx_gen_rmat / dcx_gen_chunglu, counter-based SplitMix64 streams), so the
same seed gives the same graph on every machine:

* ``er16k()``            -- directed G(n,m), V=16,384, E=131,072, seed 1 (the
                            CPU reference workload; generator contract of
                            src/graph_gen.cpp:9-34).
* ``rmat(scale, seed)``  -- R-MAT with (a,b,c,d)=(0.57,0.19,0.19,0.05),
                            16*2^scale samples, self-loops and duplicates
                            dropped (scale 20 seed 1: plain GPU workload;
                            scale 22 seed 2: serving base; scale 24 seed 4).
* ``chunglu_batch()``    -- 256 Chung-Lu digraphs of log-uniform size
                            (10^4..2*10^6 samples, average degree 8,
                            exponent 2.1, seed 1000+i): the batched workload.
"""
import math

import numpy as np

import dcx

MASK64 = (1 << 64) - 1


def splitmix64(state):
    """One SplitMix64 step: returns (new_state, output)."""
    state = (state + 0x9E3779B97F4A7C15) & MASK64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return state, z ^ (z >> 31)


def uniform01(seed):
    """First SplitMix64 draw of `seed`, as a double in [0, 1)."""
    _, z = splitmix64(seed)
    return (z >> 11) * (1.0 / (1 << 53))


def er16k(seed=1):
    V, E = 16384, 131072
    src, dst = dcx.gen_er(V, E, seed)
    return V, src, dst


def rmat(scale, seed):
    src, dst = dcx.gen_rmat(scale, 16 << scale, seed)
    return 1 << scale, src, dst


def chunglu_sizes(n=256, seed0=1000, m_lo=1e4, m_hi=2e6, avg_deg=8):
    """(V_i, samples_i) of the batch: samples_i = round(exp(U(ln m_lo, ln m_hi)))
    with U drawn from SplitMix64(seed0 + i), V_i = ceil(samples_i / avg_deg)."""
    out = []
    for i in range(n):
        u = uniform01(seed0 + i)
        m = int(round(math.exp(math.log(m_lo) + u * (math.log(m_hi) - math.log(m_lo)))))
        out.append((int(math.ceil(m / avg_deg)), m))
    return out


def chunglu_batch(n=256, seed0=1000, m_lo=1e4, m_hi=2e6, avg_deg=8, gamma=2.1):
    """List of (V, src, dst): graph i is gen_chunglu(V_i, samples_i, seed0+i)."""
    graphs = []
    for i, (V, m) in enumerate(chunglu_sizes(n, seed0, m_lo, m_hi, avg_deg)):
        src, dst = dcx.gen_chunglu(V, m, seed0 + i, gamma)
        graphs.append((V, src, dst))
    return graphs
