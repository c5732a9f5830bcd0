"""Rows derived from the row above ("chains", dcx_config.chain_rows): only
every chain_rows-th team row is an anchor and the rows between are refined
from the row above with per-vertex supports; an anchor is peeled or
(dcx_config.anchors = 1) refined from row 0's values.  Every row must equal
the CPU oracle's, bit for bit, for chain spacings 2, 3, 5 and 16 and with
chains off (1), with both kinds of anchor; graphs where every row is a team
row (lds_members = -1) and the default split; the derived rows must show in
the pass log (variant DERIVED, the first row of each chain flagged -- 1
after a peeled anchor, 2 for an anchor derived from row 0 -- its list
reads counted).  At R-MAT 20 the state hash must equal the oracle's
(tests/golden/hashes.json)."""
import json
import os
import sys

import numpy as np
import pytest

import dcx
import graphs
import oracle
from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def hub_digraph(V=3000, seed=5):
    """G(n,m) plus two hubs pointing to and from most vertices: rows whose
    supports change through the hubs' long lists (walked by workgroups)."""
    rng = np.random.default_rng(seed)
    _, s, d = graphs.random_gnm(V, 20 * V, seed)
    s, d = list(np.asarray(s)), list(np.asarray(d))
    for h in (0, 1):
        others = rng.choice(np.arange(2, V), size=V // 2 + 300, replace=False)
        for o in others:
            s.append(h)
            d.append(int(o))
            s.append(int(o))
            d.append(h)
    e = np.unique(np.stack([np.array(s, np.uint32), np.array(d, np.uint32)], 1), axis=0)
    e = e[e[:, 0] != e[:, 1]]
    return V, e[:, 0].copy(), e[:, 1].copy()


CASES = {
    "rmat12": lambda: (1 << 12,) + tuple(dcx.gen_rmat(12, 16 << 12, 3)),
    "rmat14": lambda: (1 << 14,) + tuple(dcx.gen_rmat(14, 16 << 14, 8)),
    "gnm3000": lambda: graphs.random_gnm(3000, 45000, 21),
    "hubs": hub_digraph,
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("spacing,anchors", [(2, 0), (3, 0), (5, 0), (16, 0), (1, 0), (2, 1),
                                             (3, 1), (5, 1), (16, 1), (1, 1)])
def test_chains_equal_oracle(name, spacing, anchors):
    V, s, d = CASES[name]()
    s, d = np.asarray(s, np.uint32), np.asarray(d, np.uint32)
    ref = oracle.decompose(oracle.csr_from_edges(V, s, d))
    with dcx.Context(0, lds_members=-1, chain_rows=spacing, anchors=anchors) as c:
        say(f"chains {name} V={V} E={len(s)} spacing={spacing} anchors={anchors}")
        g = c.upload_edges(V, s, d)
        K, L = g.decompose(dcx.VARIANT_BATCHED)
        t = c.timing()
        assert (K, L) == (ref[0], ref[2])
        assert np.array_equal(g.kmaxes(), ref[1])
        assert np.array_equal(g.rows(), ref[3])
        log = g.pass_log()
        derived = [p for p in log[1:] if p["variant"] == dcx.VARIANT_DERIVED]
        if anchors == 1 and K >= 2:
            # every team row but row 0 is derived: anchors from row 0
            # (flag 2), the rows between from the row above; row 1 is row
            # 0's chain, set up before the launch (flag 0)
            assert len(derived) == K, derived
            for p in derived:
                anchor = p["k"] % spacing == 0 or spacing == 1
                want = 2 if anchor and p["k"] >= 2 else 0
                assert p["levels"] == want, p
                if want:
                    assert p["alive"] > 0, p
        elif spacing > 1 and K >= 2:
            assert derived and all(p["k"] % spacing != 0 for p in derived), derived
            for p in derived:
                # first row of a chain from a peeled anchor (row 0's chain is
                # set up by whole-grid kernels before the launch)
                first = (p["k"] - 1) % spacing == 0
                assert p["levels"] == (1 if first and p["k"] > 1 else 0), p
                if first and p["k"] > 1:
                    assert p["alive"] > 0, p  # the supports' set-up read lists
        else:
            assert not derived
        assert t.rows_scheduler == 0
        # a second decomposition on the same context (reused scratch)
        K2, L2 = g.decompose(dcx.VARIANT_BATCHED)
        assert (K2, L2) == (K, L)
        assert np.array_equal(g.rows(), ref[3])
        g.free()


def test_chains_default_split_rmat16():
    """The default launch: big rows on teams (peeled or derived), small rows
    LDS-resident beside them."""
    V = 1 << 16
    s, d = dcx.gen_rmat(16, 16 << 16, 4)
    ref = oracle.decompose(oracle.csr_from_edges(V, s, d))
    for spacing in (0, 4):
        with dcx.Context(0, chain_rows=spacing) as c:
            say(f"chains rmat16 default split spacing={spacing}")
            g = c.upload_edges(V, s, d)
            K, L = g.decompose(dcx.VARIANT_BATCHED)
            assert (K, L) == (ref[0], ref[2])
            assert np.array_equal(g.rows(), ref[3])
            g.free()


@pytest.mark.timeout(300)
def test_chains_rmat20_hash():
    G = json.load(open(os.path.join(GOLDEN, "hashes.json")))["rmat20_s1"]
    for spacing in (16, 4):
        with dcx.Context(0, chain_rows=spacing) as c:
            say(f"chains rmat20 spacing={spacing}")
            s, d = c.gen_rmat_device(20, 16 << 20, 1)
            g = c.ingest(1 << 20, s, d, 0, drop_self_loops=True, dedup=True)
            s.free()
            d.free()
            K, L = g.decompose(dcx.VARIANT_BATCHED)
            assert (K, L) == (G["kmax"], G["lmax"])
            assert f"{g.state_hash():016x}" == G["hash"]
            g.free()
