"""dcx_klist_noalloc, the allocation-free variant of dcx_klist: its pass
scratch is a static device array (at most 262,144 vertices) and its output
a caller-allocated device array.  Its row must equal dcx_klist's, bit for
bit, on two calls in a row -- the second on a smaller graph, so the static
array still holds the first call's words -- and a graph above its limit is
refused with DCX_ERR_CAPACITY."""
import sys

import numpy as np
import pytest

import dcx

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def test_klist_noalloc_equals_klist():
    with dcx.Context(0) as c:
        for scale, seed in ((17, 3), (12, 5)):  # the second call is smaller
            V = 1 << scale
            s, d = dcx.gen_rmat(scale, 16 << scale, seed)
            g = c.upload_edges(V, s, d)
            out = dcx.DeviceArray.empty(V)
            for k in (0, 1, 7):
                say(f"noalloc rmat{scale} k={k}")
                lk_ref, ref = g.klist(k, dcx.VARIANT_PEEL)
                lk = g.klist_noalloc(k, out)
                got = out.numpy().view(np.int32)
                assert lk == lk_ref, (scale, k)
                assert np.array_equal(got, ref), (scale, k)
            out.free()
            g.free()
        say("noalloc refuses a graph above its limit")
        V = (1 << 18) + 1
        g = c.upload_edges(V, np.array([0, 1], np.uint32), np.array([1, 0], np.uint32))
        out = dcx.DeviceArray.empty(V)
        with pytest.raises(dcx.DcxError) as e:
            g.klist_noalloc(0, out)
        assert e.value.status == 5  # DCX_ERR_CAPACITY
        out.free()
        g.free()
