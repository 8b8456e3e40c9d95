"""Serving: requests admitted into a continuous batch produce exactly the
result they produce alone (same tiles, same arithmetic per tile)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pool():
    from setsm_amd.serve import make_pool
    return make_pool(n_scenes=2, image_px=1024, device="cuda", seed=7)


def test_alone_equals_mixed(require_gpu, pool):
    from setsm_amd import smx
    from setsm_amd.serve import Request, Server
    reqs = [Request(0, 0.0, 0, 90, 48, 11), Request(1, 0.0, 1, 60, 96, 12), Request(2, 0.01, 0, 120, 32, 13)]
    mixed, st = Server(pool, levels=3).run(reqs)
    print("mixed:", {k: v for k, v in st.items()}, flush=True)
    by_id = {d["rid"]: d for d in mixed}
    for r in reqs:
        alone, _ = Server(pool, levels=3).run([Request(r.rid, 0.0, r.scene, r.cells, r.nheights, r.seed, r.plan)])
        a, m = alone[0]["ncc"], by_id[r.rid]["ncc"]
        assert alone[0]["evaluated"] == by_id[r.rid]["evaluated"] > 0
        assert np.array_equal(a, m), f"request {r.rid} differs alone vs mixed"


def test_plain_baseline_serves(require_gpu, pool):
    from setsm_amd import smx
    from setsm_amd.serve import Server, make_stream
    stream = make_stream(4, rate=50.0, sizes=(32, 96), heights=(16, 48), n_scenes=2, seed=3)
    done_f, sf = Server(pool, levels=3).run(stream)
    done_p, sp = Server(pool, levels=3, precision=smx.PREC_FP64, batching=False).run(stream)
    assert sf["requests"] == sp["requests"] == 4
    ef = {d["rid"]: d["evaluated"] for d in done_f}
    ep = {d["rid"]: d["evaluated"] for d in done_p}
    assert ef == ep
    print("fast:", sf, "plain:", sp, flush=True)
