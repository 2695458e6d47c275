"""Host-side scheduling of the serving loop (no GPU): the step plan gives
every active request up to a quantum of tiles in arrival order, fills the
capacity with the oldest requests' remaining tiles, never exceeds the
capacity and never hands out a tile twice."""
import types

from powerfoam.serving import RenderServer, tiles_of


def _fake(ntiles_list):
    reqs = []
    for k, n in enumerate(ntiles_list):
        r = types.SimpleNamespace(ntiles=n, _next_tile=0, rid=k)
        reqs.append(r)
    return reqs


def _plan(active, capacity, quantum, oldest_share=0.0, bwd_weight=1, grouped=None):
    srv = RenderServer.__new__(RenderServer)
    srv.active, srv.capacity, srv.quantum = active, capacity, quantum
    srv.oldest_share, srv.bwd_weight = oldest_share, bwd_weight
    srv.backward, srv._records = "recorded", None
    if grouped is not None:  # (bwd_capacity, max_buckets, {scene: cells})
        srv.backward, srv._records = "grouped", []
        srv.bwd_capacity, srv.max_buckets = grouped[0], grouped[1]
        srv.scenes = {k: types.SimpleNamespace(num_cells=n) for k, n in grouped[2].items()}
    return srv._schedule()


def test_tiles_of():
    assert tiles_of(1920, 1080) == 240 * 135 and tiles_of(37, 29) == 5 * 4


def test_round_robin_then_oldest():
    a = _fake([32400, 1024, 4096])
    plan = _plan(a, 4096, 1024)
    takes = {p[0].rid: p[2] for p in plan}
    assert takes == {0: 2048, 1: 1024, 2: 1024}
    assert sum(takes.values()) == 4096


def test_every_tile_exactly_once():
    reqs = _fake([100, 7, 4096, 33, 1])
    seen = {r.rid: [] for r in reqs}
    active = list(reqs)
    while active:
        plan = _plan(active, 64, 16)
        assert 0 < sum(p[2] for p in plan) <= 64
        for r, b, take in plan:
            assert b == r._next_tile
            seen[r.rid].extend(range(b, b + take))
            r._next_tile = b + take
        active = [r for r in active if r._next_tile < r.ntiles]
    for r in reqs:
        assert seen[r.rid] == list(range(r.ntiles))


def test_oldest_share_and_backward_weight():
    # the oldest request may take up to oldest_share of the capacity first
    a = _fake([32400, 1024, 4096])
    takes = {p[0].rid: p[2] for p in _plan(a, 4096, 1024, oldest_share=0.5)}
    assert takes == {0: 2048, 1: 1024, 2: 1024}
    takes = {p[0].rid: p[2] for p in _plan(_fake([32400, 1024, 4096]), 4096, 1024, oldest_share=1.0)}
    assert takes == {0: 4096}
    # a forward+backward tile counts bwd_weight forward tiles of capacity and quantum
    b = _fake([32400, 1024])
    b[0].mode = "forward_backward"
    takes = {p[0].rid: p[2] for p in _plan(b, 4096, 1024, bwd_weight=4)}
    # 256 (quantum / 4) for the backward request, 1024 for the forward one,
    # then the remaining 2048 of capacity buys the oldest 512 more tiles
    assert takes == {0: 768, 1: 1024}


def test_grouped_backward_limits():
    # forward+backward tiles per step stay within bwd_capacity, and a step's
    # forward+backward requests' cells within max_buckets; forward requests
    # take the rest of the capacity
    a = _fake([5000, 5000, 5000, 3000])
    for r, sc in zip(a[:3], ("x", "y", "x")):
        r.mode, r.scene = "forward_backward", sc
    cells = {"x": 600, "y": 500}
    takes = {p[0].rid: p[2] for p in _plan(a, 4096, 1024, grouped=(1536, 1100, cells))}
    # request 0 (600 cells) and 1 (500) fit the 1100 buckets, request 2 waits
    assert 2 not in takes
    assert takes[0] + takes[1] <= 1536 and takes[0] == 1024 and takes[1] == 512
    assert takes[3] == 4096 - 1536
    # the same requests across steps: every tile exactly once, limits held
    seen = {r.rid: [] for r in a}
    active = list(a)
    while active:
        plan = _plan(active, 4096, 1024, grouped=(1536, 1100, cells))
        assert sum(p[2] for p in plan if getattr(p[0], "mode", "") == "forward_backward") <= 1536
        assert sum(cells[p[0].scene] for p in plan if getattr(p[0], "mode", "") == "forward_backward") <= 1100
        for r, b, take in plan:
            assert b == r._next_tile
            seen[r.rid].extend(range(b, b + take))
            r._next_tile = b + take
        active = [r for r in active if r._next_tile < r.ntiles]
    for r in a:
        assert seen[r.rid] == list(range(r.ntiles))


def test_steady_host_turns_the_collector_off_and_back():
    import gc

    from powerfoam.serving import steady_host

    gc.enable()
    with steady_host():
        assert not gc.isenabled()
    assert gc.isenabled()
    gc.disable()
    try:
        with steady_host():
            assert not gc.isenabled()
        assert not gc.isenabled()  # left as it was found
    finally:
        gc.enable()
