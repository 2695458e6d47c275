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


def _plan(active, capacity, quantum):
    srv = RenderServer.__new__(RenderServer)
    srv.active, srv.capacity, srv.quantum = active, capacity, quantum
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
