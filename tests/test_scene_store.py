"""The served-scene file format: a checked round trip of every array, and
rejection of files or arrays that do not describe one scene."""
import numpy as np
import pytest
import torch

from powerfoam import scene_store as store
from powerfoam import workloads
from powerfoam.adjacency import build_power_adjacency


@pytest.fixture(scope="module")
def arrays():
    sc = workloads.uniform_scene(256, seed=4)
    adj, off = build_power_adjacency(torch.from_numpy(sc["points"]), torch.from_numpy(sc["radii"]))
    N, K, E = len(sc["radii"]), sc["texel_sites"].shape[1], adj.numel()
    rng = np.random.default_rng(4)
    halves = torch.from_numpy(rng.standard_normal((E, 4)).astype(np.float16)).view(torch.int32)  # (E, 2)
    f = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32))
    return {
        "spheres": torch.cat([f(sc["points"]), f(sc["radii"])[:, None]], 1),
        "nsigmas": torch.cat([f(sc["normals"]), f(sc["density"])[:, None]], 1),
        "texel_sites": f(sc["texel_sites"]),
        "texel_height": f(sc["texel_height"]),
        "adjacency": adj.int(),
        "adjacency_offsets": off.int(),
        "walk_faces": torch.cat([halves, adj.int()[:, None]], 1).contiguous(),
        "sh": f(0.1 * rng.standard_normal((N, 16, 3))),
    }, (N, K, E)


def test_round_trip(arrays, tmp_path):
    a, _ = arrays
    p = tmp_path / "scene.safetensors"
    store.write_scene_arrays(p, a, "sh3")
    b, colour = store.read_scene_arrays(p)
    assert colour == "sh3" and set(a) == set(b)
    for k in a:
        assert a[k].dtype == b[k].dtype and torch.equal(a[k], b[k]), k


def test_rgb_colour(arrays, tmp_path):
    a, (N, K, E) = arrays
    a = dict(a)
    del a["sh"]
    a["texel_rgb"] = torch.rand(N, K, 3)
    p = tmp_path / "rgb.safetensors"
    store.write_scene_arrays(p, a, "rgb")
    b, colour = store.read_scene_arrays(p)
    assert colour == "rgb" and torch.equal(a["texel_rgb"], b["texel_rgb"])


@pytest.mark.parametrize("break_it", ["missing", "extra", "dtype", "shape", "offsets", "index", "neighbour"])
def test_rejects_inconsistent_arrays(arrays, break_it):
    a, (N, K, E) = arrays
    a = {k: v.clone() for k, v in a.items()}
    if break_it == "missing":
        del a["texel_height"]
    elif break_it == "extra":
        a["texel_rgb"] = torch.rand(N, K, 3)
    elif break_it == "dtype":
        a["adjacency"] = a["adjacency"].long()
    elif break_it == "shape":
        a["sh"] = a["sh"][:, :9]
    elif break_it == "offsets":
        a["adjacency_offsets"][N // 2] = E + 1
    elif break_it == "index":
        a["adjacency"][0] = N
        a["walk_faces"][0, 2] = N
    elif break_it == "neighbour":
        a["walk_faces"][3, 2] += 1
    with pytest.raises(ValueError):
        store.check_scene_arrays(a, "sh3")


def test_rejects_other_files(arrays, tmp_path):
    from safetensors.torch import save_file

    a, _ = arrays
    p = tmp_path / "other.safetensors"
    save_file({k: v.contiguous() for k, v in a.items()}, str(p), metadata={"format": "something-else"})
    with pytest.raises(ValueError):
        store.read_scene_arrays(p)
    with pytest.raises(ValueError):
        store.write_scene_arrays(tmp_path / "x.safetensors", a, "sh2")
