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
    faces = torch.from_numpy(rng.standard_normal((E, 4)).astype(np.float16))
    f = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32))
    return {
        "spheres": torch.cat([f(sc["points"]), f(sc["radii"])[:, None]], 1),
        "nsigmas": torch.cat([f(sc["normals"]), f(sc["density"])[:, None]], 1),
        "texel_sites": f(sc["texel_sites"]),
        "texel_height": f(sc["texel_height"]),
        "adjacency": adj.int(),
        "adjacency_offsets": off.int(),
        "walk_faces": faces,
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


@pytest.mark.parametrize("break_it", ["missing", "extra", "dtype", "shape", "offsets", "index", "faces"])
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
    elif break_it == "faces":
        a["walk_faces"] = a["walk_faces"].float()
    with pytest.raises(ValueError):
        store.check_scene_arrays(a, "sh3")


def _packed(a):
    """The powerfoam-scene/1 arrays: faces as (E, 3) int32 records."""
    p = dict(a)
    p["walk_faces"] = torch.cat([a["walk_faces"].view(torch.int32), a["adjacency"][:, None]], 1).contiguous()
    return p


def test_reads_packed_face_files(arrays, tmp_path):
    """Files of the earlier format (faces packed with their neighbour) are
    read with their faces unpacked to the same fp16 values."""
    from safetensors.torch import save_file

    a, _ = arrays
    p = tmp_path / "v1.safetensors"
    save_file(_packed(a), str(p), metadata={"format": store.FORMAT_PACKED, "colour": "sh3"})
    b, colour = store.read_scene_arrays(p)
    assert colour == "sh3" and b["walk_faces"].dtype == torch.float16
    assert torch.equal(b["walk_faces"], a["walk_faces"])


@pytest.mark.parametrize("break_it", ["neighbour", "unpacked"])
def test_rejects_inconsistent_packed_arrays(arrays, break_it):
    a, _ = arrays
    p = _packed(a)
    if break_it == "neighbour":
        p["walk_faces"][3, 2] += 1
    else:
        p["walk_faces"] = a["walk_faces"]
    with pytest.raises(ValueError):
        store.check_scene_arrays(p, "sh3", packed=True)


def test_rejects_other_files(arrays, tmp_path):
    from safetensors.torch import save_file

    a, _ = arrays
    p = tmp_path / "other.safetensors"
    save_file({k: v.contiguous() for k, v in a.items()}, str(p), metadata={"format": "something-else"})
    with pytest.raises(ValueError):
        store.read_scene_arrays(p)
    with pytest.raises(ValueError):
        store.write_scene_arrays(tmp_path / "x.safetensors", a, "sh2")
