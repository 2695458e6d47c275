"""A served scene written to the scene file and read back renders the same
images and gradients as the scene it came from."""
import numpy as np
import pytest
import torch

from powerfoam import raytrace as rt
from powerfoam import scene_store as store
from powerfoam import workloads
from powerfoam.power import build_power_adjacency_gpu
from powerfoam.serving import PlainRenderer, RenderRequest, ResidentScene

pytestmark = pytest.mark.gpu
DEV = "cuda"


@pytest.mark.parametrize("colour", ["sh3", "rgb"])
def test_saved_scene_renders_the_same(scene4k, tmp_path, colour):
    sc = scene4k
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    pts, rad = t(sc["points"]), t(sc["radii"])
    adj, off = build_power_adjacency_gpu(pts, rad)
    faces = rt.walk_faces(pts, rad, adj, off)
    rng = np.random.default_rng(21)
    extra = {"sh": t((0.1 * rng.standard_normal((len(sc["radii"]), 16, 3))).astype(np.float32))} if colour == "sh3" \
        else {"texel_rgb": t(sc["texel_rgb"])}
    scene = ResidentScene(pts, rad, t(sc["density"]), t(sc["normals"]), t(sc["texel_sites"]), t(sc["texel_height"]),
                          adj, off, faces, **extra)
    path = tmp_path / "scene.safetensors"
    store.save_scene(path, scene)
    loaded = store.load_scene(path, DEV)
    assert loaded.nbytes() == scene.nbytes()
    target = torch.from_numpy(rng.uniform(0, 1, (48, 64, 3)).astype(np.float32))
    outs = []
    for s in (scene, loaded):
        r = RenderRequest(scene="s", camera=workloads.front_camera(64, 48), mode="forward_backward", target=target,
                          rid=0)
        PlainRenderer({"s": s}).render(r)
        outs.append(r)
    torch.cuda.synchronize()
    a, b = outs
    assert torch.equal(a.color, b.color) and torch.equal(a.log_t, b.log_t)
    # gradients accumulate with float atomics in no fixed order: 1e-5 relative L2
    for k in a.grads:
        if a.grads[k] is not None:
            x, y = a.grads[k].double(), b.grads[k].double()
            assert torch.linalg.norm(x - y) <= 1e-5 * torch.linalg.norm(y) + 1e-12, k
