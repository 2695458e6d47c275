"""On-disk format for served scenes.

One safetensors file holds every array a resident scene reads, the
power-diagram CSR and the fp16 walk faces included, so that bringing a
scene into a server is a file read and an upload: no Qhull lift or GPU
adjacency build, no face precompute (benchmark.py:25-83, :295-302 run once,
when the file is written).  safetensors stores raw little-endian arrays and
a JSON header; loading executes nothing from the file.

Arrays (N cells, K texels, E directed edges):

=================  ===========  ====================================
name               dtype        shape
=================  ===========  ====================================
spheres            float32      (N, 4)  [p, r]
nsigmas            float32      (N, 4)  [n, sigma]
texel_sites        float32      (N, K, 3)
texel_height       float32      (N, K)
adjacency          int32        (E,)
adjacency_offsets  int32        (N + 1,)
walk_faces         float16      (E, 4)  the walk faces (benchmark.py:295-302)
sh                 float32      (N, 16, 3)  when colour is "sh3"
texel_rgb          float32      (N, K, 3)   when colour is "rgb"
=================  ===========  ====================================

Header metadata: ``format`` = ``powerfoam-scene/2``, ``colour``.  Files of
``powerfoam-scene/1`` (walk_faces as int32 (E, 3) records: the four fp16
values, then the neighbour) are read too, their faces unpacked on reading.
"""
import torch

FORMAT = "powerfoam-scene/2"
FORMAT_PACKED = "powerfoam-scene/1"

_COMMON = {
    "spheres": (torch.float32, lambda N, K, E: (N, 4)),
    "nsigmas": (torch.float32, lambda N, K, E: (N, 4)),
    "texel_sites": (torch.float32, lambda N, K, E: (N, K, 3)),
    "texel_height": (torch.float32, lambda N, K, E: (N, K)),
    "adjacency": (torch.int32, lambda N, K, E: (E,)),
    "adjacency_offsets": (torch.int32, lambda N, K, E: (N + 1,)),
    "walk_faces": (torch.float16, lambda N, K, E: (E, 4)),
}
_COLOUR = {
    "sh3": ("sh", torch.float32, lambda N, K, E: (N, 16, 3)),
    "rgb": ("texel_rgb", torch.float32, lambda N, K, E: (N, K, 3)),
}


def check_scene_arrays(arrays, colour, packed=False):
    """Raise ValueError unless arrays hold exactly the format's tensors with
    consistent shapes and dtypes (CSR offsets included); packed: the
    powerfoam-scene/1 walk faces."""
    if colour not in _COLOUR:
        raise ValueError(f"unknown colour mode {colour!r}")
    name, dtype, shape = _COLOUR[colour]
    spec = dict(_COMMON)
    spec[name] = (dtype, shape)
    if packed:
        spec["walk_faces"] = (torch.int32, lambda N, K, E: (E, 3))
    missing = sorted(set(spec) - set(arrays))
    extra = sorted(set(arrays) - set(spec))
    if missing or extra:
        raise ValueError(f"scene arrays: missing {missing}, unexpected {extra}")
    N = arrays["spheres"].shape[0]
    K = arrays["texel_sites"].shape[1] if arrays["texel_sites"].dim() == 3 else -1
    E = arrays["adjacency"].shape[0]
    for key, (dt, shape) in spec.items():
        t = arrays[key]
        if t.dtype != dt or tuple(t.shape) != shape(N, K, E):
            raise ValueError(f"scene array {key}: {t.dtype} {tuple(t.shape)}, expected {dt} {shape(N, K, E)}")
    off = arrays["adjacency_offsets"]
    if int(off[0]) != 0 or int(off[-1]) != E or bool((off[1:] < off[:-1]).any()):
        raise ValueError("adjacency_offsets is not a CSR offset array over the adjacency")
    adj = arrays["adjacency"]
    if E and (int(adj.min()) < 0 or int(adj.max()) >= N):
        raise ValueError("adjacency holds an index out of range")
    # packed records carry the neighbour in their last word
    if packed and not torch.equal(arrays["walk_faces"][:, 2], adj):
        raise ValueError("walk_faces neighbours disagree with the adjacency")


def write_scene_arrays(path, arrays, colour):
    """Check and write arrays (host or device tensors) to path."""
    from safetensors.torch import save_file

    host = {k: v.detach().to("cpu").contiguous() for k, v in arrays.items()}
    check_scene_arrays(host, colour)
    save_file(host, str(path), metadata={"format": FORMAT, "colour": colour})


def read_scene_arrays(path, device="cpu"):
    """(arrays, colour) from path, checked, on device."""
    from safetensors import safe_open

    with safe_open(str(path), framework="pt", device="cpu") as f:
        meta = f.metadata() or {}
        if meta.get("format") not in (FORMAT, FORMAT_PACKED):
            raise ValueError(f"{path}: not a {FORMAT} file (format {meta.get('format')!r})")
        packed = meta.get("format") == FORMAT_PACKED
        colour = meta.get("colour")
        arrays = {k: f.get_tensor(k) for k in f.keys()}
    check_scene_arrays(arrays, colour, packed)
    if packed:
        arrays["walk_faces"] = arrays["walk_faces"][:, :2].contiguous().view(torch.float16).view(-1, 4)
    if device != "cpu":
        arrays = {k: v.to(device, non_blocking=True) for k, v in arrays.items()}
    return arrays, colour


def save_scene(path, scene):
    """Write a serving.ResidentScene (f32 attributes) to path."""
    w = scene.walk
    if w.nsigmas.dtype != torch.float32:
        raise ValueError("the scene format stores f32 attributes")
    arrays = {
        "spheres": w.spheres,
        "nsigmas": w.nsigmas,
        "texel_sites": w.texel_sites,
        "texel_height": w.texel_height,
        "adjacency": w.adjacency,
        "adjacency_offsets": w.offsets,
        "walk_faces": w.faces,
    }
    if scene.sh is not None:
        arrays["sh"] = scene.sh
        colour = "sh3"
    else:
        arrays["texel_rgb"] = w.texel_rgb
        colour = "rgb"
    write_scene_arrays(path, arrays, colour)


def load_scene(path, device="cuda"):
    """A serving.ResidentScene read from path onto device."""
    from .serving import ResidentScene

    a, colour = read_scene_arrays(path, device)
    s, ns = a["spheres"], a["nsigmas"]
    return ResidentScene(s[:, :3], s[:, 3], ns[:, 3], ns[:, :3], a["texel_sites"], a["texel_height"],
                         a["adjacency"], a["adjacency_offsets"], a["walk_faces"], sh=a.get("sh"),
                         texel_rgb=a.get("texel_rgb"))
