"""Writes tests/golden/walk_256.npz: a seeded 256-cell scene, its walk and
raster inputs, and the outputs that the checks pin:

* fp64_*: the FP64 envelope oracle (oracle/fp64.py) for every pixel of a
  16x16 camera: visited cells (power-diagram owners along the ray up to the
  transmittance stop), colour, log transmittance and depth.  These are
  independent of the C restatement and of fp16 face storage.
* walk_*: the CPU restatement (oracle/pf_oracle.cpp) of the walk with the
  fp16 faces: visit counts, FNV-1a-64 sequence hashes, colour, log T,
  depth, and the single-thread (fixed order) gradients of
  L = sum |C - T|^2 for the stored target T.
* walk32_*: the same walk on fp32 faces (what the FP64 owners pin).
* raster_*: the restatement's raster forward on the Cech adjacency.
* sv_*: the spherical-Voronoi pre-pass on the texel sites.

Run from the repository root: python tests/golden/make_golden.py
The reference implementation is not used: it needs NVIDIA Warp, which is
not installed here; these vectors come from the restatement and the FP64
oracle, which the tests pin by closed forms and finite differences.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "dynamic-powerfoam_amd")]

import torch  # noqa: E402

from oracle import fp64  # noqa: E402
from oracle import oracle as orc  # noqa: E402
from powerfoam import workloads  # noqa: E402
from powerfoam.adjacency import build_power_adjacency  # noqa: E402
from powerfoam.camera import pack_cameras  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "walk_256.npz")


def main():
    sc = workloads.uniform_scene(256, seed=11, w_max=4e-2)
    rng = np.random.default_rng(11)
    adj, off = build_power_adjacency(torch.from_numpy(sc["points"]), torch.from_numpy(sc["radii"]))
    adj, off = adj.numpy().astype(np.int32), off.numpy().astype(np.int32)
    cam = workloads.front_camera(16, 16, hfov_deg=50.0)
    start = orc.start_cell(sc["points"], sc["radii"], cam.eye.numpy())
    words = pack_cameras([cam], [start])
    faces = orc.walk_faces(sc["points"], sc["radii"], adj, off, f16=True)
    S = orc.Scene(sc["points"], sc["radii"], sc["density"], sc["normals"], sc["texel_sites"], sc["texel_rgb"],
                  sc["texel_height"], adj, off, faces)
    w = orc.walk_forward(S, words, seq_cap=64)
    faces32 = orc.walk_faces(sc["points"], sc["radii"], adj, off, f16=False)
    S32 = orc.Scene(sc["points"], sc["radii"], sc["density"], sc["normals"], sc["texel_sites"], sc["texel_rgb"],
                    sc["texel_height"], adj, off, faces32)
    w32 = orc.walk_forward(S32, words, seq_cap=64)
    target = rng.uniform(0, 1, (1, 16, 16, 3)).astype(np.float32)
    gc = (2.0 * (w["color"] - target)).astype(np.float32)
    g = orc.walk_backward(S, words, w["color"], gc, nthreads=1)

    ref = fp64.render(sc, words[0])
    seq = np.full((256, 64), -1, np.int32)
    for k, s in enumerate(ref["seq"]):
        seq[k, :min(len(s), 64)] = s[:64]

    cadj, coff = orc.cech(np.ascontiguousarray(sc["points"]), sc["radii"])
    spheres = np.concatenate([sc["points"], sc["radii"][:, None]], 1).astype(np.float32)
    rfaces = orc.raster_faces(spheres, cadj, coff, f16=True)
    rs = orc.RasterScene(sc["points"], sc["radii"], sc["density"], sc["normals"], sc["texel_sites"],
                         sc["texel_rgb"], sc["texel_height"], cadj, coff, rfaces)
    boff, prims = orc.raster_bin(spheres, words[0])
    rf = orc.raster_forward(rs, words[0], boff, prims)

    dof = 3
    ax = rng.standard_normal((dof, 256 * 8, 3)).astype(np.float32)
    ax /= np.linalg.norm(ax, axis=-1, keepdims=True)
    vals = (0.3 * rng.standard_normal((dof, 256 * 8, 3))).astype(np.float32)
    temps = rng.uniform(0.5, 4.0, (dof, 256 * 8)).astype(np.float32)
    eye = cam.eye.numpy().astype(np.float32)
    fwd = np.cross(cam.up.numpy(), cam.right.numpy()).astype(np.float32)
    fwd /= np.linalg.norm(fwd)
    cut = 0.5
    sv_v, sv_w = orc.sv_forward(sc["texel_sites"], eye, fwd, cut, ax, vals, temps)

    np.savez_compressed(
        OUT,
        points=sc["points"], radii=sc["radii"], density=sc["density"], normals=sc["normals"],
        texel_sites=sc["texel_sites"], texel_rgb=sc["texel_rgb"], texel_height=sc["texel_height"],
        adjacency=adj, offsets=off, cam_words=words, faces_f16=faces, target=target,
        fp64_seq=seq, fp64_color=ref["color"].reshape(16, 16, 3), fp64_log_t=ref["log_t"].reshape(16, 16),
        fp64_depth=ref["depth"].reshape(16, 16), fp64_margin=ref["margin"].reshape(16, 16),
        walk_visits=w["visits"], walk_seq_hash=w["seq_hash"], walk_seq=w["seq"], walk_color=w["color"],
        walk_log_t=w["log_t"], walk_depth=w["depth"],
        walk32_visits=w32["visits"], walk32_seq=w32["seq"], walk32_color=w32["color"], walk32_log_t=w32["log_t"],
        grad_spheres=g["spheres"], grad_nsigmas=g["nsigmas"], grad_texel_sites=g["texel_sites"],
        grad_texel_rgb=g["texel_rgb"], grad_texel_height=g["texel_height"],
        cech_adjacency=cadj, cech_offsets=coff, raster_color=rf["color"], raster_log_t=rf["log_t"],
        sv_axes=ax, sv_values=vals, sv_temps=temps, sv_eye=eye, sv_forward=fwd, sv_cutoff=np.float32(cut),
        sv_value=sv_v, sv_wsum=sv_w,
    )
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
