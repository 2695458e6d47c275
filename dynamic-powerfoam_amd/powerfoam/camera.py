"""Pinhole camera, the drop-in for ``powerfoam/camera.py``'s ``TorchCamera``.

Same dataclass fields and helper methods as the reference
(``powerfoam/camera.py:181-305``).  ``to_warp()`` (camera.py:191-205), which
rebuilt an (H, W, 6) ray map on every call, is replaced by ``to_pf()``: a
48-byte POD (``pf_camera`` in ``include/pf_hip.h``) from which the gfx950
kernels generate each pixel's ray in registers (camera.py:19-24).
"""
from dataclasses import dataclass

import numpy as np
import torch

PF_CAMERA_WORDS = 12  # pf_camera = 12 x 4 bytes


@dataclass
class TorchCamera:
    eye: torch.Tensor
    right: torch.Tensor
    up: torch.Tensor
    width: int
    height: int
    ray_maps: torch.Tensor = None
    fov_cos_cutoff: float = None

    def to_pf(self, start_cell=0):
        """Pack as pf_camera words (numpy int32 view of 12 x 4 bytes)."""
        return pack_cameras([self], [start_cell])[0]

    def to_device(self, device):
        dev = torch.device(device)
        return TorchCamera(
            eye=self.eye.to(dev),
            right=self.right.to(dev),
            up=self.up.to(dev),
            width=self.width,
            height=self.height,
            ray_maps=self.ray_maps.to(dev) if self.ray_maps is not None else None,
            fov_cos_cutoff=self.fov_cos_cutoff,
        )

    # camera.py:240-249
    def get_ray_dir(self, i, j):
        forward = torch.cross(self.up, self.right, dim=-1)
        forward = forward / torch.norm(forward, dim=-1, keepdim=True)
        x = 2.0 * j / (float(self.width) - 1.0) - 1.0
        y = 1.0 - 2.0 * i / (float(self.height) - 1.0)
        return x[:, None] * self.right[None, :] + y[:, None] * self.up[None, :] + forward[None, :]

    # camera.py:207-215
    def _build_pinhole_ray_maps(self):
        i = torch.arange(self.height, dtype=self.eye.dtype, device=self.eye.device)
        j = torch.arange(self.width, dtype=self.eye.dtype, device=self.eye.device)
        ii, jj = torch.meshgrid(i, j, indexing="ij")
        ray_dirs = self.get_ray_dir(ii.reshape(-1), jj.reshape(-1))
        ray_dirs = ray_dirs / torch.linalg.norm(ray_dirs, dim=-1, keepdim=True)
        ray_origins = self.eye[None, :].expand(ray_dirs.shape[0], 3)
        return torch.cat([ray_origins, ray_dirs], dim=-1).view(self.height, self.width, 6)

    # camera.py:251-263
    def intrinsics_matrix(self):
        right_norm = torch.linalg.norm(self.right) + 1e-8
        up_norm = torch.linalg.norm(self.up) + 1e-8
        fx = 0.5 * (float(self.width) - 1.0) / right_norm
        fy = 0.5 * (float(self.height) - 1.0) / up_norm
        cx = 0.5 * (float(self.width) - 1.0)
        cy = 0.5 * (float(self.height) - 1.0)
        return torch.tensor([[fx, 0.0, cx], [0.0, fy, cy], [0.0, 0.0, 1.0]], dtype=self.eye.dtype,
                            device=self.eye.device)

    # camera.py:265-282
    def w2c(self):
        r = self.right / (torch.linalg.norm(self.right) + 1e-8)
        u = self.up / (torch.linalg.norm(self.up) + 1e-8)
        f = torch.cross(u, r, dim=-1)
        f = f / torch.linalg.norm(f + 1e-8)
        u = torch.cross(f, r, dim=-1)
        u = u / torch.linalg.norm(u + 1e-8)
        R = torch.stack([r, u, f], dim=0)
        t = -R @ self.eye
        M = torch.eye(4, dtype=self.eye.dtype, device=self.eye.device)
        M[:3, :3] = R
        M[:3, 3] = t
        return M

    # camera.py:284-300
    def c2w(self):
        r = self.right / torch.linalg.norm(self.right)
        u = self.up / torch.linalg.norm(self.up)
        f = torch.cross(u, r, dim=-1)
        f = f / torch.linalg.norm(f)
        u = torch.cross(f, r, dim=-1)
        u = u / torch.linalg.norm(u)
        M = torch.eye(4, dtype=self.eye.dtype, device=self.eye.device)
        M[:3, :3] = torch.stack([r, u, f], dim=1)
        M[:3, 3] = self.eye
        return M

    def projection_matrix(self):
        return self.intrinsics_matrix() @ self.w2c()[:3, :]


def _vec3(t):
    if isinstance(t, torch.Tensor):
        return t.detach().to("cpu", torch.float32).numpy().reshape(3)
    return np.asarray(t, dtype=np.float32).reshape(3)


def pack_cameras(cameras, start_cells=None):
    """Pack cameras as a (B, 12) int32 array with pf_camera's layout."""
    B = len(cameras)
    out = np.zeros((B, PF_CAMERA_WORDS), dtype=np.int32)
    f = out.view(np.float32)
    for b, cam in enumerate(cameras):
        f[b, 0:3] = _vec3(cam.eye)
        f[b, 3:6] = _vec3(cam.right)
        f[b, 6:9] = _vec3(cam.up)
        out[b, 9] = 0 if start_cells is None else int(start_cells[b])
        out[b, 10] = int(cam.width)
        out[b, 11] = int(cam.height)
    return out


def cameras_to_device(cameras, start_cells=None, device="cuda"):
    """pf_camera array on the device (int32 tensor (B, 12))."""
    return torch.from_numpy(pack_cameras(cameras, start_cells)).to(device)


def look_at_camera(eye, target, world_up, width, height, hfov_deg=60.0, device="cpu"):
    """Pinhole camera with |right| = tan(hfov/2) and |up| = |right|(H-1)/(W-1)
    (the loaders' convention, data_loader/blender.py:59-65)."""
    eye = np.asarray(eye, dtype=np.float64)
    fwd = np.asarray(target, dtype=np.float64) - eye
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, np.asarray(world_up, dtype=np.float64))
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    tr = np.tan(np.deg2rad(hfov_deg) / 2.0)
    right = right * tr
    up = up * tr * (height - 1) / (width - 1)
    t = lambda v: torch.tensor(v, dtype=torch.float32, device=device)
    return TorchCamera(eye=t(eye), right=t(right), up=t(up), width=int(width), height=int(height))
