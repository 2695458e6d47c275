"""Serving: resident scenes, continuous batching over tiles, memory admission.

``RenderServer`` keeps scenes resident in HBM and renders a stream of
requests (one pinhole camera each, any resolution, forward or
forward+backward against a target image).  Work advances in steps: every
step is one graph replay of fixed-shape launches (``pf_serve_step_grouped``
or ``pf_serve_step``) over up to ``tile_capacity`` 8x8 tiles drawn from all
active requests, so requests that arrive while others are on the GPU join
at the next step and their tiles run in the same launches.  By default the
oldest request may take the whole capacity and the others fill what it
leaves, in arrival order (``oldest_share=1``: first come, first served,
with backfill); ``oldest_share=0`` gives every active request up to
``quantum`` tiles first (round robin).  On the multi-scene trace the first
order finished requests sooner and freed their memory for admission:
saturation 230 against 198 Mrays/s, p99 at load 0.6 37.5 against 39.2 ms
(mean of three traces).  The request table lives
in pinned host memory; each step copies it to the device and launches the
kernels, and both are captured once in a HIP graph per table buffer.  The
buffers rotate (``streams``, default 3), so the host fills one while the
GPU runs the others, and each has its own default-priority stream, so one
step's last tiles overlap the next step's first.  A process gets four
hardware queues by default (GPU_MAX_HW_QUEUES), one of which serves torch's
current stream (admission work): with three server streams every stream
has a queue of its own.  With two, some runs had both on one queue and the
mixed trace's p99 doubled; with four, two streams shared a queue in some
runs (multi-scene p99 67-75 ms against 47-54 ms with three).

Forward+backward requests (``backward="grouped"``, the default): their
tiles in a step record each ray's contributing segments with the
transmittance before and colour after each, and the step sorts the
segments of all its forward+backward requests by (request, cell) on the
device and differentiates them in that order, summing a cell's terms
across the wave before one atomic per run, as the plain path's grouped walk
backward does for one image.  ``backward="recorded"`` differentiates each
tile from its own record through a per-wave LDS cache of gradient rows.

Colour: with ``colour="prepass"`` (default) each admitted request of an
SH-coloured scene gets its own texel colours from one SH pre-pass for its
camera, which the step kernel then reads like the plain path does; with
``colour="walk"`` the walk evaluates the SH at every shaded segment and no
per-request colour buffer is needed (less memory, more arithmetic).  A
scene coloured by spherical-Voronoi lobes (a trained reference checkpoint,
``ResidentScene.from_powerfoam``) always gets the per-request pre-pass
(``SphericalVoronoi.forward``, color_fn.py:188-246, as PowerfoamScene
colours its texels before a render, scene.py:331-335).

Admission: a request reserves its output buffers (its colour buffer, and
gradient buffers for forward+backward) against ``memory_budget``; it
waits while the budget is full and is rejected if it waited longer than
``max_wait`` seconds after ``submit`` or can never fit.  ``PlainRenderer`` is the per-request drop-in path used as the
latency baseline: SH colour pre-pass + walk (+ backward), one request at a
time.
"""
import contextlib
import ctypes
import gc
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .camera import cameras_to_device, pack_cameras
from .color_fn import sv_prepass, sv_prepass_backward
from .raytrace import (DEFAULT_RECORD_CAP, WalkScene, launch_walk_backward, launch_walk_forward, start_cells_device,
                       walk_faces)
from .shading import sh_backward, sh_forward

TILE = 8


def tiles_of(w, h):
    return ((w + TILE - 1) // TILE) * ((h + TILE - 1) // TILE)


@dataclass
class RenderRequest:
    scene: str
    camera: object  # TorchCamera
    mode: str = "forward"  # or "forward_backward"
    target: torch.Tensor = None  # (H, W, 3) for forward_backward
    arrival: float = 0.0
    rid: int = 0
    # filled by the server
    color: torch.Tensor = None
    log_t: torch.Tensor = None
    grads: dict = None
    colour_grads: object = None  # PlainRenderer: the colour pre-pass's gradients
    admitted: float = None
    done: float = None
    rejected: bool = False
    _next_tile: int = 0
    _bytes: int = 0
    _submitted: float = 0.0
    _cam_dev: torch.Tensor = None
    _texel_rgb: torch.Tensor = None
    _sv_ctx: object = None
    _events: list = field(default_factory=list)

    @property
    def width(self):
        return self.camera.width

    @property
    def height(self):
        return self.camera.height

    @property
    def ntiles(self):
        return tiles_of(self.width, self.height)


_STREAMS = {}


def server_streams(device, n):
    """The process's server streams of a device, created once and shared by
    every RenderServer: HIP maps a new stream to the next of the process's
    hardware queues, so servers that made their own streams landed on
    different queue assignments, some with two server streams on one queue
    (multi-scene p99 57-64 ms against 49-52 ms, saturation 176-216 against
    198-230 Mrays/s on the same traces)."""
    dev = torch.device(device)
    idx = dev.index if dev.index is not None else torch.cuda.current_device()
    pool = _STREAMS.setdefault(idx, [])
    while len(pool) < n:
        pool.append(torch.cuda.Stream(device=idx))  # default priority
    return pool[:n]


class ResidentScene:
    """Device arrays of one scene (f32 attributes, (E, 4) fp16 walk faces, K = 8)
    plus its colour: fixed texel_rgb, degree-3 SH coefficients (turned into
    texel colours by a per-request pre-pass, or evaluated inside the walk),
    or spherical-Voronoi lobes ``sv = (axes (D, N*K, 3), values (D, N*K, 3),
    temperatures (D, N*K))`` in the lobe-major layout of
    PowerfoamScene.get_att_sv (scene.py:290-312), coloured per request."""

    def __init__(self, points, radii, density, normals, texel_sites, texel_height, adjacency, adjacency_offsets,
                 faces, sh=None, texel_rgb=None, sv=None):
        if sh is None and texel_rgb is None and sv is None:
            raise ValueError("a scene needs SH coefficients, spherical-Voronoi lobes or texel_rgb")
        K = texel_sites.shape[1]
        if K != 8 or faces.dtype != torch.float16 or faces.dim() != 2 or faces.shape[1] != 4:
            raise ValueError("served scenes use K = 8 texels and (E, 4) fp16 walk faces")
        # with SH colour the kernel never reads texel_rgb; texel_sites (same
        # shape) stands in so the scene struct holds a valid pointer
        rgb = texel_rgb if texel_rgb is not None else texel_sites
        self.walk = WalkScene(points, radii, density, normals, texel_sites, rgb, texel_height, adjacency,
                              adjacency_offsets, faces, pack_faces=False)
        self.points = points.detach().float().contiguous()
        self.radii = radii.detach().float().contiguous()
        self.sh = sh.detach().float().contiguous() if sh is not None else None
        if self.sh is not None and self.sh.shape != (points.shape[0], 16, 3):
            raise ValueError("SH coefficients must be (N, 16, 3) (degree 3)")
        self.num_cells = points.shape[0]
        self.K = K
        self.sv = None
        if sv is not None:
            axes, vals, temps = (t.detach().float().contiguous() for t in sv)
            D, nk = axes.shape[0], self.num_cells * K
            if axes.shape != (D, nk, 3) or vals.shape != (D, nk, 3) or temps.shape != (D, nk):
                raise ValueError("spherical-Voronoi lobes must be axes (D, N*K, 3), values (D, N*K, 3) and "
                                 "temperatures (D, N*K)")
            self.sv = (axes, vals, temps)
            self.texel_sites_flat = self.walk.texel_sites.reshape(-1, 3)
        s = _lib.PfServeScene()
        s.scene = self.walk.c
        s.sh = _lib.ptr(self.sh)
        s.sh_degree = 3 if self.sh is not None else -1
        self.c = s

    @classmethod
    def from_powerfoam(cls, scene, power_adjacency=None):
        """A resident copy of a PowerfoamScene, e.g. one read from a reference
        checkpoint with load_pt (scene.py:801-821): activated attributes in
        f32, world texel sites and heights (scene.py:324-337), the lobes for
        the per-request colour pre-pass, a power-diagram adjacency (built on
        the GPU unless given as (adjacency, offsets)) and its fp16 walk
        faces.  Served requests then equal PowerfoamScene.render_walk on the
        same adjacency."""
        from .power import build_power_adjacency_gpu

        with torch.no_grad():
            f32 = lambda t: t.detach().float().contiguous()
            pts, rad = f32(scene.points), f32(scene.get_radii())
            density, normals = f32(scene.get_density()), f32(scene.get_normals())
            sites, heights = (f32(t) for t in scene.world_texels())
            sv = tuple(f32(t) for t in scene.get_att_sv())
        if power_adjacency is None:
            adj, off = build_power_adjacency_gpu(pts, rad)
        else:
            adj, off = (t.to(pts.device).int().contiguous() for t in power_adjacency)
        faces = walk_faces(pts, rad, adj, off)
        return cls(pts, rad, density, normals, sites, heights, adj, off, faces, sv=sv)

    def nbytes(self):
        w = self.walk
        tensors = [w.spheres, w.nsigmas, w.texel_sites, w.texel_height, w.adjacency, w.offsets, w.faces]
        if self.sh is not None:
            tensors.append(self.sh)
        elif self.sv is not None:
            tensors.extend(self.sv)
        else:
            tensors.append(w.texel_rgb)
        return sum(t.numel() * t.element_size() for t in tensors)

    def colour_prepass(self, camera, cams_dev, keep_for_backward=False):
        """This camera's texel colours (N, K, 3) from the SH or lobes, or None
        for a scene with fixed texel_rgb.  For lobes with keep_for_backward,
        returns (colours, context of sv_prepass_backward)."""
        if self.sv is not None:
            out = sv_prepass(self.texel_sites_flat, camera, *self.sv, keep_for_backward=keep_for_backward)
            rgb, ctx = out if keep_for_backward else (out, None)
            rgb = rgb.view(self.num_cells, self.K, 3)
            return (rgb, ctx) if keep_for_backward else rgb
        if self.sh is not None:
            rgb = sh_forward(self.sh, 3, self.walk.texel_sites, cams_dev)[0]
            return (rgb, None) if keep_for_backward else rgb
        return (None, None) if keep_for_backward else None

    def grad_bytes(self):
        n, K = self.num_cells, self.K
        return 4 * n * (4 + 4 + 3 * K + 3 * K + K)


class RenderServer:
    def __init__(self, device="cuda", memory_budget=64 << 30, tile_capacity=16384, quantum=1024, use_graph=True,
                 max_wait=2.0, transmittance_threshold=1e-3, max_steps=1 << 16, colour="prepass", bwd_weight=1,
                 oldest_share=1.0, streams=3, record_cap=128, backward="grouped", bwd_capacity=None,
                 max_buckets=1 << 24, entries_per_slot=None):
        if colour not in ("prepass", "walk"):
            raise ValueError("colour must be 'prepass' or 'walk'")
        if backward not in ("grouped", "recorded"):
            raise ValueError("backward must be 'grouped' or 'recorded'")
        if not 0 < int(tile_capacity) < 1 << 26 or int(quantum) <= 0 or int(streams) < 1:
            raise ValueError("tile_capacity must be in [1, 2^26), quantum positive and streams at least 1")
        self.colour = colour
        self.device = device
        self.budget = int(memory_budget)
        self.capacity = int(tile_capacity)
        self.quantum = int(quantum)
        self.bwd_weight = max(int(bwd_weight), 1)
        self.oldest_share = float(oldest_share)
        self.on_finish = None  # called with each finished request (e.g. to consume and drop its gradients)
        self.use_graph = use_graph
        self.max_wait = max_wait
        self.thr = float(transmittance_threshold)
        self.max_steps = int(max_steps)
        self.scenes = {}
        self.scene_index = {}
        self.reserved = 0
        self.queue = deque()
        self.active = []
        self.inflight = deque()  # (event, [requests completing in this step])
        self.finished = []
        self.rejected = []
        self.nbuf = int(streams)
        self.max_reqs = 256
        self._tables_dirty = True
        self._graphs = None
        self._step_count = 0
        self.status = torch.zeros(1, dtype=torch.int32, device=device)
        rq = ctypes.sizeof(_lib.PfServeRequest)
        self._req_bytes = rq
        self._host_tab = [torch.empty(self.max_reqs * rq + 32, dtype=torch.uint8, pin_memory=True)
                          for _ in range(self.nbuf)]
        self._dev_tab = [torch.empty(self.max_reqs * rq + 32, dtype=torch.uint8, device=device)
                         for _ in range(self.nbuf)]
        self._events = [None] * self.nbuf
        # one stream per table buffer (default priority): a step's tail overlaps
        # the next step's start
        self._streams = server_streams(device, self.nbuf)
        self._flags = _lib.PF_SERVE_SH_IN_WALK if colour == "walk" else 0
        # forward+backward tiles record their contributing segments (up to
        # record_cap per ray, a multiple of 8; 0 replays the walk) in one
        # workspace per table buffer, allocated with the first such request:
        # 16 * record_cap * 64 bytes per tile of capacity and buffer (6.4 GB
        # at the defaults), reserved from memory_budget
        self.record_cap = -(-int(record_cap) // 8) * 8 if record_cap > 0 else 0
        self._records = None
        self.record_bytes = 0
        # backward="grouped": a step's forward+backward tiles (at most
        # bwd_capacity, the whole capacity by default) record their segments
        # with partials, and the step differentiates all of them sorted by
        # (request, cell) (pf_serve_step_grouped); a step's forward+backward
        # requests number their cells from bucket0, at most max_buckets in
        # all.  Workspace per table buffer: 32 * record_cap + 4 bytes per
        # slot (64 per tile of bwd_capacity), 64 bytes per sorted entry
        # (min(record_cap, 48) per slot) and 8 per bucket: 7.6 GB at the
        # defaults.  A quarter of the capacity held large forward+backward
        # requests to many steps (multi-scene p99 of those requests 102 ms
        # against 46 ms).  It needs colours from memory, so colour="walk"
        # and record_cap 0 differentiate tile by tile ("recorded").
        self.backward = backward if (colour == "prepass" and self.record_cap > 0) else "recorded"
        self.bwd_capacity = int(bwd_capacity) if bwd_capacity else self.capacity
        if not 0 < self.bwd_capacity <= self.capacity:
            raise ValueError("bwd_capacity must be in [1, tile_capacity]")
        self.max_buckets = int(max_buckets)
        # sorted entries the grouped workspace holds per slot (rays past the
        # budget are replayed): min(record_cap, 48) by default
        self.entries_per_slot = entries_per_slot

    # ------------------------------------------------------------ scenes --
    def add_scene(self, name, scene: ResidentScene):
        if name in self.scenes:
            # requests and the scene table refer to a resident scene by its index
            raise ValueError(f"scene {name!r} is already resident")
        if self.backward == "grouped" and scene.num_cells > self.max_buckets:
            raise ValueError(f"scene {name!r} has more cells ({scene.num_cells}) than max_buckets")
        need = scene.nbytes()
        if self.reserved + need > self.budget:
            raise MemoryError(f"scene {name} ({need} B) does not fit in the memory budget")
        self.reserved += need
        self.scene_index[name] = len(self.scenes)
        self.scenes[name] = scene
        n = len(self.scenes)
        arr = (_lib.PfServeScene * n)(*[s.c for s in self.scenes.values()])
        raw = torch.frombuffer(bytearray(bytes(arr)), dtype=torch.uint8)
        new_tab = raw.to(self.device)
        # steps in flight on the server's streams read the old table: it is
        # released only once they have finished
        for ev in self._events:
            if ev is not None:
                ev.synchronize()
        self._scene_tab = new_tab
        self._graphs = None  # the scene table pointer changed

    # ---------------------------------------------------------- requests --
    def _prepass(self, r):
        sc = self.scenes[r.scene]
        return sc.sv is not None or (self.colour == "prepass" and sc.sh is not None)

    def _request_bytes(self, r):
        b = r.width * r.height * 16
        if self._prepass(r):
            sc = self.scenes[r.scene]
            # colours, and the lobes' weight sums kept for their backward
            b += sc.num_cells * sc.K * (16 if sc.sv is not None else 12)
        if r.mode == "forward_backward":
            b += self.scenes[r.scene].grad_bytes() + r.width * r.height * 12
        return b

    def _ensure_records(self):
        """The forward+backward tiles' record workspaces (one per table
        buffer), allocated with the first such request and counted against
        the memory budget.  Grouped: segments and partials of 64 slots per
        tile of bwd_capacity, counts and the sort's workspace; recorded:
        segments of 64 slots per tile of capacity.  If they do not fit,
        those tiles replay their walks instead (record_cap 0)."""
        if self._records is not None or self.record_cap <= 0:
            return
        lib = _lib.load()
        grouped = self.backward == "grouped"
        if grouped:
            slots = 64 * self.bwd_capacity
            ws = lib.pf_serve_grouped_workspace_bytes(self.bwd_capacity, self.record_cap, self.max_buckets)
            if ws <= 0:
                raise ValueError("bwd_capacity, record_cap or max_buckets out of range")
            if self.entries_per_slot is not None:  # a smaller sort budget
                dflt = min(self.record_cap, 48)
                ws -= 64 * slots * (dflt - max(min(int(self.entries_per_slot), dflt), 1))
            need = (32 * self.record_cap * slots + 4 * slots + ws) * self.nbuf
        else:
            slots = lib.pf_serve_record_slots(self.capacity)
            need = 16 * self.record_cap * slots * self.nbuf
        if self.reserved + need > self.budget:
            self.record_cap = 0
            self.backward = "recorded"
            return
        self.reserved += need
        self._records = []
        for _ in range(self.nbuf):
            segs = torch.empty(self.record_cap * slots, 4, dtype=torch.int32, device=self.device)
            rec = _lib.PfWalkRecord()
            rec.segments = segs.data_ptr()
            rec.cap = self.record_cap
            keep = [segs]
            if grouped:
                part = torch.empty(self.record_cap * slots, 4, dtype=torch.float32, device=self.device)
                counts = torch.empty(slots, dtype=torch.int32, device=self.device)
                work = torch.empty(ws, dtype=torch.uint8, device=self.device)
                rec.partials = part.data_ptr()
                rec.counts = counts.data_ptr()
                keep += [part, counts, work]
            self._records.append((keep, rec))
        self.record_bytes = need
        self._graphs = None  # the launches now carry the workspaces

    def submit(self, r: RenderRequest):
        if r.scene not in self.scenes:
            raise ValueError(f"scene {r.scene!r} is not resident")
        if r.mode not in ("forward", "forward_backward"):
            raise ValueError("mode must be 'forward' or 'forward_backward'")
        if r.mode == "forward_backward" and (r.target is None or tuple(r.target.shape) != (r.height, r.width, 3)):
            raise ValueError("a forward_backward request needs an (H, W, 3) target")
        if r.mode == "forward_backward":
            self._ensure_records()
        r._bytes = self._request_bytes(r)
        r._submitted = time.perf_counter()
        if r._bytes > self.budget - sum(s.nbytes() for s in self.scenes.values()) - self.record_bytes:
            r.rejected = True
            self.rejected.append(r)
            return False
        self.queue.append(r)
        return True

    def _admit(self, now):
        while self.queue and len(self.active) < self.max_reqs:
            r = self.queue[0]
            if self.reserved + r._bytes > self.budget:
                if now - r._submitted > self.max_wait:
                    self.queue.popleft()
                    r.rejected = True
                    self.rejected.append(r)
                    continue
                break
            self.queue.popleft()
            self.reserved += r._bytes
            sc = self.scenes[r.scene]
            dev = self.device
            r.color = torch.empty(r.height, r.width, 3, dtype=torch.float32, device=dev)
            r.log_t = torch.empty(r.height, r.width, dtype=torch.float32, device=dev)
            if r.mode == "forward_backward":
                n, K = sc.num_cells, sc.K
                r.grads = dict(spheres=torch.zeros(n, 4, device=dev), nsigmas=torch.zeros(n, 4, device=dev),
                               texel_sites=torch.zeros(n, K, 3, device=dev),
                               texel_rgb=torch.zeros(n, K, 3, device=dev), texel_height=torch.zeros(n, K, device=dev))
                r.target = r.target.to(dev, torch.float32).contiguous()
            r._cam_dev = cameras_to_device([r.camera], None, dev)
            start_cells_device(sc.points, sc.radii, r._cam_dev)
            if self._prepass(r):  # this camera's texel colours, once, ahead of its tiles on the same stream
                r._texel_rgb, r._sv_ctx = sc.colour_prepass(r.camera, r._cam_dev, keep_for_backward=True)
                if r.mode != "forward_backward":
                    r._sv_ctx = None
            r.admitted = now
            r._next_tile = 0
            self.active.append(r)

    def _cost(self, r):
        return self.bwd_weight if getattr(r, "mode", "forward") == "forward_backward" else 1

    def _schedule(self):
        """Pick this step's tiles: up to quantum per active request in arrival
        order, leftover capacity to the oldest requests.  Capacity and quantum
        count forward tiles; a forward+backward tile counts bwd_weight of them
        (it costs that much more), so a step's duration stays bounded whatever
        the mix and a short request never waits behind a long step.  With the
        grouped backward a step also holds at most bwd_capacity
        forward+backward tiles, and forward+backward requests whose cells
        together fit max_buckets."""
        plan = []
        cap = self.capacity
        grouped = self.backward == "grouped" and self._records is not None
        bwd_left = self.bwd_capacity if grouped else 1 << 62
        buckets = 0

        def bwd(r):
            return getattr(r, "mode", "forward") == "forward_backward"

        def room(r, w):
            lim = cap // w
            return min(lim, bwd_left) if bwd(r) else lim

        for k, r in enumerate(self.active):
            if cap <= 0:
                break
            w = self._cost(r)
            if grouped and bwd(r):
                n = self.scenes[r.scene].num_cells
                if bwd_left <= 0 or buckets + n > self.max_buckets:
                    continue
            quantum = self.quantum
            if k == 0:  # the oldest request may take a larger share
                quantum = max(quantum, int(self.oldest_share * self.capacity))
            take = min(r.ntiles - r._next_tile, max(quantum // w, 1), room(r, w))
            if take > 0:
                plan.append([r, r._next_tile, take])
                cap -= take * w
                if bwd(r):
                    bwd_left -= take
                    if grouped:
                        buckets += self.scenes[r.scene].num_cells
        for p in plan:
            if cap <= 0:
                break
            r, b, take = p
            w = self._cost(r)
            more = min(r.ntiles - (b + take), room(r, w))
            if more > 0:
                p[2] += more
                cap -= more * w
                if bwd(r):
                    bwd_left -= more
        return plan

    def _write_table(self, buf, plan):
        """Forward entries first, then forward+backward ones (the step's
        two launches split the tile space at the boundary)."""
        plan = sorted(plan, key=lambda p: p[0].mode == "forward_backward")
        n = len(plan)
        arr = (_lib.PfServeRequest * max(n, 1))()
        t0 = 0
        fwd_tiles = 0
        buckets = 0  # grouped backward: each forward+backward entry's cells from bucket0
        for k, (r, b, take) in enumerate(plan):
            e = arr[k]
            w = pack_cameras([r.camera])[0]
            ctypes.memmove(ctypes.addressof(e.cam), w.ctypes.data, 48)
            e.scene = self.scene_index[r.scene]
            e.mode = 1 if r.mode == "forward_backward" else 0
            e.tile_begin = b
            e.tile_count = take
            e.step_tile0 = t0
            e.start_cell = r._cam_dev.data_ptr() + 36
            e.color = r.color.data_ptr()
            e.log_t = r.log_t.data_ptr()
            e.texel_rgb = r._texel_rgb.data_ptr() if r._texel_rgb is not None else 0
            if r.mode == "forward_backward":
                e.target = r.target.data_ptr()
                for key in ("spheres", "nsigmas", "texel_sites", "texel_rgb", "texel_height"):
                    setattr(e.grads, key, r.grads[key].data_ptr())
                e.bucket0 = buckets
                buckets += self.scenes[r.scene].num_cells
            t0 += take
            if r.mode != "forward_backward":
                fwd_tiles = t0
        host = self._host_tab[buf]
        ctypes.memmove(host.data_ptr(), ctypes.addressof(arr), ctypes.sizeof(arr))
        desc = np.array([n, t0, fwd_tiles, buckets], dtype=np.int64)
        ctypes.memmove(host.data_ptr() + self.max_reqs * self._req_bytes, desc.ctypes.data, 32)
        return t0

    def _launch(self, buf):
        lib = _lib.load()
        host, dev = self._host_tab[buf], self._dev_tab[buf]
        dev.copy_(host, non_blocking=True)
        rec = ctypes.byref(self._records[buf][1]) if self._records is not None else None
        desc = dev.data_ptr() + self.max_reqs * self._req_bytes
        if self.backward == "grouped" and self._records is not None:
            work = self._records[buf][0][3]
            rc = lib.pf_serve_step_grouped(_lib.ptr(self._scene_tab), dev.data_ptr(), desc, self.capacity,
                                           self.bwd_capacity, self.thr, self.max_steps, _lib.ptr(self.status), rec,
                                           self.max_buckets, 0, work.data_ptr(), work.numel(), _lib.stream_ptr())
            _lib.check(rc, "pf_serve_step_grouped")
            return
        rc = lib.pf_serve_step(_lib.ptr(self._scene_tab), dev.data_ptr(), desc, self.capacity, self.thr,
                               self.max_steps, self._flags, _lib.ptr(self.status), rec, _lib.stream_ptr())
        _lib.check(rc, "pf_serve_step")

    def _build_graphs(self):
        # steps in flight may still read their tables, which the warm-up
        # below rewrites, and the old graphs: let them finish first
        torch.cuda.synchronize()
        self._graphs = []
        for buf in range(self.nbuf):
            s = self._streams[buf]
            s.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(s):  # warm up outside capture with an empty table
                desc = np.zeros(4, np.int64)
                ctypes.memmove(self._host_tab[buf].data_ptr() + self.max_reqs * self._req_bytes, desc.ctypes.data,
                               32)
                self._launch(buf)
        torch.cuda.synchronize()
        for buf in range(self.nbuf):
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=self._streams[buf]):
                self._launch(buf)
            self._graphs.append(g)
        torch.cuda.synchronize()

    def step(self, now=None):
        """Run one step if there is work; returns the number of tiles."""
        now = time.perf_counter() if now is None else now
        self._admit(now)
        if not self.active and self.queue and self.inflight:
            # the queue waits for memory that finished steps still hold
            self.poll(block=True)
            now = time.perf_counter()
            self._admit(now)
        if not self.active:
            return 0
        plan = self._schedule()
        buf = self._step_count % self.nbuf
        if self._events[buf] is not None:
            self._events[buf].synchronize()  # that buffer's previous step has read its table
        if self.use_graph and self._graphs is None:
            self._build_graphs()  # its warm-up writes empty tables, so it goes before this step's table
        ntiles = self._write_table(buf, plan)
        s = self._streams[buf]
        s.wait_stream(torch.cuda.current_stream())  # admission work (start cells, colour pre-pass)
        with torch.cuda.stream(s):
            if self.use_graph:
                self._graphs[buf].replay()
            else:
                self._launch(buf)
            ev = torch.cuda.Event()
            ev.record(s)
        self._events[buf] = ev
        completing = []
        for r, b, take in plan:
            r._next_tile = b + take
            r._events.append(ev)
            if r._next_tile >= r.ntiles:
                completing.append(r)
        self.active = [r for r in self.active if r._next_tile < r.ntiles]
        self.inflight.append((ev, completing))
        self._step_count += 1
        return ntiles

    def poll(self, block=False):
        while self.inflight:
            ev, reqs = self.inflight[0]
            if not block and not ev.query():
                return
            ev.synchronize()
            for r in reqs:  # its earlier steps may have run on the other stream
                for e in r._events:
                    e.synchronize()
            t = time.perf_counter()
            self.inflight.popleft()
            for r in reqs:
                r._events = []
                r.done = t
                r._texel_rgb = None  # its last step has completed
                self.reserved -= r._bytes
                self.finished.append(r)
                if self.on_finish is not None:
                    self.on_finish(r)

    def finish_gradients(self, r):
        """Colour-parameter gradients of a finished forward+backward request:
        dL/dSH (SH backward of its texel-colour gradient), or for a lobe
        scene (d axes, d values, d temperatures) (SphericalVoronoi's
        backward, color_fn.py:10-24); None for fixed texel_rgb."""
        sc = self.scenes[r.scene]
        if r.grads is None:
            return None
        if sc.sv is not None:
            ctx, r._sv_ctx = r._sv_ctx, None
            return sv_prepass_backward(ctx, r.grads["texel_rgb"].reshape(-1, 3))
        if sc.sh is None:
            return None
        return sh_backward(sc.sh, 3, sc.walk.texel_sites, r._cam_dev, r.grads["texel_rgb"][None])

    def run(self, requests, max_inflight=2):
        """Serve a trace of requests (arrival times in seconds from now)."""
        with steady_host():
            return self._run(requests, max_inflight)

    def _run(self, requests, max_inflight):
        t0 = time.perf_counter()
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        for r in pending:
            r.arrival += t0
        while pending or self.queue or self.active or self.inflight:
            now = time.perf_counter()
            while pending and pending[0].arrival <= now:
                self.submit(pending.popleft())
            if (self.active or self.queue) and len(self.inflight) < max_inflight:
                self.step(now)
            else:
                self.poll(block=bool(self.inflight) and not pending and not self.queue and not self.active)
                if not self.inflight and not self.active and not self.queue and pending:
                    time.sleep(max(0.0, min(pending[0].arrival - time.perf_counter(), 0.001)))
            self.poll()
        torch.cuda.synchronize()
        return time.perf_counter() - t0


@contextlib.contextmanager
def steady_host():
    """Python's cyclic garbage collector off while a serving loop runs (one
    full collection first).  A collection pauses the host for tens of
    milliseconds while it walks every tracked object, and the requests
    completing meanwhile are seen late: on the 8-scene trace at load 0.6
    such a pause put four of the last requests at 60-77 ms (p99 40.2 ms),
    against 2-17 ms and p99 36.3 ms with the collector off.  The loops'
    objects are freed by reference counting; anything cyclic waits for the
    collector, which is back on when the loop returns."""
    was = gc.isenabled()
    gc.collect()
    gc.disable()
    try:
        yield
    finally:
        if was:
            gc.enable()


class _SceneView:
    def __init__(self, c, num_cells, K, keep):
        self.c, self.num_cells, self.K, self._keep = c, num_cells, K, keep


class PlainRenderer:
    """Per-request drop-in path: colour pre-pass (SH or lobes) + walk forward
    (+ backward and colour backward), one request at a time, as a user of
    RayTracer would."""

    def __init__(self, scenes, transmittance_threshold=1e-3):
        self.scenes = scenes
        self.thr = transmittance_threshold
        self.on_finish = None  # as RenderServer.on_finish

    def render(self, r: RenderRequest):
        sc = self.scenes[r.scene]
        dev = sc.points.device
        cam = cameras_to_device([r.camera], None, dev)
        start_cells_device(sc.points, sc.radii, cam)
        w = sc.walk
        rgb, sv_ctx = sc.colour_prepass(r.camera, cam, keep_for_backward=True)
        if rgb is not None:
            c = _lib.PfScene.from_buffer_copy(w.c)
            c.texel_rgb = rgb.data_ptr()
            scene = _SceneView(c, w.num_cells, w.K, rgb)
        else:
            scene = w
        # a differentiable render records its contributing segments, as
        # RayTracer.forward does, and the backward reads them
        cap = DEFAULT_RECORD_CAP if r.mode == "forward_backward" else 0
        out = launch_walk_forward(scene, cam, r.width, r.height, self.thr, want_depth=False, record_cap=cap)
        r.color, r.log_t = out["color"][0], out["log_t"][0]
        if r.mode == "forward_backward":
            tgt = r.target.to(dev, torch.float32)
            gc = 2.0 * (out["color"] - tgt[None])
            r.grads = launch_walk_backward(scene, cam, r.width, r.height, out["color"], out["log_t"], gc, None,
                                           self.thr, record=out["record"])
            if sc.sv is not None:
                r.colour_grads = sv_prepass_backward(sv_ctx, r.grads["texel_rgb"].reshape(-1, 3))
            elif sc.sh is not None:
                r.colour_grads = sh_backward(sc.sh, 3, w.texel_sites, cam, r.grads["texel_rgb"][None])
        return r

    def run(self, requests):
        """FIFO over a trace; each request starts when it has arrived and the
        previous one is done (with the server's steady host, steady_host)."""
        with steady_host():
            return self._run(requests)

    def _run(self, requests):
        t0 = time.perf_counter()
        for r in sorted(requests, key=lambda q: q.arrival):
            r.arrival += t0
            while time.perf_counter() < r.arrival:
                time.sleep(max(0.0, min(r.arrival - time.perf_counter(), 0.001)))
            r.admitted = time.perf_counter()
            self.render(r)
            torch.cuda.synchronize()
            r.done = time.perf_counter()
            if self.on_finish is not None:
                self.on_finish(r)
        return time.perf_counter() - t0


def latency_stats(requests):
    lat = np.array([r.done - r.arrival for r in requests if r.done is not None])
    if lat.size == 0:
        return dict(p50_ms=None, p99_ms=None, n=0)
    return dict(p50_ms=float(np.percentile(lat, 50) * 1e3), p99_ms=float(np.percentile(lat, 99) * 1e3),
                mean_ms=float(lat.mean() * 1e3), n=int(lat.size))
