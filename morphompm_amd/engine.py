"""Batch layout and the checkpointed trajectory engine over the HIP kernels.

A ``BatchPlan`` lays out one or more independent simulations ("instances")
for single launches: every instance's particle slots and grid nodes start at
a multiple of 256 so each 256-slot chunk belongs to one instance, bins and
P2G tiles are numbered instance after instance, and the instance table plus
the chunk->instance maps live in device memory.

A ``Trajectory`` runs n_steps forward steps and their reverse.  It stores a
checkpoint of the particle state every K steps; the reverse recomputes one
segment of K steps at a time (keeping each step's state, grid and sort
record), then runs the reverse steps of that segment.  Memory is bounded by
(n_steps / K) checkpoints plus one segment.  The last segment of the forward
pass is kept, so it is not recomputed.
"""
import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import MpmBatch, MpmInstance, SLOT_ALIGN, check, ptr, stream_ptr, suffix

N_DYN = 24        # x(3) v(3) C(9) F(9)
ROW_X, ROW_V, ROW_C, ROW_F = 0, 3, 6, 15
DEFAULT_TILE = (4, 4, 8)   # best or within 0.3% of it for the batched and bioink workloads
# The payload packs a particle's cell (base + 2 per axis, 0 .. nodes + 1) in
# 10 bits per axis, so an instance grid holds at most 1022 nodes per axis.
MAX_AXIS_NODES = 1022


def check_spec(s, i=0):
    """ValueError for an instance the kernels cannot run: a grid axis outside
    1..MAX_AXIS_NODES, a negative particle count, or a non-positive or
    non-finite dx / dt / particle mass / volume."""
    if len(s.shape) != 3 or any(int(d) < 1 or int(d) > MAX_AXIS_NODES for d in s.shape):
        raise ValueError(f"instance {i}: grid {tuple(s.shape)} must have 1..{MAX_AXIS_NODES} "
                         "nodes per axis")
    if int(s.n) < 0:
        raise ValueError(f"instance {i}: particle count {s.n} is negative")
    for name in ("dx", "dt", "pmass", "vol0"):
        v = float(getattr(s, name))
        if not (np.isfinite(v) and v > 0.0):
            raise ValueError(f"instance {i}: {name} = {v} must be positive and finite")


def model_spec(model):
    """(kind, prm6) of a constitutive model object; NotImplementedError if the
    model has no device implementation (there is no CPU fallback)."""
    kind = getattr(model, "kind", None)
    if kind in ("nh", "hencky"):
        return (_lib.NEO_HOOKEAN if kind == "nh" else _lib.HENCKY,
                (float(model.mu), float(model.lam), 0.0, 0.0, 0.0, 0.0))
    if kind == "hb":
        return (_lib.HERSCHEL_BULKLEY,
                (float(model.kappa), float(model.K), float(model.n), float(model.tau_y),
                 float(model.m), float(model.eps)))
    raise NotImplementedError(f"model {type(model).__name__} has no HIP implementation")


@dataclass
class InstanceSpec:
    n: int
    shape: tuple
    dx: float
    dt: float
    damping: float
    pmass: float
    vol0: float
    kind: int
    prm: tuple
    bin_lo: tuple = None   # lowest occupied bin per axis at the start (tile alignment)
    origin: tuple = (0, 0, 0)   # positions stored relative to origin * dx (whole cells)

    @staticmethod
    def from_cfg(n, cfg, model, pts=None, center=False):
        """``pts`` (optional rest positions) aligns the P2G tiles with the lowest
        occupied bin of each axis; ``center`` also stores positions relative to
        the whole cell nearest their mean (more FP32 bits for displacements)."""
        kind, prm = model_spec(model)
        lo, origin = None, (0, 0, 0)
        if pts is not None:
            x = np.asarray(pts.detach().cpu() if isinstance(pts, torch.Tensor) else pts,
                           dtype=np.float64)
            if len(x):
                lo = tuple(int(v) for v in
                           np.maximum(np.floor(x * (1.0 / cfg.dx) - 0.5).min(0) + 2, 0))
                if center:
                    origin = tuple(int(v) for v in np.rint(x.mean(0) / cfg.dx))
        return InstanceSpec(int(n), tuple(cfg.shape), float(cfg.dx), float(cfg.dt),
                            float(cfg.damping), float(cfg.pmass), float(cfg.vol0), kind, prm, lo,
                            origin)


def _ceil(a, b):
    return -(-a // b)


def bin_keys(x, dx, shape, tile, shift):
    """Unit-level bin key (tile id * cells per tile + place in the tile) of
    positions x (n, 3), numpy or torch (int64 result of the same kind), for a
    grid of ``shape`` nodes, the P2G tile shape and the instance's tile shift.
    Positions off the grid are clamped to its border bins."""
    if isinstance(x, torch.Tensor):
        arr = lambda v: torch.tensor(v, dtype=torch.int64, device=x.device)
        b = torch.floor(x * (1.0 / dx) - 0.5).to(torch.int64) + 2
        clip = lambda v, lo, hi: torch.minimum(torch.maximum(v, lo), hi)
    else:
        arr = lambda v: np.array(v, dtype=np.int64)
        b = np.floor(x * (1.0 / dx) - 0.5).astype(np.int64) + 2
        clip = np.clip
    dims = arr(list(shape)) + 2
    b = clip(b, arr([0, 0, 0]), dims - 1)
    T, sh = arr(list(tile)), arr(list(shift))
    b = b + sh
    t = b // T
    nt = -(-(dims + sh) // T)
    tile_id = (t[:, 0] * nt[1] + t[:, 1]) * nt[2] + t[:, 2]
    l = b - t * T
    local = (l[:, 0] * T[1] + l[:, 1]) * T[2] + l[:, 2]
    return tile_id * int(np.prod(tile)) + local


class BatchPlan:
    def __init__(self, specs, tile=DEFAULT_TILE, block=256, device="cuda"):
        self.specs = list(specs)
        if not self.specs:
            raise ValueError("a batch needs at least one instance")
        for i, s in enumerate(self.specs):
            check_spec(s, i)
        _lib.require_gpu()
        self.tile = tuple(int(t) for t in tile)
        self.block = int(block)
        self.device = torch.device(device)
        if self.tile not in _lib.tile_shapes():
            raise ValueError(f"tile {self.tile} not built; available {_lib.tile_shapes()}")
        insts = (MpmInstance * len(self.specs))()
        slot_inst, node_inst, tile_inst = [], [], []
        p_off = g_off = t_off = 0
        self.p_off, self.g_off = [], []
        for i, s in enumerate(self.specs):
            nx, ny, nz = s.shape
            sh = [0, 0, 0] if s.bin_lo is None else [(-int(b)) % t for b, t in zip(s.bin_lo, self.tile)]
            tx, ty, tz = (_ceil(nx + 2 + sh[0], self.tile[0]), _ceil(ny + 2 + sh[1], self.tile[1]),
                          _ceil(nz + 2 + sh[2], self.tile[2]))
            I = insts[i]
            I.shift[0], I.shift[1], I.shift[2] = sh
            I.origin[0], I.origin[1], I.origin[2] = (int(o) for o in s.origin)
            I.p_off, I.g_off, I.t_off = p_off, g_off, t_off
            I.n, I.nx, I.ny, I.nz, I.tx, I.ty, I.tz = s.n, nx, ny, nz, tx, ty, tz
            I.model = s.kind
            I.dx, I.dt, I.damping, I.pmass, I.vol0 = s.dx, s.dt, s.damping, s.pmass, s.vol0
            for k in range(6):
                I.prm[k] = s.prm[k]
            self.p_off.append(p_off)
            self.g_off.append(g_off)
            ps = max(1, _ceil(s.n, SLOT_ALIGN)) * SLOT_ALIGN
            gs = _ceil(nx * ny * nz, SLOT_ALIGN) * SLOT_ALIGN
            slot_inst += [i] * (ps // SLOT_ALIGN)
            node_inst += [i] * (gs // SLOT_ALIGN)
            tile_inst += [i] * (tx * ty * tz + 1)   # tiles, then the dead unit
            p_off += ps
            g_off += gs
            t_off += tx * ty * tz + 1
        self.n_slots, self.n_nodes, self.n_tiles = p_off, g_off, t_off
        self.n_bins = t_off * int(np.prod(self.tile))
        self.n_inst = len(self.specs)
        self._insts = insts
        raw = bytes(insts)
        self.d_inst = torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(self.device)
        self.d_slot_inst = torch.tensor(slot_inst, dtype=torch.int32, device=self.device)
        self.d_node_inst = torch.tensor(node_inst, dtype=torch.int32, device=self.device)
        self.d_tile_inst = torch.tensor(tile_inst, dtype=torch.int32, device=self.device)
        c = MpmBatch()
        c.n_inst = self.n_inst
        c.tile[0], c.tile[1], c.tile[2] = self.tile
        c.n_slots, c.n_nodes, c.n_bins, c.n_tiles = (self.n_slots, self.n_nodes, self.n_bins,
                                                     self.n_tiles)
        c.block = self.block
        c.model_mask = int(np.bitwise_or.reduce([1 << s.kind for s in self.specs]))
        c.inst = self.d_inst.data_ptr()
        c.slot_inst = self.d_slot_inst.data_ptr()
        c.node_inst = self.d_node_inst.data_ptr()
        c.tile_inst = self.d_tile_inst.data_ptr()
        # chunks grouped by model: model-specialised kernels, one launch per model
        kinds = [int(s.kind) for s in self.specs]
        self.d_model_chunk = None
        if len(set(kinds)) > 1:
            chunk_kind = np.array([kinds[i] for i in slot_inst], dtype=np.int64)
            order = np.argsort(chunk_kind, kind="stable").astype(np.int32)
            counts = np.bincount(chunk_kind, minlength=3)
            self.d_model_chunk = torch.from_numpy(order).to(self.device)
            c.model_chunk = self.d_model_chunk.data_ptr()
            off = np.concatenate([[0], np.cumsum(counts)])
            for m in range(4):
                c.model_chunk_off[m] = int(off[m])
        self.c = c
        self.sortlen = 2 * self.n_slots + self.n_bins + 1 + 2 * self.n_tiles + 1
        # each instance's chunks are contiguous: [start, end) in chunk units
        cnt = np.bincount(np.array(slot_inst, dtype=np.int64), minlength=self.n_inst)
        self._chunk_end = torch.from_numpy(np.cumsum(cnt)).to(self.device)
        self._chunk_start = torch.from_numpy(np.cumsum(cnt) - cnt).to(self.device)
        self._ws = {}
        # slot -> live mask and per-chunk instance (host copies for packing)
        self.slot_inst_np = np.repeat(np.array(slot_inst, dtype=np.int64), SLOT_ALIGN)
        live = np.zeros(self.n_slots, dtype=bool)
        for i, s in enumerate(self.specs):
            live[self.p_off[i]:self.p_off[i] + s.n] = True
        self.live = torch.from_numpy(live).to(self.device)

    def set_prm(self, i, prm6):
        """Replace instance i's model parameters (e.g. Herschel-Bulkley tau_y, K)
        in the device instance table, stream-ordered on the current stream."""
        for k in range(6):
            self._insts[i].prm[k] = float(prm6[k])
        raw = torch.frombuffer(bytearray(bytes(self._insts)), dtype=torch.uint8)
        self.d_inst.copy_(raw.to(self.device, non_blocking=False))

    # ------------------------------------------------------------ buffers
    def workspace(self, dtype):
        if dtype not in self._ws:
            nb = _lib.lib().mpm_workspace_bytes(ctypes.byref(self.c), 1 if dtype == torch.float64 else 0)
            if nb < 0:
                check(int(nb), "workspace_bytes")
            self._ws[dtype] = torch.zeros(int(nb), dtype=torch.uint8, device=self.device)
        return self._ws[dtype]

    def new_dyn(self, dtype):
        return torch.zeros(N_DYN, self.n_slots, dtype=dtype, device=self.device)

    def new_mat(self, dtype):
        return torch.zeros(9, self.n_slots, dtype=dtype, device=self.device)

    def new_grid(self, dtype):
        return torch.zeros(self.n_nodes, 4, dtype=dtype, device=self.device)

    def new_sort(self):
        return torch.zeros(self.sortlen, dtype=torch.int32, device=self.device)

    def new_payload(self, dtype):
        return torch.zeros(self.n_slots, 16, dtype=dtype, device=self.device)

    # ------------------------------------------------------------ packing
    # ``perm`` (optional, per instance): slot j of instance i holds particle
    # perm[i][j] of the caller's arrays.  ``cell_order`` gives the order that
    # sorts particles by bin key, so a kernel's neighbouring lanes touch
    # neighbouring grid nodes and the sorted payload writes are near-coalesced.
    def sl(self, i):
        return slice(self.p_off[i], self.p_off[i] + self.specs[i].n)

    def cell_order(self, pts_list):
        """Per instance, the stable order of its particles by bin key (tile,
        then place in the tile).  A CUDA tensor of positions gets a device
        tensor back, sorted on the device; anything else a numpy array."""
        perms = []
        for i, pts in enumerate(pts_list):
            s = self.specs[i]
            sh = [self._insts[i].shift[a] for a in range(3)]
            if isinstance(pts, torch.Tensor) and pts.is_cuda:
                key = bin_keys(pts.detach().to(torch.float64), s.dx, s.shape, self.tile, sh)
                perms.append(torch.argsort(key, stable=True))
            else:
                x = np.asarray(pts.detach().cpu() if isinstance(pts, torch.Tensor) else pts,
                               dtype=np.float64)
                perms.append(np.argsort(bin_keys(x, s.dx, s.shape, self.tile, sh), kind="stable"))
        return perms

    def _cols(self, a, i, n, k, dtype, perm):
        t = _dev(a, dtype, self.device).reshape(n, k)
        if perm is not None and perm[i] is not None:
            t = t[torch.as_tensor(perm[i], device=self.device)]
        return t.T

    def origin_offset(self, i, dtype=torch.float64):
        s = self.specs[i]
        return torch.tensor([o * s.dx for o in s.origin], dtype=dtype, device=self.device)

    def pack_dyn(self, parts, dtype, perm=None, positions=True):
        """parts: list over instances of (x, v, C, F) with shapes (n,3),(n,3),(n,3,3),(n,3,3).
        ``positions``: x is a position (stored relative to the instance origin);
        False for cotangents."""
        dyn = self.new_dyn(dtype)
        for i, (x, v, C, F) in enumerate(parts):
            s = self.sl(i)
            n = self.specs[i].n
            if positions and any(self.specs[i].origin):
                x = _dev(x, torch.float64, self.device).reshape(n, 3) - self.origin_offset(i)
            dyn[ROW_X:ROW_X + 3, s] = self._cols(x, i, n, 3, dtype, perm)
            dyn[ROW_V:ROW_V + 3, s] = self._cols(v, i, n, 3, dtype, perm)
            dyn[ROW_C:ROW_C + 9, s] = self._cols(C, i, n, 9, dtype, perm)
            dyn[ROW_F:ROW_F + 9, s] = self._cols(F, i, n, 9, dtype, perm)
        return dyn

    def pack_mat(self, mats, dtype, perm=None):
        out = self.new_mat(dtype)
        for i, M in enumerate(mats):
            out[:, self.sl(i)] = self._cols(M, i, self.specs[i].n, 9, dtype, perm)
        return out

    def _uncols(self, t, i, perm):
        if perm is None or perm[i] is None:
            return t
        inv = torch.empty(len(perm[i]), dtype=torch.int64, device=self.device)
        inv[torch.as_tensor(perm[i], device=self.device)] = torch.arange(len(perm[i]),
                                                                          device=self.device)
        return t[inv]

    def unpack_dyn(self, dyn, i=0, perm=None, positions=True):
        s = self.sl(i)
        n = self.specs[i].n
        u = lambda t: self._uncols(t, i, perm)
        x = u(dyn[ROW_X:ROW_X + 3, s].T)
        if positions and any(self.specs[i].origin):
            x = (x.to(torch.float64) + self.origin_offset(i)).to(dyn.dtype)
        return (x, u(dyn[ROW_V:ROW_V + 3, s].T),
                u(dyn[ROW_C:ROW_C + 9, s].T).reshape(n, 3, 3),
                u(dyn[ROW_F:ROW_F + 9, s].T).reshape(n, 3, 3))

    def unpack_mat(self, M, i=0, perm=None):
        n = self.specs[i].n
        return self._uncols(M[:, self.sl(i)].T, i, perm).reshape(n, 3, 3)

    def per_instance_sum(self, v):
        """Sum of a per-slot float64 vector over each instance's live slots, in a
        fixed order (cumulative sum over slots), -> (n_inst,)."""
        v = torch.where(self.live, v.to(torch.float64), torch.zeros((), dtype=torch.float64,
                                                                      device=self.device))
        cs = torch.cumsum(torch.cat([v.new_zeros(1), v]), 0)
        starts = torch.tensor(self.p_off, device=self.device)
        ends = starts + torch.tensor([s.n for s in self.specs], device=self.device)
        return cs[ends] - cs[starts]

    # ------------------------------------------------------------ kernels
    def step_fwd(self, dyn_in, G, sched_row, dyn_out, grid, sortbuf, payload, advect,
                 stream=None):
        f = getattr(_lib.lib(), f"mpm_step_fwd_{suffix(dyn_in.dtype)}")
        rc = f(ctypes.byref(self.c), ptr(dyn_in), ptr(G), sched_row, ptr(dyn_out), ptr(grid),
               ptr(sortbuf), ptr(payload), int(bool(advect)), ptr(self.workspace(dyn_in.dtype)),
               stream_ptr(stream))
        check(rc, "step_fwd")

    def step_fwd_chain(self, dyn_in, G, sched_row, sched_next, dyn_out, grid, sortbuf, payload,
                       payload_next, advect, ready, out=_lib.OUT_FULL, stream=None):
        """One step of a chain: ``ready`` -- this step's payload and sort keys
        come from the previous chained step; ``payload_next`` (or None)
        receives the next step's payload; ``out`` -- the rows of dyn_out
        written (OUT_FULL, OUT_RECORD: all but v, OUT_NEXT: x and F)."""
        f = getattr(_lib.lib(), f"mpm_step_fwd_chain_{suffix(dyn_in.dtype)}")
        rc = f(ctypes.byref(self.c), ptr(dyn_in), ptr(G), sched_row, sched_next, ptr(dyn_out),
               ptr(grid), ptr(sortbuf), ptr(payload), ptr(payload_next), int(bool(advect)),
               int(bool(ready)), int(out), ptr(self.workspace(dyn_in.dtype)),
               stream_ptr(stream))
        check(rc, "step_fwd_chain")

    def step_bwd(self, dyn_in, G, sched_row, dyn_out, grid, sortbuf, payload, bar_out, bar_in,
                 G_bar, theta_acc, advect, stream=None):
        f = getattr(_lib.lib(), f"mpm_step_bwd_{suffix(dyn_in.dtype)}")
        rc = f(ctypes.byref(self.c), ptr(dyn_in), ptr(G), sched_row, ptr(dyn_out), ptr(grid),
               ptr(sortbuf), ptr(payload), ptr(bar_out), ptr(bar_in), ptr(G_bar), ptr(theta_acc),
               int(bool(advect)), ptr(self.workspace(dyn_in.dtype)), stream_ptr(stream))
        check(rc, "step_bwd")

    def step_bwd_chain(self, dyn_in, G, sched_row, dyn_out, grid, sortbuf, payload, bar_out,
                       bar_in, G_bar, theta_acc, advect, ready, dyn_prev, sortbuf_prev,
                       stream=None):
        """One reverse step of a chain: ``ready`` -- 0 for the chain's first
        call (its adjoint payload is computed), else 1 + (i % 2) for its i-th
        call (the payload comes from the previous, later, chained step, in one
        of the workspace's two payload areas); ``dyn_prev``/``sortbuf_prev``
        (or None) -- also prepare the previous step's."""
        f = getattr(_lib.lib(), f"mpm_step_bwd_chain_{suffix(dyn_in.dtype)}")
        rc = f(ctypes.byref(self.c), ptr(dyn_in), ptr(G), sched_row, ptr(dyn_out), ptr(grid),
               ptr(sortbuf), ptr(payload), ptr(bar_out), ptr(bar_in), ptr(G_bar), ptr(theta_acc),
               int(bool(advect)), int(ready), ptr(dyn_prev), ptr(sortbuf_prev),
               ptr(self.workspace(dyn_in.dtype)), stream_ptr(stream))
        check(rc, "step_bwd_chain")

    def sort_by_cell(self, dyn, sortbuf, stream=None):
        f = getattr(_lib.lib(), f"mpm_sort_by_cell_{suffix(dyn.dtype)}")
        check(f(ctypes.byref(self.c), ptr(dyn), ptr(sortbuf), ptr(self.workspace(dyn.dtype)),
                stream_ptr(stream)), "sort_by_cell")

    def theta_per_instance(self, theta_acc):
        """theta_acc [n_slots/256, 2] float64 -> (n_inst, 2), summed over each
        instance's chunks in chunk order (device work only: safe under graph
        capture)."""
        cs = torch.cumsum(torch.cat([theta_acc.new_zeros(1, 2), theta_acc]), 0)
        return cs[self._chunk_end] - cs[self._chunk_start]


def _dev(a, dtype, device):
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=dtype)
    return torch.as_tensor(np.asarray(a), dtype=dtype, device=device)


class Trajectory:
    """n_steps forward steps with checkpoints every K steps, and their reverse.

    ``sched``: None (growth tensor used as is, the reference semantics) or a
    (n_steps, n_inst) tensor a_t: step t uses Fg = I + a_t (G - I).

    ``keep_segments`` (M, default 1): the forward pass keeps the step records
    of its last M segments, and the reverse pass recomputes only the others
    (M >= the number of segments: no recompute).  Each kept segment costs one
    segment store (K steps of state, grid, sort record and payload: about
    10 GB for 2.7 M particles at K = 20); checkpoints are written every K
    steps either way.  "auto": as many as fit in 60% of the free device
    memory.
    """

    def __init__(self, plan, dtype, n_steps, advect, checkpoint_every=None, sched=None,
                 stream=None, keep_segments=1):
        self.plan = plan
        self.dtype = dtype
        self.n_steps = int(n_steps)
        self.advect = bool(advect)
        K = int(checkpoint_every) if checkpoint_every else self.n_steps
        self.K = max(1, min(K, max(1, self.n_steps)))
        self.nseg = _ceil(self.n_steps, self.K) if self.n_steps > 0 else 0
        self.stream = stream
        if sched is not None:
            sched = torch.as_tensor(sched, dtype=dtype, device=plan.device)
            if sched.ndim == 1:
                sched = sched[:, None].expand(self.n_steps, plan.n_inst)
            sched = sched.contiguous()
            if sched.shape != (self.n_steps, plan.n_inst):
                raise ValueError(f"sched shape {tuple(sched.shape)} != ({self.n_steps}, {plan.n_inst})")
        self.sched = sched
        p = plan
        self.ckpt = torch.zeros(max(1, self.nseg), N_DYN, p.n_slots, dtype=dtype, device=p.device)
        Ks = self.K
        if keep_segments == "auto":
            keep_segments = self.auto_keep_segments(plan, dtype, Ks, max(1, self.nseg))
        self.M = max(1, min(int(keep_segments), max(1, self.nseg)))
        self._stores = [(torch.zeros(Ks, N_DYN, p.n_slots, dtype=dtype, device=p.device),
                         torch.zeros(Ks, p.n_nodes, 4, dtype=dtype, device=p.device),
                         torch.zeros(Ks, p.sortlen, dtype=torch.int32, device=p.device),
                         torch.zeros(Ks, p.n_slots, 16, dtype=dtype, device=p.device))
                        for _ in range(self.M)]
        # the store that holds the last segment
        self.seg_dyn, self.seg_grid, self.seg_sort, self.seg_payload = self._stores[-1]
        self.G = p.new_mat(dtype)
        self.bar_a = p.new_dyn(dtype)
        self.bar_b = p.new_dyn(dtype)
        self.G_bar = p.new_mat(dtype)
        self.theta_acc = torch.zeros(p.n_slots // SLOT_ALIGN, 2, dtype=torch.float64,
                                     device=p.device)
        self._held = [-1] * self.M     # segment whose step records each store holds

    @staticmethod
    def store_bytes(plan, dtype, K):
        """Bytes of one segment store (K steps of state, grid, sort record and
        payload)."""
        e = torch.tensor([], dtype=dtype).element_size()
        return K * (plan.n_slots * (N_DYN + 16) * e + plan.n_nodes * 4 * e + plan.sortlen * 4)

    @staticmethod
    def auto_keep_segments(plan, dtype, K, nseg, fraction=0.6):
        """As many segment stores as fit in ``fraction`` of the device's free
        memory (at least 1, at most nseg)."""
        free, _ = torch.cuda.mem_get_info(plan.device)
        return max(1, min(nseg, int(fraction * free // Trajectory.store_bytes(plan, dtype, K))))

    # segment-store bookkeeping (graph replays restore it as the captured run left it)
    def store_state(self):
        return tuple(self._held)

    def set_store_state(self, state):
        self._held = list(state)

    def _a(self, t):
        if self.sched is None:
            return None
        row = self.sched[t]
        return ctypes.c_void_p(row.data_ptr())

    def _seg(self, s):
        t0 = s * self.K
        return t0, min(self.K, self.n_steps - t0)

    def _store(self, s):
        """Store index of segment s: the last M segments have one each; the
        others are (re)computed in store 0, which is free by then."""
        return max(0, s - (self.nseg - self.M))

    def _dst(self, s, k):
        """The state after step k of segment s: the next checkpoint after the
        segment's last step (the final state stays in its store), else the
        segment's store."""
        _, L = self._seg(s)
        if k == L - 1 and s < self.nseg - 1:
            return self.ckpt[s + 1]
        return self._stores[self._store(s)][0][k]

    def _src(self, s, k):
        return self.ckpt[s] if k == 0 else self._dst(s, k - 1)

    def _run_segment(self, s, keep, recompute=False):
        """The segment's steps as one chain: each step's G2P writes the next
        step's payload and sort keys; the first step computes its own.
        States written: checkpoints and the final state in full; a kept
        segment's step records without v (the reverse pass reads x, C, F);
        the intermediate states of a segment that is not kept with x and F
        only (the next step reads nothing else).  A segment the reverse pass
        recomputes (``recompute``) stops its last step after the step record:
        its output state is the closing checkpoint, already held (the final
        state, after the last segment, is written as usual)."""
        t0, L = self._seg(s)
        m = self._store(s)
        _, grid, sort, payload = self._stores[m]
        for k in range(L):
            last = k == L - 1
            if last:
                out = _lib.OUT_NONE if recompute and s < self.nseg - 1 else _lib.OUT_FULL
            else:
                out = _lib.OUT_RECORD if keep else _lib.OUT_NEXT
            self.plan.step_fwd_chain(self._src(s, k), self.G, self._a(t0 + k),
                                     None if last else self._a(t0 + k + 1), self._dst(s, k),
                                     grid[k], sort[k], payload[k],
                                     None if last else payload[k + 1], self.advect,
                                     k > 0, out, self.stream)
        self._held[m] = s if keep else -1

    def forward(self, dyn0, G=None):
        """Runs the trajectory from dyn0 ([24, n_slots]); returns the final state
        (a view into this trajectory's storage)."""
        if G is not None:
            self.G.copy_(G)
        self.ckpt[0].copy_(dyn0)
        if self.n_steps == 0:
            return self.ckpt[0]
        for s in range(self.nseg):
            self._run_segment(s, keep=(s >= self.nseg - self.M))
        return self.final()

    def final(self):
        if self.n_steps == 0:
            return self.ckpt[0]
        s = self.nseg - 1
        _, L = self._seg(s)
        if self._held[self._store(s)] != s:
            self._run_segment(s, keep=True)
        return self._dst(s, L - 1)

    def backward(self, bar_final, G_bar_final=None, want_theta=False):
        """Cotangents of (dyn0, G) from the cotangent of the final state.
        Returns (bar0 [24, n_slots], G_bar [9, n_slots], theta [n_inst, 2] or None).
        Returned tensors are this trajectory's buffers (copy to keep)."""
        self.bar_a.copy_(bar_final)
        if G_bar_final is None:
            self.G_bar.zero_()
        elif self.sched is None:
            self.G_bar.copy_(G_bar_final)
        else:
            # the final state's Fg is the last step's I + a (G - I)
            a = self.sched[self.n_steps - 1]
            chunk = torch.from_numpy(self.plan.slot_inst_np).to(self.plan.device)
            self.G_bar.copy_(G_bar_final * a[chunk][None, :])
        self.theta_acc.zero_()
        bar_out, bar_in = self.bar_a, self.bar_b
        th = self.theta_acc if want_theta else None
        for s in range(self.nseg - 1, -1, -1):
            m = self._store(s)
            if self._held[m] != s:
                self._run_segment(s, keep=True, recompute=True)
            t0, L = self._seg(s)
            _, grid, sort, payload = self._stores[m]
            # one chain per segment: each reverse step's VJP kernel prepares
            # the adjoint payload of the step before it
            for k in range(L - 1, -1, -1):
                self.plan.step_bwd_chain(self._src(s, k), self.G, self._a(t0 + k),
                                         self._dst(s, k), grid[k], sort[k], payload[k],
                                         bar_out, bar_in, self.G_bar, th, self.advect,
                                         0 if k == L - 1 else 1 + (L - 1 - k) % 2,
                                         self._src(s, k - 1) if k > 0 else None,
                                         sort[k - 1] if k > 0 else None, self.stream)
                bar_out, bar_in = bar_in, bar_out
        theta = self.plan.theta_per_instance(self.theta_acc) if want_theta else None
        return bar_out, self.G_bar, theta
