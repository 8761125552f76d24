"""One MLS-MPM growth step and its reverse on the GPU
(python/morphompm/transfer.py:33-158).

    step(state, model, cfg, advect=True)                      -> (state', cache)
    step_vjp(cache, xn_b, vn_b, Cn_b, Fn_b, Fgn_b, advect=True) -> (x_b, v_b, C_b, F_b, Fg_b)

Same signatures and semantics as the reference: particles with stencil nodes
outside the grid skip those nodes, grid velocity is (1 - damping) mom / m
where m > 1e-12, F is updated with the new C, Fg passes through, and with
advect=False x_b is returned as 0.  ``cache`` is an opaque handle holding the
step's device buffers (input state, grid, sort record); like the reference's
cache it refers to the step's inputs, so it stays valid after the caller
changes its own arrays.  float64 states run the FP64 kernels, float32 the
FP32 kernels.
"""
from collections import OrderedDict

import torch

from ._convert import Like, to_dev
from .engine import BatchPlan, InstanceSpec, N_DYN
from .state import ParticleState

_PLANS = OrderedDict()
_MAX_PLANS = 16


def plan_for(n, cfg, model, tile=None, block=256):
    spec = InstanceSpec.from_cfg(n, cfg, model)
    key = (spec.n, spec.shape, spec.dx, spec.dt, spec.damping, spec.pmass, spec.vol0, spec.kind,
           spec.prm, tile, block)
    p = _PLANS.get(key)
    if p is None:
        from .engine import DEFAULT_TILE
        p = BatchPlan([spec], tile=tile or DEFAULT_TILE, block=block)
        _PLANS[key] = p
        while len(_PLANS) > _MAX_PLANS:
            _PLANS.popitem(last=False)
    else:
        _PLANS.move_to_end(key)
    return p


class StepCache:
    __slots__ = ("plan", "dtype", "dyn_in", "G", "dyn_out", "grid", "sort", "payload", "like",
                 "advect", "cfg", "model")


def _pack_state(plan, state, dtype):
    dyn = plan.pack_dyn([(state.x, state.v, state.C, state.F)], dtype)
    G = plan.pack_mat([state.Fg], dtype)
    return dyn, G


def step(state, model, cfg, advect=True):
    like = Like(state.x)
    dt = like.dtype
    plan = plan_for(state.n, cfg, model)
    dyn_in, G = _pack_state(plan, state, dt)
    dyn_out = plan.new_dyn(dt)
    grid = plan.new_grid(dt)
    sort = plan.new_sort()
    payload = plan.new_payload(dt)
    plan.step_fwd(dyn_in, G, None, dyn_out, grid, sort, payload, advect)
    x, v, C, F = plan.unpack_dyn(dyn_out)
    new = ParticleState(like.out(x), like.out(v), like.out(C), like.out(F),
                        like.out(plan.unpack_mat(G)))
    c = StepCache()
    c.plan, c.dtype, c.dyn_in, c.G, c.dyn_out, c.grid, c.sort = plan, dt, dyn_in, G, dyn_out, grid, sort
    c.payload = payload
    c.like, c.advect, c.cfg, c.model = like, bool(advect), cfg, model
    return new, c


def step_vjp(cache, xn_b, vn_b, Cn_b, Fn_b, Fgn_b, advect=True):
    """Reverse of ``step``; ``advect`` must match the forward call
    (python/morphompm/transfer.py:83-86)."""
    p, dt = cache.plan, cache.dtype
    bar_out = p.pack_dyn([(xn_b, vn_b, Cn_b, Fn_b)], dt, positions=False)
    G_bar = p.pack_mat([Fgn_b], dt)
    bar_in = p.new_dyn(dt)
    p.step_bwd(cache.dyn_in, cache.G, None, cache.dyn_out, cache.grid, cache.sort, cache.payload,
               bar_out, bar_in, G_bar, None, advect)
    x, v, C, F = p.unpack_dyn(bar_in, positions=False)
    o = cache.like.out
    return o(x), o(v), o(C), o(F), o(p.unpack_mat(G_bar))
