"""torch.autograd front of the displacement map and the filtered variance.

The reference has no gradients; these wrap the library's forward map
w -> psi (the Zel'dovich path of ICGenerator::done, genetIC/src/ic.hpp:
1633-1648 and zeldovich.hpp:152-253, with the level combination of
field.hpp:816-849) and its hand-written transpose, and the filtered variance
of QuadraticModification (quadraticmodification.hpp:50-92) with its gradient
2 Q w.  Tensors live on the context's device in its precision (float32 for a
32-bit context), one dense (n, n, n) tensor per level, z fastest.

The library runs on its own HIP stream and returns synchronised; torch's
stream is synchronised before each call so the inputs are complete.
"""
import torch

from .gic import FLOAT32


def _dtype(gen):
    return torch.float32 if gen.precision == FLOAT32 else torch.float64


def _check_levels(gen, tensors, what):
    if len(tensors) != len(gen.sizes):
        raise ValueError("%s: one tensor per level expected (%d levels)" % (what, len(gen.sizes)))
    for t, n in zip(tensors, gen.sizes):
        if t.shape != (n, n, n):
            raise ValueError("%s: level tensor of shape %s, expected %s" % (what, tuple(t.shape), (n, n, n)))
        if t.dtype != _dtype(gen) or t.device.type != "cuda":
            raise ValueError("%s: tensors must be %s on the GPU" % (what, _dtype(gen)))


class Displacement(torch.autograd.Function):
    """psi = Psi(w): white noise per level -> 3 displacement components per
    level (Mpc/h).  backward applies the transposed chain."""

    @staticmethod
    def forward(ctx, gen, *noise):
        noise = [t.contiguous() for t in noise]
        _check_levels(gen, noise, "displacement")
        psi = [torch.empty_like(noise[l]) for l in range(len(noise)) for _ in range(3)]
        torch.cuda.current_stream().synchronize()
        gen.displacement_forward([t.data_ptr() for t in noise], [p.data_ptr() for p in psi])
        ctx.gen = gen
        ctx.shapes = [t.shape for t in noise]
        return tuple(psi)

    @staticmethod
    def backward(ctx, *grad_psi):
        gen = ctx.gen
        dt = _dtype(gen)
        gp = []
        for i, g in enumerate(grad_psi):
            shape = ctx.shapes[i // 3]
            gp.append(torch.zeros(shape, dtype=dt, device="cuda") if g is None else g.to(dt).contiguous())
        out = [torch.empty(s, dtype=dt, device="cuda") for s in ctx.shapes]
        torch.cuda.current_stream().synchronize()
        gen.displacement_adjoint([g.data_ptr() for g in gp], [o.data_ptr() for o in out])
        return (None, *out)


def displacement(gen, *noise):
    """Differentiable Zel'dovich displacements: a list of (psi_x, psi_y,
    psi_z) per level."""
    psi = Displacement.apply(gen, *noise)
    return [tuple(psi[3 * l:3 * l + 3]) for l in range(len(noise))]


class FilteredVariance(torch.autograd.Function):
    """V(w): the filtered variance of the flagged cells (calculate variance at
    the context's filtering_scale) as a float64 scalar; d V / d w = 2 Q w."""

    @staticmethod
    def forward(ctx, gen, *noise):
        noise = [t.contiguous() for t in noise]
        _check_levels(gen, noise, "variance")
        grads = [torch.empty_like(t) for t in noise]
        torch.cuda.current_stream().synchronize()
        v = gen.variance_gradient([t.data_ptr() for t in noise], [g.data_ptr() for g in grads])
        ctx.save_for_backward(*grads)
        return torch.tensor(v, dtype=torch.float64, device=noise[0].device)

    @staticmethod
    def backward(ctx, grad_v):
        grads = ctx.saved_tensors
        return (None, *[g * grad_v.to(g.dtype) for g in grads])


def filtered_variance(gen, *noise):
    return FilteredVariance.apply(gen, *noise)


def sphere_loss(gen, noise, mask, v_target):
    """The differentiation workload's scalar loss on one level:
    sum over the masked cells of |psi|^2 plus (V - v_target)^2, in float64."""
    (px, py, pz), = displacement(gen, noise)
    m = mask.to(px.dtype)
    l1 = ((px * m) ** 2 + (py * m) ** 2 + (pz * m) ** 2).double().sum()
    v = filtered_variance(gen, noise)
    return l1 + (v - v_target) ** 2
