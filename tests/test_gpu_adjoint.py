"""The displacement map's transpose and the filtered-variance gradient
through the C-ABI and the torch.autograd layer: the forward matches the
product path, <Psi w, g> = <w, Psi^T g> on one, two and three levels,
central finite differences of the differentiation workload's loss, and the
exact line search for quadratic modifications."""
import os

import numpy as np
import pytest
import torch

from genetic_amd import autodiff, gic

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EH_TABLE = os.path.join(HERE, "..", "genetic_amd", "data", "eh_nowiggle_512.dat")
COSMO = dict(Om=0.31, Ol=0.69, hubble=0.677, s8=0.81, ns=0.967, z=99)

ONE = [(50.0, 64)]
TWO = [(50.0, 64), (24, 24, 24, 4, 64)]
THREE = [(50.0, 64), (16, 16, 16, 2, 64), (16, 16, 16, 2, 64)]


def _gen(levels, precision=gic.FLOAT64):
    g = gic.ICGenerator(precision=precision)
    g.set_cosmology(**COSMO)
    g.set_camb(EH_TABLE)
    g.init_base_grid(*levels[0])
    for z in levels[1:]:
        g.init_zoom_grid_with_origin_at(*z)
    return g


def _dt(precision):
    return torch.float32 if precision == gic.FLOAT32 else torch.float64


def _noise(gen, seed, precision):
    rng = np.random.default_rng(seed)
    return [torch.tensor(rng.standard_normal((n, n, n)), dtype=_dt(precision), device="cuda") for n in gen.sizes]


def _forward(gen, noise):
    psi = [torch.empty_like(noise[l]) for l in range(len(noise)) for _ in range(3)]
    torch.cuda.synchronize()
    gen.displacement_forward([t.data_ptr() for t in noise], [p.data_ptr() for p in psi])
    return psi


def _adjoint(gen, psi_bar):
    out = [torch.empty_like(psi_bar[3 * l]) for l in range(len(psi_bar) // 3)]
    torch.cuda.synchronize()
    gen.displacement_adjoint([p.data_ptr() for p in psi_bar], [o.data_ptr() for o in out])
    return out


def _sphere_mask(gen, level, centre, radius):
    n, _, off, dx = gen.level_geometry(level)
    box = gen.level_geometry(0)[1]
    x = np.arange(n) * dx + dx / 2
    d = [((x + off[a] - centre[a] + box / 2) % box) - box / 2 for a in range(3)]
    r2 = d[0][:, None, None] ** 2 + d[1][None, :, None] ** 2 + d[2][None, None, :] ** 2
    return r2 < radius * radius


@pytest.mark.parametrize("levels", [ONE, TWO], ids=["one", "two"])
def test_forward_matches_product_path(levels):
    """gic_displacement_forward is done's Zel'dovich path: same psi as
    set_field + apply_power_spectrum + zeldovich + copy_field."""
    print("forward vs product path", levels, flush=True)
    g = _gen(levels)
    w = _noise(g, 1, gic.FLOAT64)
    psi = _forward(g, w)
    p = _gen(levels)
    for l, t in enumerate(w):
        p.set_field(l, t.cpu().numpy())
    p.apply_power_spectrum()
    p.zeldovich()
    for l in range(len(levels)):
        for a in range(3):
            ref = p.field(gic.FIELD_PSI_X + a, l)
            got = psi[3 * l + a].cpu().numpy()
            np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("precision,tol", [(gic.FLOAT64, 1e-12), (gic.FLOAT32, 2e-5)], ids=["f64", "f32"])
@pytest.mark.parametrize("levels", [ONE, TWO, THREE], ids=["one", "two", "three"])
def test_adjoint_dot_product(levels, precision, tol):
    """<Psi w, g> = <w, Psi^T g> for random w and g on every level."""
    print("dot-product test", levels, precision, flush=True)
    gen = _gen(levels, precision)
    w = _noise(gen, 2, precision)
    gbar = _noise(gen, 3, precision)
    gbar = [gbar[l] * (a + 1) for l in range(len(gbar)) for a in range(3)]
    psi = _forward(gen, w)
    wbar = _adjoint(gen, gbar)
    lhs = sum(float((p.double() * q.double()).sum()) for p, q in zip(psi, gbar))
    rhs = sum(float((a.double() * b.double()).sum()) for a, b in zip(w, wbar))
    scale = sum(float(p.double().norm() * q.double().norm()) for p, q in zip(psi, gbar))
    assert abs(lhs - rhs) <= tol * scale, (lhs, rhs, scale)
    assert all(bool(torch.isfinite(t).all()) for t in wbar)


def test_adjoint_f32_matches_f64():
    """The FP32 transpose stays within float rounding of the FP64 one."""
    print("adjoint f32 vs f64", flush=True)
    g64, g32 = _gen(TWO, gic.FLOAT64), _gen(TWO, gic.FLOAT32)
    gbar = _noise(g64, 4, gic.FLOAT64)
    gbar = [gbar[l] for l in range(len(gbar)) for _ in range(3)]
    a = _adjoint(g64, gbar)
    b = _adjoint(g32, [t.float() for t in gbar])
    for x, y in zip(a, b):
        assert float((x - y.double()).norm()) < 1e-5 * float(x.norm())


def _fd_check(levels, precision, scale, radius, ndir, eps, tol):
    gen = _gen(levels, precision)
    ref = _gen(levels, gic.FLOAT64) if precision == gic.FLOAT32 else gen
    box = levels[0][0]
    centre = (box / 2, box / 2, box / 2)
    mask = torch.tensor(_sphere_mask(gen, 0, centre, radius), device="cuda")
    cells = np.flatnonzero(mask.cpu().numpy().ravel())
    for g in {id(gen): gen, id(ref): ref}.values():
        g.set_flags(0, cells)
        g.set_variance_filter_scale(scale)
    w = _noise(gen, 5, precision)[0].requires_grad_(True)
    with torch.no_grad():
        v0 = autodiff.filtered_variance(ref, w.detach().double())
    vt = 1.5 * float(v0)
    loss = autodiff.sphere_loss(gen, w, mask, vt)
    loss.backward()
    grad = w.grad.double()
    rng = np.random.default_rng(1)
    errs = []
    w64 = w.detach().double()
    for _ in range(ndir):
        d = torch.tensor(rng.standard_normal(w.shape), dtype=torch.float64, device="cuda")
        d /= d.norm()
        with torch.no_grad():
            lp = float(autodiff.sphere_loss(ref, w64 + eps * d, mask, vt))
            lm = float(autodiff.sphere_loss(ref, w64 - eps * d, mask, vt))
        fd = (lp - lm) / (2 * eps)
        an = float((grad * d).sum())
        errs.append(abs(fd - an) / max(abs(fd), 1e-30))
    print("relative FD errors: max %.3g median %.3g" % (max(errs), float(np.median(errs))), flush=True)
    assert max(errs) < tol


def test_finite_differences_one_level_f64():
    """Central differences (eps 1e-3) of |psi|^2 over a sphere plus the
    variance penalty along random unit directions, 64^3 FP64."""
    print("FD 64^3 f64", flush=True)
    _fd_check(ONE, gic.FLOAT64, 1.0, 6.0, 8, 1e-3, 1e-6)


def test_finite_differences_workload_f32():
    """The differentiation workload: 256^3 in a 50 Mpc/h box, FP32 adjoint,
    sphere of 3 Mpc/h, variance at filtering scale 0.5 Mpc/h with target
    1.5x its value, 32 random unit directions, eps 1e-3 (differences taken
    in FP64)."""
    print("FD 256^3 f32 gradient", flush=True)
    _fd_check([(50.0, 256)], gic.FLOAT32, 0.5, 3.0, 32, 1e-3, 1e-3)


def test_variance_gradient_value_matches_calculate():
    print("variance value vs calculate", flush=True)
    g = _gen(TWO)
    w = _noise(g, 6, gic.FLOAT64)
    cells = np.flatnonzero(_sphere_mask(g, 1, (25.0, 25.0, 25.0), 2.0).ravel())
    for l, t in enumerate(w):
        g.set_field(l, t.cpu().numpy())
    g.set_flags(1, cells)
    g.set_variance_filter_scale(0.5)
    want = g.calculate("variance")
    got = g.variance_gradient([t.data_ptr() for t in w])
    assert abs(got - want) <= 1e-10 * abs(want)


def test_quadratic_line_search_hits_target():
    """With the line search each quadratic step lands on its target: the
    final variance is the target to rounding, the linear constraint holds and
    chi^2 stays close to that of the reference's steps."""
    print("quadratic line search", flush=True)
    res = {}
    for ls in (False, True):
        g = _gen(ONE)
        g.set_seed(8896131)
        cells = np.flatnonzero(_sphere_mask(g, 0, (25.0, 25.0, 25.0), 6.0).ravel())
        g.set_flags(0, cells)
        g.set_variance_filter_scale(1.0)
        g.set_quadratic_line_search(ls)
        g.modify("overdensity", False, 0.1)
        g.modify("variance", True, 1.5)
        target = g.calculate("variance") * 1.5
        g.apply_modifications()
        res[ls] = (g.calculate("variance"), g.calculate("overdensity"), g.chi2()[0], target)
    v, d, chi, target = res[True]
    assert abs(v - target) < 1e-8 * target
    assert abs(d - 0.1) < 1e-9
    vr, dr, chir, _ = res[False]
    assert abs(vr - target) < 2e-3 * target   # the reference's precision 1e-3
    assert abs(chi - chir) < 0.01 * chir, (chi, chir)
