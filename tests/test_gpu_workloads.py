"""GPU checks at the benchmark workloads' own sizes.

- 1024^3 + 2 x 256^3 zoom with 2LPT and four constraints (the headline
  workload; too large for the oracle): the production draw equals the
  serial stream cell by cell, the constraints are reached, FP32 psi agrees
  with FP64 psi to the 1e-5-cell target, 2LPT matches the two-wave closed
  form at 1024^3.
- 128^3 + 128^3 x4 zoom with an overdensity of 2.5 in a 1 Mpc/h sphere:
  FP32 and FP64 runs against the FP64 oracle, fields and particles.
- 64 realisations of 128^3 with 16 constraints each in one batch: two
  realisations against the oracle, every constraint of all 64 reached.
"""
import os

import numpy as np
import pytest

import pyoracle as po
from genetic_amd import gic
from helpers import select_sphere_cells

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EH = os.path.join(ROOT, "genetic_amd", "data", "eh_nowiggle_512.dat")
COSMO = dict(Om=0.31, Ol=0.69, s8=0.81, ns=0.967, z=99.0, hubble=0.677)
SEED = 8896131

# the headline workload: base_grid 200 1024; zoom_grid_with_origin_at
# 448 448 448 8 256; zoom_grid_with_origin_at 64 64 64 2 256; constraints in
# a 2 Mpc/h sphere at the box centre on the innermost level
W5_BOX, W5_N, W5_ZOOMS = 200.0, 1024, [(448, 448, 448, 8, 256), (64, 64, 64, 2, 256)]
W5_MODS = [("overdensity", 2.5), ("Lx", 0.0), ("Ly", 0.0), ("Lz", 0.0)]


def _gpu(precision, box, n, zooms=(), lpt=1):
    ic = gic.ICGenerator(precision=precision)
    ic.set_cosmology(**COSMO)
    ic.set_camb(EH)
    ic.init_base_grid(box, n)
    for z in zooms:
        ic.init_zoom_grid_with_origin_at(*z)
    if lpt == 2:
        ic.set_lpt_order(2)
    return ic


def _oracle(box, n, zooms=()):
    o = po.OracleIC()
    o.set_cosmology(**COSMO)
    tab, nc = po.read_camb(EH)
    o.set_camb(tab, nc)
    o.init_base_grid(box, n)
    for z in zooms:
        o.init_zoom_grid_at(*z)
    return o


def _sphere(ic, level, centre, radius):
    n, _, off, dx = ic.level_geometry(level)
    box0 = ic.level_geometry(0)[1]
    return select_sphere_cells(n, box0, centre, radius, offset=tuple(off), dx=dx, periodic=box0)


# ------------------------------------------------------------ 1024^3 zoom ---
def test_workload_draw_is_the_serial_stream():
    """The FP32 draw of all three levels (one stream continued across them)
    equals the FP64 Gaussian stream cell by cell, less each level's mean
    (the k = 0 mode the draw zeroes), to FP32 transform rounding.  The
    stream itself is checked against the serial generator in
    tests/test_gpu_rng_chunks.py."""
    sizes = [W5_N] + [z[4] for z in W5_ZOOMS]
    total = sum(s ** 3 for s in sizes)
    print("workload draw vs stream: %d Gaussians" % total, flush=True)
    ref = gic.ICGenerator(precision=gic.FLOAT64)
    stream = ref.gaussian_stream(SEED, total)
    ref.close()
    ic = _gpu(gic.FLOAT32, W5_BOX, W5_N, W5_ZOOMS)
    ic.set_seed(SEED)
    ic.draw()
    start = 0
    for level, n in enumerate(sizes):
        seg = stream[start:start + n ** 3]
        mean = float(np.mean(seg))
        got = ic.field(gic.FIELD_OVERDENSITY, level, dtype=np.float32).reshape(-1)
        worst = 0.0
        step = 1 << 24
        for i in range(0, n ** 3, step):
            d = np.abs(got[i:i + step].astype(np.float64) - (seg[i:i + step] - mean))
            worst = max(worst, float(d.max()))
        print("level %d: max |w - (g - mean)| = %.3g" % (level, worst), flush=True)
        assert worst < 5e-5, (level, worst)
        start += n ** 3
        del got
    ic.close()


@pytest.mark.parametrize("precision", [gic.FLOAT32, gic.FLOAT64])
def test_workload_constraints_reached(precision):
    """modify overdensity absolute 2.5 and Lx, Ly, Lz absolute 0 on the
    innermost level: after apply_modifications every calculated value is
    its target (FP64 to 1e-6 of the quantity's scale before the
    modification, FP32 to 2e-4), the chi^2 rises, and the 2LPT displacements
    of every level are finite."""
    print("workload constraints precision=%d" % precision, flush=True)
    ic = _gpu(precision, W5_BOX, W5_N, W5_ZOOMS, lpt=2 if precision == gic.FLOAT32 else 1)
    ic.set_seed(SEED)
    ic.set_flags(2, _sphere(ic, 2, (100.0, 100.0, 100.0), 2.0))
    before = {name: ic.calculate(name) for name, _ in W5_MODS}
    for name, target in W5_MODS:
        ic.modify(name, False, target)
    pre, post = ic.apply_modifications()
    assert post > pre
    tol = 1e-6 if precision == gic.FLOAT64 else 2e-4
    for name, target in W5_MODS:
        got = ic.calculate(name)
        scale = max(abs(before[name]), abs(target))
        print("%s: before %.6g after %.9g target %g" % (name, before[name], got, target), flush=True)
        assert abs(got - target) <= tol * scale, name
    if precision == gic.FLOAT32:
        ic.apply_power_spectrum()
        ic.zeldovich()
        for level in range(3):
            for f in (gic.FIELD_PSI_X, gic.FIELD_PSI2_Z):
                a = ic.field(f, level, dtype=np.float32)
                assert np.isfinite(a).all() and np.abs(a).max() > 0, (level, f)
                del a
    ic.close()


def test_workload_fp32_psi_matches_fp64():
    """rms over every cell of the three levels of |psi_FP32 - psi_FP64| / dx
    below 1e-5 cells (the target), from the same white noise."""
    import torch
    print("workload psi FP32 vs FP64", flush=True)
    sizes = [W5_N] + [z[4] for z in W5_ZOOMS]
    gen = torch.Generator(device="cuda").manual_seed(11)
    noise = [torch.randn((n, n, n), generator=gen, device="cuda", dtype=torch.float64) for n in sizes]
    psi = {}
    ctx = {}
    for prec in (64, 32):
        ctx[prec] = _gpu(prec, W5_BOX, W5_N, W5_ZOOMS)
        dt = torch.float64 if prec == 64 else torch.float32
        w = [t.to(dt) for t in noise]
        p = [torch.empty_like(w[l]) for l in range(len(w)) for _ in range(3)]
        torch.cuda.synchronize()
        ctx[prec].displacement_forward([t.data_ptr() for t in w], [q.data_ptr() for q in p])
        torch.cuda.synchronize()
        psi[prec] = p
        del w
    del noise
    sq, cells, worst = 0.0, 0, 0.0
    for level, n in enumerate(sizes):
        dx = ctx[64].level_geometry(level)[3]
        d2 = sum((psi[32][3 * level + a].double() - psi[64][3 * level + a]) ** 2 for a in range(3))
        sq += float(d2.sum()) / dx ** 2
        worst = max(worst, float(d2.max().sqrt()) / dx)
        cells += n ** 3
        del d2
    for c in ctx.values():
        c.close()
    rms = (sq / cells) ** 0.5
    print("psi FP32 vs FP64: rms %.3g cells, max %.3g cells" % (rms, worst), flush=True)
    assert rms < 1e-5
    assert worst < 1e-3


def test_2lpt_two_waves_closed_form_1024():
    """2LPT at 1024^3 FP32 (the workload's base size): delta = A cos(k1 x) +
    B cos(k2 y) gives psi2 = -(3/7) A B (k1 sin k1x cos k2y, k2 cos k1x
    sin k2y, 0) / (k1^2 + k2^2)."""
    n, box, A, B = 1024, W5_BOX, 0.3, 0.2
    print("2lpt closed form n=%d" % n, flush=True)
    ic = _gpu(gic.FLOAT32, box, n, lpt=2)
    ic.set_seed(1)
    ic.apply_power_spectrum()
    x = np.arange(n) * box / n
    k1, k2 = 2 * np.pi * 5 / box, 2 * np.pi * 7 / box
    cx, sx = np.cos(k1 * x), np.sin(k1 * x)
    cy, sy = np.cos(k2 * x), np.sin(k2 * x)
    delta = np.empty((n, n, n), dtype=np.float32)
    delta[:] = (A * cx[:, None] + B * cy[None, :])[:, :, None]
    ic.set_field(0, delta)
    del delta
    ic.zeldovich()
    ex = -(3 / 7) * A * B * k1 * np.outer(sx, cy) / (k1 ** 2 + k2 ** 2)
    ey = -(3 / 7) * A * B * k2 * np.outer(cx, sy) / (k1 ** 2 + k2 ** 2)
    s = np.abs(ex).max()
    for f, e in ((gic.FIELD_PSI2_X, ex), (gic.FIELD_PSI2_Y, ey), (gic.FIELD_PSI2_Z, None)):
        got = ic.field(f, 0, dtype=np.float32)
        if e is None:
            err = float(np.abs(got).max())
        else:
            err = max(float(np.abs(got[:, :, z] - e).max()) for z in (0, 1, 511, 1023))
            # every z plane equals plane 0
            err = max(err, float(np.abs(got - got[:, :, :1]).max()))
        print("psi2 component %d: max error %.3g of %.3g" % (f, err, s), flush=True)
        assert err < 2e-5 * s, f
        del got
    ic.close()


# ------------------------------------------------------------- 128^3 zoom ---
C2_ZOOM = (48, 48, 48, 4, 128)


@pytest.mark.parametrize("precision,tol", [(gic.FLOAT64, 1e-9), (gic.FLOAT32, 1e-4)])
def test_zoom128_constrained_matches_oracle(precision, tol):
    """base_grid 50 128; zoom_grid_with_origin_at 48 48 48 4 128; centre 25
    25 25; select_sphere 1; modify overdensity absolute 2.5; done -- against
    the FP64 oracle: chi^2 before and after, the constrained value, delta
    and psi on both levels, and the particles of both levels."""
    print("zoom128 constrained precision=%d" % precision, flush=True)
    ic = _gpu(precision, 50.0, 128, [C2_ZOOM])
    o = _oracle(50.0, 128, [C2_ZOOM])
    ic.set_seed(SEED)
    o.seed(SEED)
    cells = _sphere(ic, 1, (25.0, 25.0, 25.0), 1.0)
    assert cells.size > 4000
    ic.set_flags(1, cells)
    o.set_flags(1, cells)
    ic.modify("overdensity", False, 2.5)
    o.add_modification("overdensity", False, 2.5)
    g_pre, g_post = ic.apply_modifications()
    o_pre, o_post = o.apply_modifications()
    ctol = 1e-9 if precision == gic.FLOAT64 else 2e-5
    assert abs(g_pre - o_pre) <= ctol * o_pre and abs(g_post - o_post) <= ctol * o_post
    assert abs(ic.calculate("overdensity") - 2.5) <= (1e-9 if precision == gic.FLOAT64 else 1e-4) * 2.5
    ic.zeldovich()
    o.zeldovich()
    for level in range(2):
        psi_max = 0.0
        for which in range(4):
            a = ic.field(which, level)
            b = o.field(which, level)
            err = np.abs(a - b).max() / np.abs(b).max()
            print("level %d field %d: max rel error %.3g" % (level, which, err), flush=True)
            assert err <= tol, (level, which, err)
            if which:
                psi_max = max(psi_max, float(np.abs(b).max()))
        # particles are written as float32 in both precisions
        pos, vel = ic.particles(level)
        opos, ovel, _, _ = o.particles(level)
        box = 50.0
        dpos = np.abs(pos - opos)
        dpos = np.minimum(dpos, box - dpos)   # positions wrap into the box
        assert dpos.max() <= 2e-6 * box + 3 * tol * psi_max, level
        assert np.abs(vel - ovel).max() <= (3 * tol + 1e-6) * np.abs(ovel).max(), level
    ic.close()


# --------------------------------------------------------- 64 x 128^3 batch ---
def test_batch_64x128_against_oracle_and_targets():
    """The batching workload as benchmarked (64 realisations of 128^3, seeds
    1..64, 16 overdensity constraints each from the splitmix64 stream, FP32
    fields, one HIP graph): realisations 1 and 64 against the FP64 oracle,
    and every one of the 64 x 16 constraints reached in the batch output."""
    import sys
    import torch
    sys.path.insert(0, ROOT)
    from bench import batch_constraints
    n, box, nreal, nmods = 128, 50.0, 64, 16
    print("batch 64 x 128^3 x 16 constraints", flush=True)
    plan = batch_constraints(nreal, nmods, box)
    cons = np.array([[(c[0], c[1], c[2], r, t) for c, r, t in mods] for mods in plan])
    seeds = np.arange(1, nreal + 1, dtype=np.uint64)
    ic = _gpu(gic.FLOAT32, box, n)
    out = torch.full((nreal, 4, n, n, n), float("nan"), dtype=torch.float32, device="cuda")
    ic.batch_constrained(seeds, cons, out.data_ptr(), lanes=8, graph=True)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    del out
    ic.close()
    assert np.isfinite(got).all()
    worst = 0.0
    for r in range(nreal):
        d = got[r, 0].reshape(-1)
        for (x, y, z, rad, t) in cons[r]:
            cl = select_sphere_cells(n, box, (x, y, z), rad)
            worst = max(worst, abs(float(d[cl].astype(np.float64).mean()) - t) / max(1.0, abs(t)))
    print("batch: worst constraint error %.3g" % worst, flush=True)
    assert worst < 1e-4
    for r in (0, nreal - 1):
        o = _oracle(box, n)
        o.seed(int(seeds[r]))
        for (x, y, z, rad, t) in cons[r]:
            o.set_flags(0, select_sphere_cells(n, box, (x, y, z), rad))
            o.add_modification("overdensity", False, float(t))
        o.apply_modifications()
        o.zeldovich()
        for f in range(4):
            b = o.field(f, 0)
            err = float(np.abs(got[r, f].astype(np.float64) - b).max() / np.abs(b).max())
            print("realisation %d field %d: max rel error %.3g" % (r + 1, f, err), flush=True)
            assert err < 1e-4, (r, f, err)
        del o
