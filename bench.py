"""Benchmark of genetIC's field-generation hot path on one MI355X.

A step is one realisation: seed -> white noise (bit-exact ranlxs2 +
ziggurat) -> Hoffman-Ribak modifications of the noise -> delta = sqrt(P) * w
-> Zel'dovich and second-order psi on every level, with the zoom levels
combined, all resident in HBM (no host copies in the timed region).
Default workload: a 1024^3 base grid in a 200 Mpc/h box with 256^3 zooms at
refinement 2 and 4 nested around the box centre, FP32 fields, 2LPT, and an
overdensity plus three angular-momentum constraints on the innermost level.

Prints one JSON line (see the keys at the end of main()).
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Mparticles/s (δ+ψ) at 256³ & 1024³; FP32-vs-FP64 ψ err; adjoint/fwd ratio"
COSMO = dict(Om=0.31, Ol=0.69, s8=0.81, ns=0.967, z=99.0, hubble=0.677)
EH_TABLE = os.path.join(ROOT, "genetic_amd", "data", "eh_nowiggle_512.dat")

# innermost-level constraints of the large zoom: a 2 Mpc/h sphere at the box
# centre, mean overdensity 2.5 and zero angular momentum
LARGE_MODS = dict(centre=(100.0, 100.0, 100.0), radius=2.0,
                  mods=[("overdensity", 2.5), ("Lx", 0.0), ("Ly", 0.0), ("Lz", 0.0)])

WORKLOADS = {
    # base box, base n, zooms (x0, y0, z0, zoomfac, n) relative to the level above
    "zoom1024_2lpt": dict(box=200.0, n=1024, zooms=[(448, 448, 448, 8, 256), (64, 64, 64, 2, 256)], lpt=2,
                          mods=LARGE_MODS,
                          desc="1024^3 base in 200 Mpc/h + 256^3 zooms at refinement 2 and 4 around the box centre, "
                               "ZA + 2LPT, overdensity 2.5 and Lx=Ly=Lz=0 in a 2 Mpc/h sphere on the innermost level"),
    "zoom1024": dict(box=200.0, n=1024, zooms=[(448, 448, 448, 8, 256), (64, 64, 64, 2, 256)],
                     desc="1024^3 base in 200 Mpc/h + 256^3 zooms at refinement 2 and 4 around the box centre, ZA"),
    "single256": dict(box=50.0, n=256, zooms=[], desc="256^3 single level in 50 Mpc/h, ZA"),
    "zoom128": dict(box=50.0, n=128, zooms=[(48, 48, 48, 4, 128)],
                    desc="128^3 base in 50 Mpc/h + 128^3 zoom at refinement 4 centred at (25,25,25), ZA"),
}


def make_context(wl, precision):
    from genetic_amd import gic
    ic = gic.ICGenerator(precision=precision)
    ic.set_cosmology(**COSMO)
    ic.set_camb(EH_TABLE)
    ic.init_base_grid(wl["box"], wl["n"])
    for z in wl["zooms"]:
        ic.init_zoom_grid_with_origin_at(*z)
    if wl.get("lpt", 1) == 2:
        ic.set_lpt_order(2)
    return ic


def sphere_cells(ic, level, centre, radius):
    """Cells of a level within radius of centre (select_sphere, ic.hpp:1490-1500)."""
    n, _, off, dx = ic.level_geometry(level)
    box0 = ic.level_geometry(0)[1]
    return sphere_cells_geom(n, dx, off, box0, centre, radius)


def sphere_cells_geom(n, dx, off, box0, centre, radius):
    import numpy as np
    idx = np.arange(n)
    d = []
    for a in range(3):
        r = idx * dx + off[a] + dx / 2 - centre[a]
        r = np.where(r > box0 / 2, r - box0, r)
        d.append(np.where(r < -box0 / 2, r + box0, r))
    r2 = np.float64(np.float32(radius) * np.float32(radius))
    inside = d[0][:, None, None] ** 2 + d[1][None, :, None] ** 2 + d[2][None, None, :] ** 2 < r2
    return np.flatnonzero(inside.ravel())


def progress(msg):
    """A progress line on stderr (the JSON result goes to stdout)."""
    print("[bench] %s" % msg, file=sys.stderr, flush=True)


def one_step(ic, seed, mods):
    ic.reseed(seed)
    ic.draw()
    if mods:
        ic.clear_modifications()
        for name, target in mods["mods"]:
            ic.modify(name, False, target)
        ic.apply_modifications()
    ic.apply_power_spectrum()
    ic.zeldovich()


def cpu_sample_workload(wl, scale):
    """The CPU sample of a zoom workload: the same nesting, refinements, box
    resolution and constraints with every grid n/scale cells a side (the box
    shrinks with it, so cell sizes are unchanged); the constraint sphere
    shrinks to fit the innermost level."""
    n = wl["n"] // scale
    zooms = []
    for (x0, y0, z0, fac, nz) in wl["zooms"]:
        zooms.append((x0 // scale, y0 // scale, z0 // scale, fac, nz // scale))
    out = dict(box=wl["box"] / scale, n=n, zooms=zooms, lpt=wl.get("lpt", 1))
    if wl.get("mods"):
        m = wl["mods"]
        out["mods"] = dict(centre=tuple(c / scale for c in m["centre"]), radius=m["radius"] / scale,
                           mods=m["mods"])
    return out


def host_cpu():
    """lscpu's model name, sockets, cores per socket and threads per core."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for ln in out.splitlines():
            if ":" in ln:
                k, v = ln.split(":", 1)
                info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return dict(model=info.get("Model name"), sockets=info.get("Socket(s)"),
                cores_per_socket=info.get("Core(s) per socket"), threads_per_core=info.get("Thread(s) per core"),
                cpus=info.get("CPU(s)"))


def gpu_fft_ms(n=256, precision=32, reps=20):
    """Median wall ms of an in-place r2c + c2r pair of a padded n^3 field on
    the GPU (synchronous calls), after 3 warm-up pairs."""
    import numpy as np
    import torch
    from genetic_amd import gic
    g = gic.ICGenerator(precision=precision)
    dt = torch.float32 if precision == 32 else torch.float64
    buf = torch.randn(n * n * (n + 2), dtype=dt, device="cuda")
    torch.cuda.synchronize()
    ts = []
    for r in range(reps + 3):
        t0 = time.perf_counter()
        g.fft_inplace(buf.data_ptr(), n, -1)
        g.fft_inplace(buf.data_ptr(), n, +1)
        g.synchronize()
        if r >= 3:
            ts.append(time.perf_counter() - t0)
    g.close()
    return float(np.median(ts)) * 1e3


def cpu_baseline(wl, scale, reps=5):
    """The oracle (FP64 C++ restatement of the reference's stage order:
    serial real-space stream, threaded FFTs, modifications, sqrt P,
    Zel'dovich, 2LPT, level combination) timed on the workload's CPU sample
    on all of the host's cores: one untimed realisation, then the median of
    `reps`.  On AVX-512 hosts the x86-64-v4 build of the oracle runs.  Also
    an FFT-only comparison: an r2c + c2r pair at 256^3 on the host (the
    oracle's own FFT; FFTW is not installed) against the GPU's."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    build = "liboracle_v4.so" if po.host_has_avx512() else "liboracle.so"
    po.lib(build)
    s = cpu_sample_workload(wl, scale)
    tab, nc = po.read_camb(EH_TABLE)
    times = []
    ncells = 0
    for r in range(reps + 1):
        # the oracle, like the reference, seeds once per generator: one
        # generator per realisation, set up outside the timed region
        o = po.OracleIC()
        o.set_cosmology(Om=COSMO["Om"], Ol=COSMO["Ol"], s8=COSMO["s8"], ns=COSMO["ns"], z=COSMO["z"],
                        hubble=COSMO["hubble"])
        o.set_camb(tab, nc)
        o.init_base_grid(s["box"], s["n"])
        geo = [(s["n"], s["box"] / s["n"], (0.0, 0.0, 0.0))]
        for (x0, y0, z0, fac, nz) in s["zooms"]:
            o.init_zoom_grid_at(x0, y0, z0, fac, nz)
            na, dxa, offa = geo[-1]
            dxz = dxa * (na // fac) / nz
            geo.append((nz, dxz, tuple(offa[i] + (x0, y0, z0)[i] * dxa for i in range(3))))
        o.set_lpt_order(s["lpt"])
        ncells = sum(g[0] ** 3 for g in geo)
        if s.get("mods"):
            m = s["mods"]
            nz, dxz, offz = geo[-1]
            o.set_flags(len(geo) - 1, sphere_cells_geom(nz, dxz, offz, s["box"], m["centre"], m["radius"]))
        t0 = time.perf_counter()
        o.seed(1 + r)
        o.draw()
        if s.get("mods"):
            for name, target in s["mods"]["mods"]:
                o.add_modification(name, False, target)
            o.apply_modifications()
        o.apply_power_spectrum()
        o.zeldovich()
        if r > 0:
            times.append(time.perf_counter() - t0)
        del o
    med = float(np.median(times))
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    desc = "%d^3 in %.1f Mpc/h + zooms %s" % (s["n"], s["box"], ", ".join("%d^3" % z[4] for z in s["zooms"]))
    cpu_fft = po.time_fft(256, 3)
    gpu_fft32 = gpu_fft_ms(256, 32)
    gpu_fft64 = gpu_fft_ms(256, 64)
    return dict(value=round(ncells / med / 1e6, 3), unit="Mparticles/s", cores=threads, kind="port",
                host=host_cpu(), build=build,
                sample="oracle FP64 (%s), the workload scaled by 1/%d per side (%s, same cell sizes, %s, "
                       "constraint sphere %.2f Mpc/h): %d cells, median of %d realisations after 1 warm-up, "
                       "%.2f s each"
                       % (build, scale, desc, "2LPT" if s["lpt"] == 2 else "ZA",
                          s["mods"]["radius"] if s.get("mods") else 0.0, ncells, reps, med),
                fft_only=dict(n=256, cpu_ms=round(cpu_fft, 2), gpu_ms_f32=round(gpu_fft32, 4),
                              gpu_ms_f64=round(gpu_fft64, 4), speedup_f32=round(cpu_fft / gpu_fft32, 1),
                              speedup_f64=round(cpu_fft / gpu_fft64, 1),
                              note="r2c + c2r pair of a 256^3 field; CPU: the oracle's own FP64 FFT on all host "
                                   "threads (FFTW is not installed); GPU: the product's transforms, wall time of "
                                   "synchronous calls"))


def single_level_throughput(name="single256", steps=20, warmup=3, precision=32):
    """The metric's 256^3 point: seed -> noise -> delta -> Zel'dovich psi on
    one 256^3 level, resident in HBM, steps after warmup, wall clock."""
    wl = WORKLOADS[name]
    ic = make_context(wl, precision)
    for w in range(warmup):
        one_step(ic, 5000 + w, None)
    ic.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        one_step(ic, 1 + k, None)
    ic.synchronize()
    dt = (time.perf_counter() - t0) / steps
    ic.close()
    return dict(value=round(wl["n"] ** 3 / dt / 1e6, 2), unit="Mparticles/s", ms_per_step=round(dt * 1e3, 4),
                steps=steps, config=wl["desc"] + ", FP32, seeds 1..%d real space" % steps)


def adjoint_ratio(n=256, box=50.0, reps=20):
    """Forward w -> psi and its transpose psi_bar -> w_bar on one n^3 FP32
    level (dense device arrays in and out): median wall time of reps
    synchronous calls each after 3 warm-up calls."""
    import numpy as np
    import torch
    from genetic_amd import gic
    g = gic.ICGenerator(precision=32)
    g.set_cosmology(**COSMO)
    g.set_camb(EH_TABLE)
    g.init_base_grid(box, n)
    gen = torch.Generator(device="cuda").manual_seed(7)
    w = torch.randn((n, n, n), generator=gen, device="cuda", dtype=torch.float32)
    psi = [torch.empty_like(w) for _ in range(3)]
    bar = [torch.randn((n, n, n), generator=gen, device="cuda", dtype=torch.float32) for _ in range(3)]
    wbar = torch.empty_like(w)
    torch.cuda.synchronize()
    fwd = lambda: g.displacement_forward([w.data_ptr()], [p.data_ptr() for p in psi])
    adj = lambda: g.displacement_adjoint([b.data_ptr() for b in bar], [wbar.data_ptr()])
    res = {}
    for name, fn in (("forward", fwd), ("adjoint", adj)):
        for _ in range(3):
            fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        res[name] = float(np.median(ts)) * 1e3
    g.close()
    return dict(forward_ms=round(res["forward"], 4), adjoint_ms=round(res["adjoint"], 4),
                ratio=round(res["adjoint"] / res["forward"], 3),
                config="%d^3 in %.0f Mpc/h, FP32, dense device arrays in and out, median of %d" % (n, box, reps))


def psi_precision(wl):
    """rms over all cells of |psi_fp32 - psi_fp64| / dx_level: the workload's
    Zel'dovich displacements from the same white noise in both precisions."""
    import torch
    from genetic_amd import gic
    ctx = {}
    out = {}
    gen = torch.Generator(device="cuda").manual_seed(11)
    for prec in (64, 32):
        ctx[prec] = make_context(dict(wl, lpt=1), prec)
    sizes = [wl["n"]] + [z[4] for z in wl["zooms"]]
    noise = [torch.randn((n, n, n), generator=gen, device="cuda", dtype=torch.float64) for n in sizes]
    psi = {}
    for prec in (64, 32):
        dt = torch.float64 if prec == 64 else torch.float32
        w = [t.to(dt) for t in noise]
        p = [torch.empty_like(w[l]) for l in range(len(w)) for _ in range(3)]
        torch.cuda.synchronize()
        ctx[prec].displacement_forward([t.data_ptr() for t in w], [q.data_ptr() for q in p])
        psi[prec] = p
        del w
    sq, cells, worst = 0.0, 0, 0.0
    for l, n in enumerate(sizes):
        dx = ctx[64].level_geometry(l)[3]
        d2 = sum((psi[32][3 * l + a].double() - psi[64][3 * l + a]) ** 2 for a in range(3))
        sq += float(d2.sum()) / dx ** 2
        worst = max(worst, float(d2.max().sqrt()) / dx)
        cells += n ** 3
    for c in ctx.values():
        c.close()
    return dict(rms_cells=float("%.4g" % (sq / cells) ** 0.5), max_cells=float("%.4g" % worst),
                config="Zel'dovich psi of the workload's levels, same white noise in FP32 and FP64")


def splitmix64(state):
    """The synthetic-choice stream: splitmix64 (Steele, Lea and Flood 2014)."""
    M = (1 << 64) - 1
    while True:
        state = (state + 0x9E3779B97F4A7C15) & M
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        yield z ^ (z >> 31)


def batch_constraints(nreal=64, nmods=16, box=50.0, seed=20261202):
    """Per realisation, nmods top-hat overdensity constraints: centre
    U[0, box)^3, radius U[0.5, 4] Mpc/h, target N(0, 1.5) (Box-Muller), all
    from one splitmix64 stream."""
    import math
    g = splitmix64(seed)
    u = lambda: (next(g) >> 11) * 2.0 ** -53
    out = []
    for _ in range(nreal):
        mods = []
        for _ in range(nmods):
            centre = (u() * box, u() * box, u() * box)
            radius = 0.5 + 3.5 * u()
            u1, u2 = u(), u()
            target = 1.5 * math.sqrt(-2.0 * math.log(1.0 - u1)) * math.cos(2 * math.pi * u2)
            mods.append((centre, radius, target))
        out.append(mods)
    return out


def batch_throughput(nreal=64, n=128, box=50.0, nmods=16, precision=32, lanes=8, reps=3):
    """The batching workload: nreal independent n^3 realisations (real-space
    seeds 1..nreal), each with nmods overdensity constraints, then delta and
    Zel'dovich psi; FP32 fields, FP64 Gram solve.  The batch runs as one HIP
    graph per call (gic_batch_constrained, lanes concurrent chains), graph
    recording included in the time; the same work one realisation at a time
    through the ordinary calls is timed beside it."""
    import numpy as np
    import torch
    wl = dict(box=box, n=n, zooms=[])
    ic = make_context(wl, precision)
    plan = batch_constraints(nreal, nmods, box)
    cons = np.array([[(c[0], c[1], c[2], r, t) for c, r, t in mods] for mods in plan])
    seeds = np.arange(1, nreal + 1, dtype=np.uint64)
    out = torch.empty((nreal, 4, n, n, n), dtype=torch.float32 if precision == 32 else torch.float64,
                      device="cuda")
    ic.batch_constrained(seeds, cons, out.data_ptr(), lanes=lanes, graph=True)   # warm-up: buffers, tables
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        ic.batch_constrained(seeds, cons, out.data_ptr(), lanes=lanes, graph=True)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    ok = bool(torch.isfinite(out).all())

    # the same realisations through the one-at-a-time calls
    cells = [[sphere_cells(ic, 0, c, r) for c, r, _ in mods] for mods in plan]
    t1 = time.perf_counter()
    for k, mods in enumerate(plan):
        ic.reseed(k + 1)
        ic.draw()
        ic.clear_modifications()
        for (centre, radius, target), cl in zip(mods, cells[k]):
            ic.set_flags(0, cl)
            ic.modify("overdensity", False, target)
        ic.apply_modifications()
        ic.apply_power_spectrum()
        ic.zeldovich()
    ic.synchronize()
    dts = time.perf_counter() - t1
    ic.close()
    del out
    return dict(value=round(nreal * n ** 3 / dt / 1e6, 2), unit="Mparticles/s",
                ms_per_batch=round(dt * 1e3, 3), ms_per_realisation=round(dt / nreal * 1e3, 3),
                lanes=lanes, graph=True, outputs_finite=ok,
                one_at_a_time=dict(value=round(nreal * n ** 3 / dts / 1e6, 2),
                                   ms_per_realisation=round(dts / nreal * 1e3, 3)),
                config="%d realisations of %d^3 in %.0f Mpc/h (seeds 1..%d), %d overdensity constraints each "
                       "(splitmix64 seed 20261202), FP32 fields, FP64 Gram solve, one HIP graph per batch"
                       % (nreal, n, box, nreal, nmods))


def pmc_traffic(workload, kernel):
    """HBM bytes per launch of the roofline kernel from the committed rocprofv3
    counter summary (FETCH_SIZE + WRITE_SIZE, scaled by the same counters'
    reading for a plain copy of known size), or None when not measured."""
    path = os.path.join(ROOT, "profiles", "%s_hbm_traffic.md" % workload)
    if not os.path.exists(path):
        return None
    for ln in open(path):
        if ln.startswith("traffic_bytes_per_launch[%s]" % kernel):
            return float(ln.split("=")[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="zoom1024_2lpt", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", type=int, default=32, choices=[32, 64])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the adjoint/forward ratio and the FP32-vs-FP64 displacement error")
    ap.add_argument("--cpu-scale", type=int, default=8,
                    help="the CPU baseline runs the workload with every grid 1/scale as many cells a side")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device copy (reference reading for the FETCH_SIZE/WRITE_SIZE counters)")
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the batching workload (64 constrained 128^3 realisations per HIP graph)")
    args = ap.parse_args()
    if args.batch_only:
        print(json.dumps(batch_throughput()))
        return
    if args.copy_calibration:
        import torch
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps(dict(copy_bytes=x.numel() * 4, launches=3)))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: the field-generation path runs on a single MI355X "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    wl = WORKLOADS[args.workload]
    ic = make_context(wl, args.precision)
    ncells = wl["n"] ** 3 + sum(z[4] ** 3 for z in wl["zooms"])
    mods = wl.get("mods")
    if mods:
        finest = len(wl["zooms"])
        ic.set_flags(finest, sphere_cells(ic, finest, mods["centre"], mods["radius"]))
    progress("context ready: %s, %d cells" % (args.workload, ncells))
    for w in range(args.warmup):
        one_step(ic, 1000 + w, mods)
    progress("%d warm-up steps done" % args.warmup)
    ic.set_event_timing(True)
    ic.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        one_step(ic, 1 + k, mods)
    ic.synchronize()
    t1 = time.perf_counter()
    timings = ic.event_timings(with_bytes=True)
    dt = (t1 - t0) / args.steps
    progress("%d timed steps: %.3f ms per step" % (args.steps, dt * 1e3))

    # Roofline of the dominant kernel: of every timed launch that states its
    # algorithmic HBM bytes (each input element read once, each output
    # element written once; a windowed pass counts its whole input and the
    # window of its output), the kernel with the largest total time.  Its
    # achieved rate is those bytes over its mean launch time, measured with
    # hipEvents on the context's stream around each launch.
    kern = {k: v for k, v in timings.items() if v[2] > 0}
    dom = max(kern, key=lambda k: kern[k][0])
    ms, cnt, nbytes = kern[dom]
    avg_ms = ms / cnt
    per_launch = nbytes / cnt
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    roofline = dict(bound="hbm", kernel=dom, achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=pmc_traffic(args.workload, dom),
                    bytes_per_launch=round(per_launch), avg_launch_ms=round(avg_ms, 4))
    # every byte-stating kernel, largest total time first
    top = sorted(kern.items(), key=lambda kv: -kv[1][0])
    kernels = {k: dict(ms_per_step=round(v[0] / args.steps, 4), launches_per_step=round(v[1] / args.steps, 2),
                       GBps=round(v[2] / (v[0] * 1e-3) / 1e9, 1), frac=round(v[2] / (v[0] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4))
               for k, v in top}
    # the whole step: the algorithmic bytes of all its timed kernels over the
    # step time, and the share of the step those kernels cover
    step_bytes = sum(v[2] for v in kern.values()) / args.steps
    step_kernel_ms = sum(v[0] for v in kern.values()) / args.steps
    step = dict(algorithmic_bytes=round(step_bytes), GBps=round(step_bytes / dt / 1e9, 1),
                frac=round(step_bytes / dt / 1e9 / HBM_PEAK_GBS, 4),
                kernels_ms=round(step_kernel_ms, 3), step_ms=round(dt * 1e3, 3))
    rng_ms = sum(timings[k][0] for k in ("rng_summary", "rng_scan", "rng_emit") if k in timings) / args.steps
    rng = dict(gaussians_per_s=round(ncells / (rng_ms * 1e-3), 1) if rng_ms > 0 else None, ms_per_step=round(rng_ms, 4),
               note="integer-ALU bound (ranlxs2 at luxury 397: 16.5 48-bit steps per 24-bit number)")
    stages = {k: round(v[0] / args.steps, 4) for k, v in sorted(timings.items())}
    out = dict(
        metric=METRIC,
        value=round(ncells / dt / 1e6, 2),
        unit="Mparticles/s",
        n_gpus=1,
        steps=args.steps,
        warmup=args.warmup,
        ms_per_step=round(dt * 1e3, 3),
        higher_is_better=True,
        scaling="weak",
        vs_baseline=None,
        dtype="f32" if args.precision == 32 else "f64",
        data="synthetic (seeded white noise; Eisenstein-Hu no-wiggle T(k) on 512 log-spaced k)",
        config=dict(workload=args.workload, description=wl["desc"], cells=ncells,
                    cosmology=COSMO, seeds="1..%d real space" % args.steps),
        roofline=roofline,
        step_roofline=step,
        rng=rng,
        kernels=kernels,
        stage_ms_per_step=stages,
    )
    if mods and not args.no_extras:
        # the same workload without the modifications (SURVEY C5 as stated)
        nm_steps = max(2, args.steps // 2)
        one_step(ic, 3000, None)
        ic.synchronize()
        t2 = time.perf_counter()
        for k in range(nm_steps):
            one_step(ic, 2000 + k, None)
        ic.synchronize()
        dtn = (time.perf_counter() - t2) / nm_steps
        out["without_modifications"] = dict(value=round(ncells / dtn / 1e6, 2), unit="Mparticles/s",
                                            ms_per_step=round(dtn * 1e3, 3), steps=nm_steps)
        progress("without modifications: %.3f ms per step" % (dtn * 1e3))
    ic.close()
    if not args.no_extras:
        out["throughput_256"] = single_level_throughput()
        progress("single-level 256^3 done")
        out["adjoint_over_forward"] = adjoint_ratio()
        progress("adjoint ratio done")
        out["batch"] = batch_throughput()
        progress("batch done")
        out["psi_fp32_error"] = psi_precision(wl)
        progress("FP32 displacement error done")
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(wl, args.cpu_scale)
        progress("CPU baseline done")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
