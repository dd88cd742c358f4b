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
    import numpy as np
    n, _, off, dx = ic.level_geometry(level)
    box0 = ic.level_geometry(0)[1]
    idx = np.arange(n)
    d = []
    for a in range(3):
        r = idx * dx + off[a] + dx / 2 - centre[a]
        r = np.where(r > box0 / 2, r - box0, r)
        d.append(np.where(r < -box0 / 2, r + box0, r))
    r2 = np.float64(np.float32(radius) * np.float32(radius))
    inside = d[0][:, None, None] ** 2 + d[1][None, :, None] ** 2 + d[2][None, None, :] ** 2 < r2
    return np.flatnonzero(inside.ravel())


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


def cpu_baseline(n=256):
    """The oracle's restatement of the reference's stage order (serial
    real-space stream, threaded FFTs, sqrt P, Zel'dovich) at n^3."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    tab, nc = po.read_camb(EH_TABLE)
    tt, tr = ctypes.c_double(), ctypes.c_double()
    po._check(po.lib().orc_time_pipeline(50.0, n, 8896131, po._ptr(tab), tab.size, nc, COSMO["Om"], COSMO["Ol"],
                                         COSMO["s8"], COSMO["ns"], COSMO["z"], ctypes.byref(tt), ctypes.byref(tr)))
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=n ** 3 / tt.value / 1e6, unit="Mparticles/s", cores=threads, kind="port",
                sample="oracle FP64 pipeline seed->delta,psi at %d^3 single level (50 Mpc/h), %.2f s total, "
                       "%.2f s in the serial RNG + first r2c" % (n, tt.value, tr.value))


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
    ap.add_argument("--cpu-n", type=int, default=256)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device copy (reference reading for the FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
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
    for w in range(args.warmup):
        one_step(ic, 1000 + w, mods)
    ic.set_event_timing(True)
    ic.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        one_step(ic, 1 + k, mods)
    ic.synchronize()
    t1 = time.perf_counter()
    timings = ic.event_timings()
    dt = (t1 - t0) / args.steps

    # roofline of the dominant HBM-bound kernel: a strided FFT pass on the
    # base grid reads and writes the padded field once per launch.
    esize = 4 if args.precision == 32 else 8
    n0 = wl["n"]
    field_bytes = n0 * n0 * (n0 + 2) * esize
    # full-field passes only (the windowed passes write a sub-volume)
    hbm = {k: v for k, v in timings.items()
           if k.startswith("fft_") and k.endswith("." + str(n0)) and "mirror" not in k and "_win" not in k}
    dom = max(hbm, key=lambda k: hbm[k][0])
    avg_ms = hbm[dom][0] / hbm[dom][1]
    achieved = 2 * field_bytes / (avg_ms * 1e-3) / 1e9
    roofline = dict(bound="hbm", kernel=dom, achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=pmc_traffic(args.workload, dom),
                    bytes_per_launch=2 * field_bytes, avg_launch_ms=round(avg_ms, 4))
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
        stage_ms_per_step=stages,
    )
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_n)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
