"""Tile-shape tuner of the strided FFT passes.

The strided y and x passes run tiles of `cols` adjacent kz columns on blocks
of `threads` threads; the best shape depends on the size and the precision.
The tuner works offline, in three steps, so that the library never switches
kernels at run time:

    python -m genetic_amd.tune_fft build            # compile one timing library per candidate shape
    python -m genetic_amd.tune_fft measure OUT.json # on the GPU: time every shape at every size
    python -m genetic_amd.tune_fft apply OUT.json   # write csrc/fft_tiles.h from the timings

`build` links csrc/fft.hip, compiled for one shape (GIC_TUNE_THREADS,
GIC_TUNE_COLS), with the harness csrc/fft_tune.hip into _tune/libfft_t*_c*.so.
`measure` loads each library in a process of its own and times, on a padded
n^3 field, the forms of the transform the path uses: forward + inverse, the
Zel'dovich inverse (i k/k^2, dense output), a windowed inverse with a
low-pass filter and a windowed forward that adds a filtered spectrum.  The
objective is their sum (medians of three timings of `reps` calls each); a
shape whose round trip does not give the field back is discarded.  `apply`
keeps, per precision and size, the fastest shape if it beats the default
(512 threads, 8 columns FP32 / 4 FP64) by more than 2 %, and the library is
then rebuilt with the table (make -C genetic_amd/csrc).
"""
import ctypes
import itertools
import json
import os
import statistics
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
TUNE = os.path.join(HERE, "_tune")
THREADS = (256, 512, 1024)
COLS = (4, 8, 16)
SIZES = (64, 128, 256, 512, 1024)
LDS_BYTES = 160 * 1024          # LDS per CU on gfx950
TOL = {32: 1e-5, 64: 1e-12}     # round-trip error bound on values in [-1, 1)


def candidates():
    return list(itertools.product(THREADS, COLS))


def lib_path(t, c):
    return os.path.join(TUNE, "libfft_t%d_c%d.so" % (t, c))


def fits(t, c, n, prec):
    """The tile (cols x n complex) and the line's n twiddles fit the LDS, and
    FP64 tiles take at most 8 columns."""
    cols = c if prec == 32 else min(c, 8)
    return (cols + 1) * n * (8 if prec == 32 else 16) <= LDS_BYTES


def build(jobs=6):
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(TUNE, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

    def one(tc):
        t, c = tc
        cmd = [hipcc, "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-DGIC_TUNE_THREADS=%d" % t, "-DGIC_TUNE_COLS=%d" % c,
               os.path.join(CSRC, "fft.hip"), os.path.join(CSRC, "fft_tune.hip"), "-o", lib_path(t, c)]
        return subprocess.run(cmd).returncode

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        rc = list(ex.map(one, candidates()))
    if any(rc):
        raise SystemExit("tune_fft build: a compile failed")
    print("built %d libraries in %s" % (len(candidates()), TUNE))


def time_one(path, reps=5):
    """Run in a process of its own: every size and precision the shape fits."""
    L = ctypes.CDLL(path)
    L.gic_tune_fft.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                               ctypes.POINTER(ctypes.c_double)]
    base = os.path.basename(path)
    t = int(base.split("_t")[1].split("_")[0])
    c = int(base.split("_c")[1].split(".")[0])
    out = []
    for prec in (32, 64):
        for n in SIZES:
            if not fits(t, c, n, prec):
                continue
            ms, err = (ctypes.c_double * 4)(), ctypes.c_double()
            runs = []
            for _ in range(3):
                rc = L.gic_tune_fft(n, prec, reps, ms, ctypes.byref(err))
                if rc != 0:
                    break
                runs.append(list(ms))
            ok = len(runs) == 3 and err.value < TOL[prec]
            parts = [statistics.median(r[i] for r in runs) for i in range(4)] if runs else None
            out.append(dict(threads=t, cols=c if prec == 32 else min(c, 8), prec=prec, n=n, parts=parts,
                            ms=sum(parts) if parts else None, err=err.value, ok=ok))
            print("t=%d c=%d prec=%d n=%d ms=%s parts=%s err=%.2e" % (t, c, prec, n, out[-1]["ms"], parts, err.value),
                  flush=True)
    return out


def measure(out_path):
    results = []
    for t, c in candidates():
        path = lib_path(t, c)
        if not os.path.exists(path):
            raise SystemExit("tune_fft measure: %s is not built (run the build step first)" % path)
        p = subprocess.run([sys.executable, "-m", "genetic_amd.tune_fft", "_one", path], capture_output=True,
                           text=True, timeout=600, cwd=os.path.dirname(HERE))
        sys.stdout.write(p.stdout)
        if p.returncode != 0:
            raise SystemExit("tune_fft measure: %s failed:\n%s" % (path, p.stderr[-2000:]))
        results += json.loads(p.stdout.strip().splitlines()[-1])
    with open(out_path, "w") as f:
        json.dump(results, f, indent=1)
    print("wrote", out_path)


DEFAULT = {32: (512, 8), 64: (512, 4)}
# the y passes' own shapes, chosen from whole-step bench timings rather than
# by this tuner; apply() writes them back unchanged
Y_SHAPES = [
    "// The y passes (lines 4 KB apart) where they take a shape of their own:",
    "// declared only for the sizes listed after the primary template.",
    "template <typename T, int N> struct YTileShape;",
    "// float n=1024: 16-column tiles (128-byte segments, one block per CU) took",
    "// the step's y passes from 36.8 to 34.8 ms against the x passes' 8 columns",
    "template <> struct YTileShape<float, 1024> {",
    "  static constexpr int threads = 1024;",
    "  static constexpr int cols = 16;",
    "};",
]
MARGIN = 0.02   # a shape replaces the default only if it is this much faster (run-to-run noise is about 1-2 %)


def apply(in_path, header=os.path.join(CSRC, "fft_tiles.h")):
    res = [r for r in json.load(open(in_path)) if r["ok"]]
    best, default = {}, {}
    for r in res:
        key = (r["prec"], r["n"])
        if (r["threads"], r["cols"]) == DEFAULT[r["prec"]]:
            default[key] = r
        if key not in best or r["ms"] < best[key]["ms"]:
            best[key] = r
    for key in list(best):
        if key in default and best[key]["ms"] > default[key]["ms"] * (1 - MARGIN):
            del best[key]   # within the noise of the default: keep it
    lines = [
        "// Tile shapes of the strided FFT passes (threads per block, adjacent kz",
        "// columns per tile) per precision and size.  Written by",
        "// genetic_amd/tune_fft.py from timings of every candidate shape on the GPU;",
        "// sizes without an entry take the default.",
        "#pragma once",
        "",
        "namespace gic {",
        "",
    ]
    lines += Y_SHAPES
    lines += [
        "",
        "template <typename T, int N> struct TileShape {",
        "  static constexpr int threads = 512;",
        "  static constexpr int cols = sizeof(T) == 4 ? 8 : 4;",
        "};",
    ]
    for (prec, n), r in sorted(best.items()):
        tname = "float" if prec == 32 else "double"
        lines.append("// %s n=%d: %.4f ms for the four timed transforms (round trip, ZA, windowed c2r, windowed r2c)"
                     % (tname, n, r["ms"]))
        lines.append("template <> struct TileShape<%s, %d> {" % (tname, n))
        lines.append("  static constexpr int threads = %d;" % r["threads"])
        lines.append("  static constexpr int cols = %d;" % r["cols"])
        lines.append("};")
    lines += ["", "}  // namespace gic", ""]
    with open(header, "w") as f:
        f.write("\n".join(lines))
    print("wrote", header)


def main(argv):
    if len(argv) < 2:
        raise SystemExit(__doc__)
    cmd = argv[1]
    if cmd == "build":
        build()
    elif cmd == "measure":
        measure(argv[2] if len(argv) > 2 else "fft_tune.json")
    elif cmd == "apply":
        apply(argv[2])
    elif cmd == "_one":
        print(json.dumps(time_one(argv[2])))
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv)
