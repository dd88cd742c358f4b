"""ctypes bindings of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The product never imports this.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

SEED_REAL, SEED_FOURIER_SERIAL, SEED_FOURIER_PARALLEL, SEED_FOURIER_REVERSE = 0, 1, 2, 3


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def host_has_avx512():
    """The x86-64-v4 feature set (AVX-512 F, BW, CD, DQ, VL) on this host."""
    try:
        flags = set()
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("flags"):
                flags = set(ln.split(":", 1)[1].split())
                break
        return {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"} <= flags
    except OSError:
        return False


def lib(name=None):
    """Load the oracle library once: liboracle.so, or the build named by
    `name` (or GIC_ORACLE_LIB), e.g. liboracle_v4.so on AVX-512 hosts."""
    global _LIB
    if _LIB is None:
        name = name or os.environ.get("GIC_ORACLE_LIB", "liboracle.so")
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        d, i, u, v, s = ctypes.c_double, ctypes.c_int, ctypes.c_ulong, ctypes.c_void_p, ctypes.c_size_t
        P = ctypes.POINTER
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_ic_create.restype = v
        L.orc_ic_destroy.argtypes = [v]
        sigs = {
            "orc_ranlxs_raw": [u, s, v],
            "orc_gaussians": [u, s, v, v],
            "orc_zig_tables": [v, v, v],
            "orc_ranlxs_raw_from_state": [v, i, s, v],
            "orc_gaussians_from_state": [v, i, s, s, v],
            "orc_fft_r2c": [i, v],
            "orc_fft_c2r": [i, v],
            "orc_growth_factor": [d, d, d, P(d)],
            "orc_ic_set_cosmology": [v, d, d, d, d, d, d, d],
            "orc_ic_set_camb": [v, v, s, i],
            "orc_ic_camb_info": [v, P(d), P(d)],
            "orc_ic_pk": [v, i, v, v],
            "orc_ic_base_grid": [v, d, i],
            "orc_ic_zoom_at": [v, i, i, i, i, i],
            "orc_ic_seed": [v, u, i],
            "orc_ic_draw": [v],
            "orc_ic_zero_level": [v, i],
            "orc_ic_apply_ps": [v],
            "orc_ic_zeldovich": [v],
            "orc_ic_get_field": [v, i, i, i, v],
            "orc_ic_particles": [v, i, v, v, P(d), P(d)],
            "orc_ic_velocity_ratio": [v, P(d)],
            "orc_ic_set_flags": [v, i, v, s],
            "orc_ic_set_lpt_order": [v, i],
            "orc_ic_set_field": [v, i, v],
            "orc_ic_set_variance_scale": [v, d],
            "orc_ic_calculate": [v, ctypes.c_char_p, P(d)],
            "orc_ic_add_modification": [v, ctypes.c_char_p, i, d],
            "orc_ic_apply_modifications": [v, P(d), P(d)],
            "orc_ic_clear_modifications": [v],
            "orc_ic_chi2": [v, P(d), P(s)],
            "orc_time_pipeline": [d, i, u, v, s, i, d, d, d, d, d, P(d), P(d)],
            "orc_time_fft": [i, i, P(d)],
        }
        for name, args in sigs.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def read_camb(path):
    """Flat numeric table and column count, as io::getBufferIgnoringColumnHeaders
    and io::getNumberOfColumns read it (genetIC/src/io/input.hpp:33-84)."""
    with open(path) as f:
        lines = f.read().splitlines()
    vals = []
    first = lines[0].split()
    try:
        vals.extend(float(x) for x in first)
    except ValueError:
        pass
    for ln in lines[1:]:
        vals.extend(float(x) for x in ln.split())
    ncols = len(lines[1].split())
    return np.array(vals, dtype=np.float64), ncols


def time_fft(n, reps=3):
    """Median ms of an r2c + c2r pair of a random n^3 field (host threads)."""
    ms = ctypes.c_double()
    _check(lib().orc_time_fft(n, reps, ctypes.byref(ms)))
    return ms.value


def ranlxs_raw(seed, n):
    out = np.empty(n, dtype=np.uint32)
    _check(lib().orc_ranlxs_raw(seed, n, _ptr(out)))
    return out


def gaussians(seed, n):
    out = np.empty(n, dtype=np.float64)
    used = ctypes.c_size_t(0)
    _check(lib().orc_gaussians(seed, n, _ptr(out), ctypes.byref(used)))
    return out, used.value


def ranlxs_raw_from_state(words, carry, n):
    """n raw numbers of the serial generator continued from a block boundary
    with the 12 given 48-bit words (oldest first) and carry."""
    w = np.ascontiguousarray(words, dtype=np.int64)
    out = np.empty(n, dtype=np.uint32)
    _check(lib().orc_ranlxs_raw_from_state(_ptr(w), int(carry), n, _ptr(out)))
    return out


def gaussians_from_state(words, carry, skip, n):
    """n ziggurat Gaussians parsed from that state after skipping `skip` raw numbers."""
    w = np.ascontiguousarray(words, dtype=np.int64)
    out = np.empty(n, dtype=np.float64)
    _check(lib().orc_gaussians_from_state(_ptr(w), int(carry), skip, n, _ptr(out)))
    return out


def zig_tables():
    y = np.empty(128)
    w = np.empty(128)
    k = np.empty(128, dtype=np.uint32)
    _check(lib().orc_zig_tables(_ptr(y), _ptr(w), _ptr(k)))
    return y, w, k


def fft_r2c(real):
    """real: (N,N,N) float64 -> complex (N,N,N/2+1), reference normalisation."""
    n = real.shape[0]
    buf = np.zeros(2 * n * n * (n // 2 + 1))
    buf[: n ** 3] = real.ravel()
    _check(lib().orc_fft_r2c(n, _ptr(buf)))
    return buf.view(np.complex128).reshape(n, n, n // 2 + 1)


def fft_c2r(spec):
    n = spec.shape[0]
    buf = np.ascontiguousarray(spec, dtype=np.complex128).view(np.float64).ravel().copy()
    _check(lib().orc_fft_c2r(n, _ptr(buf)))
    return buf[: n ** 3].reshape(n, n, n)


class OracleIC:
    """The reference ICGenerator's field path, restated on the CPU in FP64."""

    def __init__(self):
        self.h = lib().orc_ic_create()
        self.sizes = []

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_ic_destroy(self.h)
            self.h = None

    def set_cosmology(self, Om, Ol, s8, ns, z, hubble=0.701, Ob=0.0):
        _check(lib().orc_ic_set_cosmology(self.h, Om, Ol, Ob, hubble, z, s8, ns))

    def set_camb(self, table, ncols):
        table = np.ascontiguousarray(table, dtype=np.float64)
        _check(lib().orc_ic_set_camb(self.h, _ptr(table), table.size, ncols))

    def camb_info(self):
        a, s = ctypes.c_double(), ctypes.c_double()
        _check(lib().orc_ic_camb_info(self.h, ctypes.byref(a), ctypes.byref(s)))
        return a.value, s.value

    def pk(self, k):
        k = np.ascontiguousarray(k, dtype=np.float64)
        out = np.empty_like(k)
        _check(lib().orc_ic_pk(self.h, k.size, _ptr(k), _ptr(out)))
        return out

    def init_base_grid(self, box, n):
        _check(lib().orc_ic_base_grid(self.h, box, n))
        self.sizes.append(n)

    def init_zoom_grid_at(self, x0, y0, z0, fac, n):
        _check(lib().orc_ic_zoom_at(self.h, x0, y0, z0, fac, n))
        self.sizes.append(n)

    def seed(self, seed, mode=SEED_REAL):
        _check(lib().orc_ic_seed(self.h, seed, mode))

    def draw(self):
        _check(lib().orc_ic_draw(self.h))

    def zero_level(self, level):
        _check(lib().orc_ic_zero_level(self.h, level))

    def apply_power_spectrum(self):
        _check(lib().orc_ic_apply_ps(self.h))

    def zeldovich(self):
        _check(lib().orc_ic_zeldovich(self.h))

    def field(self, which, level, fourier=False):
        n = self.sizes[level]
        if fourier:
            out = np.empty(2 * n * n * (n // 2 + 1))
            _check(lib().orc_ic_get_field(self.h, which, level, 1, _ptr(out)))
            return out.view(np.complex128).reshape(n, n, n // 2 + 1)
        out = np.empty(n ** 3)
        _check(lib().orc_ic_get_field(self.h, which, level, 0, _ptr(out)))
        return out.reshape(n, n, n)

    def particles(self, level):
        n = self.sizes[level]
        pos = np.empty((n ** 3, 3))
        vel = np.empty((n ** 3, 3))
        m, e = ctypes.c_double(), ctypes.c_double()
        _check(lib().orc_ic_particles(self.h, level, _ptr(pos), _ptr(vel), ctypes.byref(m), ctypes.byref(e)))
        return pos, vel, m.value, e.value

    def set_lpt_order(self, order):
        _check(lib().orc_ic_set_lpt_order(self.h, int(order)))

    def set_field(self, level, data):
        d = np.ascontiguousarray(np.asarray(data, dtype=np.float64).ravel())
        _check(lib().orc_ic_set_field(self.h, level, _ptr(d)))

    # --- modifications ---------------------------------------------------
    def set_flags(self, level, cells):
        c = np.ascontiguousarray(np.asarray(cells, dtype=np.uint64))
        _check(lib().orc_ic_set_flags(self.h, level, _ptr(c), c.size))

    def set_variance_scale(self, scale):
        _check(lib().orc_ic_set_variance_scale(self.h, scale))

    def calculate(self, name):
        out = ctypes.c_double()
        _check(lib().orc_ic_calculate(self.h, name.encode(), ctypes.byref(out)))
        return out.value

    def add_modification(self, name, relative, target):
        _check(lib().orc_ic_add_modification(self.h, name.encode(), int(relative), float(target)))

    def apply_modifications(self):
        pre, post = ctypes.c_double(), ctypes.c_double()
        _check(lib().orc_ic_apply_modifications(self.h, ctypes.byref(pre), ctypes.byref(post)))
        return pre.value, post.value

    def clear_modifications(self):
        _check(lib().orc_ic_clear_modifications(self.h))

    def chi2(self):
        out, dof = ctypes.c_double(), ctypes.c_size_t()
        _check(lib().orc_ic_chi2(self.h, ctypes.byref(out), ctypes.byref(dof)))
        return out.value, dof.value

    def velocity_ratio(self):
        r = ctypes.c_double()
        _check(lib().orc_ic_velocity_ratio(self.h, ctypes.byref(r)))
        return r.value
