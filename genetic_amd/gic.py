"""ctypes binding of libgic_hip.so and a Python mirror of genetIC's
ICGenerator command surface for the field-generation path.

The method names follow the reference's paramfile commands and ICGenerator
methods (genetIC/src/main.cpp:22-172, genetIC/src/ic.hpp), so scripts and
tests read like the reference's own.  All arithmetic runs in the HIP
library; if it is missing this module raises instead of falling back.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgic_hip.so")
_LIB = None

FLOAT32, FLOAT64 = 32, 64
SEED_REAL_SPACE, SEED_FOURIER_SERIAL, SEED_FOURIER_PARALLEL, SEED_FOURIER_REVERSE = 0, 1, 2, 3
FIELD_OVERDENSITY, FIELD_PSI_X, FIELD_PSI_Y, FIELD_PSI_Z = 0, 1, 2, 3
FIELD_PSI2_X, FIELD_PSI2_Y, FIELD_PSI2_Z = 4, 5, 6
REAL_SPACE, FOURIER_SPACE = 0, 1
DTYPE_F32, DTYPE_F64 = 0, 1

# every entry point declared in include/gic.h with its argument types
_d, _i, _u, _v, _s = ctypes.c_double, ctypes.c_int, ctypes.c_ulong, ctypes.c_void_p, ctypes.c_size_t
_P = ctypes.POINTER
SIGNATURES = {
    "gic_create": ([_P(_v), _i, _i], _i),
    "gic_destroy": ([_v], None),
    "gic_last_error": ([_v], ctypes.c_char_p),
    "gic_set_cosmology": ([_v, _d, _d, _d, _d, _d, _d, _d], _i),
    "gic_set_camb_table": ([_v, _v, _s, _i], _i),
    "gic_init_base_grid": ([_v, _d, _i], _i),
    "gic_init_zoom_grid_with_origin_at": ([_v, _i, _i, _i, _i, _i], _i),
    "gic_num_levels": ([_v, _P(_i)], _i),
    "gic_level_geometry": ([_v, _i, _P(_i), _P(_d), _v, _P(_d)], _i),
    "gic_seed": ([_v, _u, _i], _i),
    "gic_reseed": ([_v, _u, _i], _i),
    "gic_draw": ([_v], _i),
    "gic_zero_level": ([_v, _i], _i),
    "gic_set_exact_power_spectrum": ([_v, _i], _i),
    "gic_reverse": ([_v], _i),
    "gic_reverse_small_k": ([_v, _d], _i),
    "gic_import_level": ([_v, _i, _v, _s], _i),
    "gic_apply_power_spectrum": ([_v], _i),
    "gic_zeldovich": ([_v], _i),
    "gic_copy_field": ([_v, _i, _i, _i, _v, _i], _i),
    "gic_set_field": ([_v, _i, _i, _i, _v, _i], _i),
    "gic_mode_power": ([_v, _i, _v, _v, _v, _s], _i),
    "gic_set_lpt_order": ([_v, _i], _i),
    "gic_set_flags": ([_v, _i, _v, _s], _i),
    "gic_set_variance_filter_scale": ([_v, _d], _i),
    "gic_calculate": ([_v, ctypes.c_char_p, _P(_d)], _i),
    "gic_add_modification": ([_v, ctypes.c_char_p, _i, _d, _i, _d], _i),
    "gic_apply_modifications": ([_v, _P(_d), _P(_d)], _i),
    "gic_clear_modifications": ([_v], _i),
    "gic_chi2": ([_v, _P(_d), _P(ctypes.c_uint64)], _i),
    "gic_displacement_forward": ([_v, _v, _v], _i),
    "gic_displacement_adjoint": ([_v, _v, _v], _i),
    "gic_variance_gradient": ([_v, _v, _P(_d), _v], _i),
    "gic_batch_constrained": ([_v, _i, _v, _i, _v, _v, _i, _i, _P(_d)], _i),
    "gic_splice": ([_v, ctypes.c_ulong, _P(_i)], _i),
    "gic_set_power_law": ([_v, _d], _i),
    "gic_set_quadratic_line_search": ([_v, _i], _i),
    "gic_particles": ([_v, _i, _v, _s, _s, _v, _v], _i),
    "gic_particle_mass_eps": ([_v, _i, _P(_d), _P(_d)], _i),
    "gic_velocity_ratio": ([_v, _P(_d)], _i),
    "gic_fft_inplace": ([_v, _v, _i, _i], _i),
    "gic_raw_stream": ([_v, _u, _v, _s], _i),
    "gic_gaussian_stream": ([_v, _u, _v, _s], _i),
    "gic_gaussian_stream_noalloc": ([_v, _u, _v, _s], _i),
    "gic_synchronize": ([_v], _i),
    "gic_set_event_timing": ([_v, _i], _i),
    "gic_get_event_timing": ([_v, ctypes.c_char_p, _P(_d), _P(ctypes.c_long)], _i),
    "gic_reset_event_timing": ([_v], _i),
    "gic_list_event_timing": ([_v, _v, _s, _P(_s)], _i),
    "gic_device_bytes": ([_v, _P(_s)], _i),
}


def lib():
    """Load libgic_hip.so (built by __graft_entry__.build()); raise if absent."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libgic_hip.so is not built (%s); run __graft_entry__.build() first" % LIB_PATH)
        # torch ships its own HIP runtime under the same soname as the
        # system's; whichever loads first serves the whole process.  Loading
        # torch first keeps one runtime for both when torch tensors are used.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def read_camb(path):
    """Numbers of a CAMB transfer file and its column count, as
    io::getBufferIgnoringColumnHeaders / io::getNumberOfColumns read them
    (genetIC/src/io/input.hpp:33-84)."""
    with open(path) as f:
        lines = f.read().splitlines()
    vals = []
    try:
        vals.extend(float(x) for x in lines[0].split())
    except ValueError:
        pass
    for ln in lines[1:]:
        vals.extend(float(x) for x in ln.split())
    if len(lines) < 2:
        raise RuntimeError("Error reading file " + path + ". Could not determine number of columns.")
    return np.asarray(vals, dtype=np.float64), len(lines[1].split())


class GICError(RuntimeError):
    pass


class ICGenerator:
    """The reference's ICGenerator field path on one GPU.

    precision 64 matches the reference's default build, 32 its -f mode.
    """

    def __init__(self, precision=FLOAT64, device=0):
        L = lib()
        h = ctypes.c_void_p()
        rc = L.gic_create(ctypes.byref(h), device, precision)
        if rc != 0:
            raise GICError("gic_create failed (code %d): no usable HIP device %d" % (rc, device))
        self._h = h
        self.precision = precision
        self.sizes = []
        # reference defaults (ic.hpp:196-206)
        self.cosmo = dict(Om=0.3, Ol=0.0, Ob=0.0, hubble=0.701, s8=0.0, ns=0.96, z=0.0)
        self._push_cosmology()

    def close(self):
        if getattr(self, "_h", None):
            if _LIB is not None:   # at interpreter exit the library may already be gone
                _LIB.gic_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != 0:
            raise GICError(lib().gic_last_error(self._h).decode())

    @property
    def handle(self):
        return self._h

    # --- cosmology ------------------------------------------------------
    def _push_cosmology(self):
        c = self.cosmo
        self._check(lib().gic_set_cosmology(self._h, c["Om"], c["Ol"], c["Ob"], c["hubble"], c["s8"], c["ns"], c["z"]))

    def set_cosmology(self, **kw):
        for k, v in kw.items():
            if k not in self.cosmo:
                raise KeyError(k)
            self.cosmo[k] = float(v)
        self._push_cosmology()

    def set_camb(self, path_or_table, ncols=None):
        if isinstance(path_or_table, (str, os.PathLike)):
            table, ncols = read_camb(path_or_table)
        else:
            table = np.ascontiguousarray(path_or_table, dtype=np.float64)
        self._check(lib().gic_set_camb_table(self._h, _ptr(table), table.size, int(ncols)))

    def set_power_law(self, amplitude):
        """powerlaw_amplitude: P(k) = amplitude k^ns (ns as set now)."""
        self._check(lib().gic_set_power_law(self._h, float(amplitude)))

    # --- grids ----------------------------------------------------------
    def init_base_grid(self, box, n):
        self._check(lib().gic_init_base_grid(self._h, float(box), int(n)))
        self.sizes.append(int(n))

    def init_zoom_grid_with_origin_at(self, x0, y0, z0, zoomfac, n):
        self._check(lib().gic_init_zoom_grid_with_origin_at(self._h, int(x0), int(y0), int(z0), int(zoomfac), int(n)))
        self.sizes.append(int(n))

    def level_geometry(self, level):
        n = ctypes.c_int()
        box = ctypes.c_double()
        dx = ctypes.c_double()
        off = np.zeros(3)
        self._check(lib().gic_level_geometry(self._h, level, ctypes.byref(n), ctypes.byref(box), _ptr(off),
                                             ctypes.byref(dx)))
        return n.value, box.value, off, dx.value

    # --- noise ----------------------------------------------------------
    def set_seed(self, seed, mode=SEED_REAL_SPACE):
        self._check(lib().gic_seed(self._h, int(seed), int(mode)))

    def reseed(self, seed, mode=SEED_REAL_SPACE):
        self._check(lib().gic_reseed(self._h, int(seed), int(mode)))

    def draw(self):
        self._check(lib().gic_draw(self._h))

    def zero_level(self, level):
        self._check(lib().gic_zero_level(self._h, int(level)))

    def fix_power(self, on=True):
        self._check(lib().gic_set_exact_power_spectrum(self._h, int(bool(on))))

    def reverse(self):
        self._check(lib().gic_reverse(self._h))

    def reverse_small_k(self, kmax):
        self._check(lib().gic_reverse_small_k(self._h, float(kmax)))

    def import_level(self, level, data):
        d = np.ascontiguousarray(np.asarray(data, dtype=np.float64)).ravel()
        self._check(lib().gic_import_level(self._h, int(level), _ptr(d), d.size))

    def splice(self, new_seed):
        """Keep the flagged region's overdensity from a new real-space draw of
        new_seed and splice it into the old noise elsewhere; returns the
        conjugate-gradient iterations."""
        it = ctypes.c_int()
        self._check(lib().gic_splice(self._h, int(new_seed), ctypes.byref(it)))
        return it.value

    def apply_power_spectrum(self):
        self._check(lib().gic_apply_power_spectrum(self._h))

    def zeldovich(self):
        self._check(lib().gic_zeldovich(self._h))

    # --- fields ---------------------------------------------------------
    def field(self, which, level, fourier=False, dtype=np.float64):
        n = self.sizes[level]
        dt = DTYPE_F64 if np.dtype(dtype) == np.float64 else DTYPE_F32
        if fourier:
            out = np.empty(2 * n * n * (n // 2 + 1), dtype=dtype)
            self._check(lib().gic_copy_field(self._h, which, level, FOURIER_SPACE, _ptr(out), dt))
            cdt = np.complex128 if np.dtype(dtype) == np.float64 else np.complex64
            return out.view(cdt).reshape(n, n, n // 2 + 1)
        out = np.empty(n ** 3, dtype=dtype)
        self._check(lib().gic_copy_field(self._h, which, level, REAL_SPACE, _ptr(out), dt))
        return out.reshape(n, n, n)

    def set_field(self, level, data, fourier=False):
        data = np.ascontiguousarray(data)
        dt = DTYPE_F64 if data.dtype in (np.float64, np.complex128) else DTYPE_F32
        if np.iscomplexobj(data):
            data = data.view(np.float64 if dt == DTYPE_F64 else np.float32)
        self._check(lib().gic_set_field(self._h, FIELD_OVERDENSITY, level, FOURIER_SPACE if fourier else REAL_SPACE,
                                        _ptr(data), dt))

    def particles(self, level, cells=None, start=0, count=None):
        n = self.sizes[level]
        if cells is not None:
            cells = np.ascontiguousarray(cells, dtype=np.uint64)
            count = cells.size
        elif count is None:
            count = n ** 3 - start
        pos = np.empty((count, 3), dtype=np.float32)
        vel = np.empty((count, 3), dtype=np.float32)
        cp = _ptr(cells) if cells is not None else None
        self._check(lib().gic_particles(self._h, level, cp, start, count, _ptr(pos), _ptr(vel)))
        return pos, vel

    def set_lpt_order(self, order):
        """1: Zel'dovich (the reference), 2: plus second-order LPT."""
        self._check(lib().gic_set_lpt_order(self._h, int(order)))

    # --- constrained realisations (ic.hpp:1592-1631) ---------------------
    def set_flags(self, level, cells):
        c = np.ascontiguousarray(np.asarray(cells, dtype=np.uint64))
        self._check(lib().gic_set_flags(self._h, level, _ptr(c), c.size))

    def set_variance_filter_scale(self, scale):
        self._check(lib().gic_set_variance_filter_scale(self._h, float(scale)))

    def calculate(self, name):
        out = ctypes.c_double()
        self._check(lib().gic_calculate(self._h, name.encode(), ctypes.byref(out)))
        return out.value

    def modify(self, name, relative, target, init_steps=10, precision=0.001):
        self._check(lib().gic_add_modification(self._h, name.encode(), int(bool(relative)), float(target),
                                               int(init_steps), float(precision)))

    def apply_modifications(self):
        pre, post = ctypes.c_double(), ctypes.c_double()
        self._check(lib().gic_apply_modifications(self._h, ctypes.byref(pre), ctypes.byref(post)))
        return pre.value, post.value

    def clear_modifications(self):
        self._check(lib().gic_clear_modifications(self._h))

    def chi2(self):
        out, dof = ctypes.c_double(), ctypes.c_uint64()
        self._check(lib().gic_chi2(self._h, ctypes.byref(out), ctypes.byref(dof)))
        return out.value, dof.value

    def set_quadratic_line_search(self, on=True):
        self._check(lib().gic_set_quadratic_line_search(self._h, int(bool(on))))

    # --- differentiation: device pointers, dense n^3 arrays per level ---------
    @staticmethod
    def _ptrs(ptrs):
        ptrs = [int(p) for p in ptrs]
        return (ctypes.c_void_p * len(ptrs))(*ptrs)

    def displacement_forward(self, noise_ptrs, psi_ptrs):
        """psi[3l+a] <- component a of the Zel'dovich displacement of the white
        noise noise[l] (device pointers of the context's precision)."""
        self._check(lib().gic_displacement_forward(self._h, self._ptrs(noise_ptrs), self._ptrs(psi_ptrs)))

    def displacement_adjoint(self, psi_bar_ptrs, noise_bar_ptrs):
        """noise_bar[l] <- (d psi / d w_l)^T psi_bar."""
        self._check(lib().gic_displacement_adjoint(self._h, self._ptrs(psi_bar_ptrs), self._ptrs(noise_bar_ptrs)))

    def variance_gradient(self, noise_ptrs, grad_ptrs=None):
        """Filtered variance of the flagged region for the given white noise;
        with grad_ptrs its gradient 2 Q w is written there."""
        v = ctypes.c_double()
        g = self._ptrs(grad_ptrs) if grad_ptrs is not None else None
        self._check(lib().gic_variance_gradient(self._h, self._ptrs(noise_ptrs), ctypes.byref(v), g))
        return v.value

    def batch_constrained(self, seeds, cons, out_ptr, lanes=4, graph=True):
        """Batched constrained realisations of a one-level context: seeds[r]
        (real-space draws), cons[r, i] = (x, y, z, radius, target) top-hat
        mean-overdensity constraints with absolute targets; out_ptr (device,
        the context's precision) receives delta, psi_x, psi_y, psi_z per
        realisation as dense n^3 arrays at out + r*4*n^3.  Returns the elapsed
        milliseconds from the first recorded work to the end of the batch."""
        seeds = np.ascontiguousarray(seeds, dtype=np.uint64).ravel()
        cons = np.ascontiguousarray(cons, dtype=np.float64)
        if cons.ndim != 3 or cons.shape[2] != 5 or cons.shape[0] != seeds.size:
            raise ValueError("cons must have shape (len(seeds), ncons, 5)")
        ms = ctypes.c_double()
        self._check(lib().gic_batch_constrained(self._h, int(seeds.size), _ptr(seeds), int(cons.shape[1]), _ptr(cons),
                                                ctypes.c_void_p(int(out_ptr)), int(lanes), int(bool(graph)),
                                                ctypes.byref(ms)))
        return ms.value

    def mode_power(self, level):
        """Per-|k|^2 sums of |delta_k|^2, mode counts and theory covariance."""
        h = self.sizes[level] // 2
        nm = 3 * h * h + 1
        s, c, t = (np.empty(nm) for _ in range(3))
        self._check(lib().gic_mode_power(self._h, level, _ptr(s), _ptr(c), _ptr(t), nm))
        return s, c, t

    def particle_mass_eps(self, level):
        m, e = ctypes.c_double(), ctypes.c_double()
        self._check(lib().gic_particle_mass_eps(self._h, level, ctypes.byref(m), ctypes.byref(e)))
        return m.value, e.value

    def velocity_ratio(self):
        r = ctypes.c_double()
        self._check(lib().gic_velocity_ratio(self._h, ctypes.byref(r)))
        return r.value

    # --- raw streams and transforms ----------------------------------------
    def raw_stream(self, seed, n):
        out = np.empty(n, dtype=np.uint32)
        self._check(lib().gic_raw_stream(self._h, int(seed), _ptr(out), n))
        return out

    def gaussian_stream(self, seed, n):
        out = np.empty(n, dtype=np.float64)
        self._check(lib().gic_gaussian_stream(self._h, int(seed), _ptr(out), n))
        return out

    def gaussian_stream_noalloc(self, seed, device_ptr, n):
        """The first n Gaussians of the stream into a device array of n
        doubles, with the library's fixed workspace (no device allocation;
        n <= 1048576)."""
        self._check(lib().gic_gaussian_stream_noalloc(self._h, int(seed), ctypes.c_void_p(int(device_ptr)), int(n)))

    def fft_inplace(self, device_ptr, n, direction):
        self._check(lib().gic_fft_inplace(self._h, ctypes.c_void_p(device_ptr), int(n), int(direction)))

    # --- timing -------------------------------------------------------------
    def synchronize(self):
        self._check(lib().gic_synchronize(self._h))

    def set_event_timing(self, on=True):
        self._check(lib().gic_set_event_timing(self._h, 1 if on else 0))

    def event_timing(self, kernel):
        t = ctypes.c_double()
        c = ctypes.c_long()
        self._check(lib().gic_get_event_timing(self._h, kernel.encode(), ctypes.byref(t), ctypes.byref(c)))
        return t.value, c.value

    def event_timings(self, with_bytes=False):
        """{kernel: (total_ms, launches)} for every timed kernel; with_bytes
        adds the launches' summed algorithmic HBM bytes as a third entry."""
        need = ctypes.c_size_t()
        self._check(lib().gic_list_event_timing(self._h, None, 0, ctypes.byref(need)))
        buf = ctypes.create_string_buffer(need.value + 16)
        self._check(lib().gic_list_event_timing(self._h, buf, len(buf), ctypes.byref(need)))
        out = {}
        for ln in buf.value.decode().splitlines():
            name, ms, cnt, nbytes = ln.split("\t")
            out[name] = (float(ms), int(cnt), float(nbytes)) if with_bytes else (float(ms), int(cnt))
        return out

    def reset_event_timing(self):
        self._check(lib().gic_reset_event_timing(self._h))

    def device_bytes(self):
        b = ctypes.c_size_t()
        self._check(lib().gic_device_bytes(self._h, ctypes.byref(b)))
        return b.value
