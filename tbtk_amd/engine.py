"""Python binding of the tbtkx C ABI (include/tbtkx.h) via ctypes.

This is a thin host-side view of the HIP engine used by tests and bench.py;
the C++ drop-in classes (tbtk_amd/host) call the same ABI. There is no CPU
fallback: if the native library is missing or no GPU is present the calls
raise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libtbtkx.so")

ADVANCED, RETARDED, PRINCIPAL, NONPRINCIPAL = 0, 1, 2, 3
H_FP64, DOUBLING, VEC_FP64 = 1, 2, 4
DAMP_NONE, DAMP_LORENTZ, DAMP_JACKSON = 0, 1, 2
(OPT_PLAIN_KERNEL, OPT_GRAPH, OPT_BLOCK_WIDTH, OPT_XCD_REMAP, OPT_TILE_ROWS, OPT_NT_STREAM,
 OPT_ELL, OPT_AUTOTUNE, OPT_MFMA_PIPE, OPT_STRIP_ROWS, OPT_GRID_SHARE,
 OPT_STRIP_CHUNK, OPT_ROWS_IN_FLIGHT, OPT_REAL, OPT_COL16, OPT_PALETTE,
 OPT_MULTI, OPT_ROW_PATTERNS, OPT_SLOT_ROWS, OPT_PATTERN_BYTE_ROWS) = range(1, 21)
ANY_DEVICE = -1

# (name, argtypes) of every exported function, in header order.
_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I = ctypes.c_int
_D = ctypes.c_double
SIGNATURES = {
    "tbtkx_last_error": [],
    "tbtkx_device_count": [_P],
    "tbtkx_device_acquire": [_P],
    "tbtkx_device_release": [_I],
    "tbtkx_create": [_I, _P],
    "tbtkx_destroy": [_P],
    "tbtkx_set_csr": [_P, _I64, _I64, _P, _P, _P, _D],
    "tbtkx_moments": [_P, _I64, _I, _P, _I, _P],
    "tbtkx_moments_noalloc": [_P, _I64, _I, _P, _I, _P],
    "tbtkx_lut_upload": [_P, _I, _I, _P],
    "tbtkx_lut_generate": [_P, _I, _I, _D, _D, _D],
    "tbtkx_lut_destroy": [_P],
    "tbtkx_contract": [_P, _I, _P, _I, _P],
    "tbtkx_contract_direct": [_P, _I, _I, _P, _I, _I, _D, _D, _D, _P],
    "tbtkx_coeffs_upload": [_P, _I, _I, _P],
    "tbtkx_contract_mfma": [_P, _I, _P],
    "tbtkx_set_damping": [_P, _I, _D],
    "tbtkx_moments_block": [_P, _I, _P, _I, _I, _P],
    "tbtkx_moments_multi": [_P, _I, _P, _P, _P, _I, _P],
    "tbtkx_moments_trace": [_P, _I, ctypes.c_uint64, _I, _I, _P],
    "tbtkx_set_lattice": [_P, _P, _P, _D],
    "tbtkx_lattice_gershgorin": [_P, _P, _P, _P],
    "tbtkx_csr_size": [_P, _P, _P],
    "tbtkx_csr_copy": [_P, _P, _P, _P],
    "tbtkx_multi_set": [_P, _I, _P, _P, _P, _P, _P, _P],
    "tbtkx_multi_set_lattices": [_P, _I, _P, _P, _P],
    "tbtkx_multi_moments": [_P, _P, _P, _P, _I, _P],
    "tbtkx_last_timing": [_P, _P, _P],
    "tbtkx_set_option": [_P, _I, _I64],
    "tbtkx_get_option": [_P, _I, _P],
}

# tbtkx_lattice_term and tbtkx_lattice (include/tbtkx.h).
LATTICE_TERM = np.dtype([("d", "<i4", (3,)), ("to", "<i4"), ("frm", "<i4"),
                         ("re", "<f8"), ("im", "<f8")], align=True)


class _Lattice(ctypes.Structure):
    _fields_ = [("cells", ctypes.c_int32 * 3), ("periodic", ctypes.c_int32 * 3),
                ("orbitals", ctypes.c_int32), ("num_terms", ctypes.c_int32),
                ("terms", ctypes.c_void_p)]


def _lattice_struct(spec):
    """(ctypes struct, arrays to keep alive) for a lattice.LatticeSpec."""
    terms = np.ascontiguousarray(spec.terms, dtype=LATTICE_TERM)
    s = _Lattice()
    for a in range(3):
        s.cells[a] = int(spec.cells[a])
        s.periodic[a] = int(bool(spec.periodic[a]))
    s.orbitals = int(spec.orbitals)
    s.num_terms = int(terms.size)
    s.terms = terms.ctypes.data if terms.size else None
    return s, terms


_lib = None


class EngineError(RuntimeError):
    pass


def load():
    """Load libtbtkx.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise EngineError(
                f"native engine {LIB_PATH} is missing: run `python -m tbtk_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_char_p if name == "tbtkx_last_error" else ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _check(rc):
    if rc != 0:
        raise EngineError(f"tbtkx error {rc}: {load().tbtkx_last_error().decode()}")


def device_count() -> int:
    n = ctypes.c_int(0)
    _check(load().tbtkx_device_count(ctypes.byref(n)))
    return n.value


class Engine:
    """One device handle: a HIP stream plus the resident Hamiltonian."""

    def __init__(self, device: int = 0):
        self._lib = load()
        h = ctypes.c_void_p()
        _check(self._lib.tbtkx_create(device, ctypes.byref(h)))
        self._h = h
        self.N = 0
        self.lut_shape = None

    def close(self):
        if self._h:
            self._lib.tbtkx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------- model --
    def set_csr(self, csr, scale: float):
        rp, col, val = csr
        rp = np.ascontiguousarray(rp, dtype=np.uint32)
        col = np.ascontiguousarray(col, dtype=np.uint32)
        val = np.ascontiguousarray(val, dtype=np.complex128)
        _check(self._lib.tbtkx_set_csr(self._h, rp.size - 1, col.size, _p(rp),
                                       _p(col), _p(val), float(scale)))
        self.N = rp.size - 1

    def set_option(self, option: int, value: int):
        _check(self._lib.tbtkx_set_option(self._h, option, int(value)))

    def get_option(self, option: int) -> int:
        v = ctypes.c_int64(0)
        _check(self._lib.tbtkx_get_option(self._h, option, ctypes.byref(v)))
        return v.value

    # ------------------------------------------------------------ plain ---
    def moments(self, frm: int, to, M: int) -> np.ndarray:
        to = np.ascontiguousarray(np.atleast_1d(to), dtype=np.int64)
        out = np.zeros((to.size, M), dtype=np.complex128)
        _check(self._lib.tbtkx_moments(self._h, int(frm), to.size, _p(to), M, _p(out)))
        return out

    def moments_noalloc(self, frm: int, to, M: int) -> np.ndarray:
        """tbtkx_moments_noalloc: tbtkx_moments on the fixed per-GPU scratch."""
        to = np.ascontiguousarray(np.atleast_1d(to), dtype=np.int64)
        out = np.zeros((to.size, M), dtype=np.complex128)
        _check(self._lib.tbtkx_moments_noalloc(self._h, int(frm), to.size, _p(to), M, _p(out)))
        return out

    # ---------------------------------------------------------- contraction
    def lut_upload(self, table: np.ndarray):
        table = np.ascontiguousarray(table, dtype=np.complex128)
        _check(self._lib.tbtkx_lut_upload(self._h, table.shape[0], table.shape[1], _p(table)))
        self.lut_shape = table.shape

    def lut_generate(self, M, E, lo, hi, scale):
        _check(self._lib.tbtkx_lut_generate(self._h, M, E, lo, hi, scale))
        self.lut_shape = (M, E)

    def lut_destroy(self):
        _check(self._lib.tbtkx_lut_destroy(self._h))
        self.lut_shape = None

    def contract(self, coeffs, gtype: int) -> np.ndarray:
        c = np.ascontiguousarray(np.atleast_2d(coeffs), dtype=np.complex128)
        # Without a table the C call reports the error; E = 1 keeps the
        # output pointer valid for it.
        E = self.lut_shape[1] if self.lut_shape else 1
        out = np.zeros((c.shape[0], E), dtype=np.complex128)
        _check(self._lib.tbtkx_contract(self._h, c.shape[0], _p(c), gtype, _p(out)))
        return out

    def contract_direct(self, coeffs, gtype: int, E: int, e0: float, de: float,
                        scale: float) -> np.ndarray:
        c = np.ascontiguousarray(np.atleast_2d(coeffs), dtype=np.complex128)
        out = np.zeros((c.shape[0], E), dtype=np.complex128)
        _check(self._lib.tbtkx_contract_direct(self._h, c.shape[0], c.shape[1], _p(c), gtype,
                                               E, e0, de, scale, _p(out)))
        return out

    def coeffs_upload(self, coeffs):
        c = np.ascontiguousarray(np.atleast_2d(coeffs), dtype=np.complex128)
        _check(self._lib.tbtkx_coeffs_upload(self._h, c.shape[0], c.shape[1], _p(c)))
        self.coeffs_shape = c.shape

    def set_damping(self, kind: int, broadening: float = 0.0):
        """Damping applied on the device (FP64) to the coefficients of every
        later contraction call: DAMP_NONE, DAMP_LORENTZ (with broadening)
        or DAMP_JACKSON."""
        _check(self._lib.tbtkx_set_damping(self._h, int(kind), float(broadening)))

    def contract_mfma(self, gtype: int, keep_on_device: bool = False):
        """Split-BF16 MFMA contraction of the uploaded coefficients."""
        if keep_on_device:
            _check(self._lib.tbtkx_contract_mfma(self._h, gtype, None))
            return None
        out = np.zeros((self.coeffs_shape[0], self.lut_shape[1]), dtype=np.complex128)
        _check(self._lib.tbtkx_contract_mfma(self._h, gtype, _p(out)))
        return out

    # ---------------------------------------------------------- batched ---
    def moments_block(self, sites, M: int, flags: int = 0) -> np.ndarray:
        sites = np.ascontiguousarray(np.atleast_1d(sites), dtype=np.int64)
        out = np.zeros((sites.size, M), dtype=np.complex128)
        _check(self._lib.tbtkx_moments_block(self._h, sites.size, _p(sites), M, flags, _p(out)))
        return out

    def moments_multi(self, froms, tos, M: int) -> list:
        """Moments <to|T_n|from> for every from-index and its to-list
        (tbtkx_moments_multi): a list of (len(to_j), M) arrays."""
        froms = np.ascontiguousarray(np.atleast_1d(froms), dtype=np.int64)
        offs = np.zeros(froms.size + 1, dtype=np.int32)
        offs[1:] = np.cumsum([len(t) for t in tos])
        flat = np.ascontiguousarray(np.concatenate([np.asarray(t, dtype=np.int64) for t in tos])
                                    if len(tos) else np.zeros(0, np.int64), dtype=np.int64)
        out = np.zeros((max(int(offs[-1]), 1), M), dtype=np.complex128)
        _check(self._lib.tbtkx_moments_multi(self._h, froms.size, _p(froms), _p(offs), _p(flat),
                                             M, _p(out)))
        return [out[offs[j]:offs[j + 1]] for j in range(froms.size)]

    def moments_trace(self, R: int, seed: int, M: int, flags: int = 0) -> np.ndarray:
        out = np.zeros(M, dtype=np.complex128)
        _check(self._lib.tbtkx_moments_trace(self._h, R, seed, M, flags, _p(out)))
        return out

    def set_lattice(self, spec, scale: float):
        """Build the Hamiltonian of a lattice.LatticeSpec on the device."""
        lat, keep = _lattice_struct(spec)
        onsite = None
        if spec.onsite is not None:
            onsite = np.ascontiguousarray(spec.onsite, dtype=np.complex128)
        _check(self._lib.tbtkx_set_lattice(self._h, ctypes.byref(lat),
                                           _p(onsite) if onsite is not None else None,
                                           float(scale)))
        self.N = spec.basis_size
        return keep

    def lattice_gershgorin(self, spec) -> float:
        """max_r sum_c |H_rc| of a lattice.LatticeSpec, computed on the device."""
        lat, keep = _lattice_struct(spec)
        onsite = None
        if spec.onsite is not None:
            onsite = np.ascontiguousarray(spec.onsite, dtype=np.complex128)
        out = ctypes.c_double(0)
        _check(self._lib.tbtkx_lattice_gershgorin(
            self._h, ctypes.byref(lat), _p(onsite) if onsite is not None else None,
            ctypes.byref(out)))
        return out.value

    def csr_size(self):
        """(rows, non-zeros) of the resident model."""
        n = ctypes.c_int64(0)
        z = ctypes.c_int64(0)
        _check(self._lib.tbtkx_csr_size(self._h, ctypes.byref(n), ctypes.byref(z)))
        return n.value, z.value

    def csr_copy(self):
        """(rowptr, col, H/s) of the resident model, as the device holds them."""
        n = ctypes.c_int64(0)
        z = ctypes.c_int64(0)
        _check(self._lib.tbtkx_csr_size(self._h, ctypes.byref(n), ctypes.byref(z)))
        rp = np.zeros(n.value + 1, dtype=np.uint32)
        col = np.zeros(max(z.value, 1), dtype=np.uint32)
        val = np.zeros(max(z.value, 1), dtype=np.complex128)
        _check(self._lib.tbtkx_csr_copy(self._h, _p(rp), _p(col), _p(val)))
        return rp, col[:z.value], val[:z.value]

    # ----------------------------------------------------- uneven batch ---
    def multi_set_lattices(self, specs, scales):
        """Upload several lattice models, each built on the device."""
        K = len(specs)
        lats = (_Lattice * K)()
        keep = []
        ptrs = (ctypes.c_void_p * K)()
        any_onsite = False
        for m, spec in enumerate(specs):
            s, terms = _lattice_struct(spec)
            lats[m] = s
            keep.append(terms)
            if spec.onsite is not None:
                o = np.ascontiguousarray(spec.onsite, dtype=np.complex128)
                keep.append(o)
                ptrs[m] = o.ctypes.data
                any_onsite = True
        sc = np.ascontiguousarray(scales, dtype=np.float64)
        _check(self._lib.tbtkx_multi_set_lattices(
            self._h, K, ctypes.cast(lats, ctypes.c_void_p),
            ctypes.cast(ptrs, ctypes.c_void_p) if any_onsite else None, _p(sc)))
        self.multi_N = np.array([sp.basis_size for sp in specs], dtype=np.int64)

    def multi_set(self, csrs, scales):
        """Upload several models at once: csrs = [(rowptr, col, val), ...]."""
        rps = [np.ascontiguousarray(c[0], dtype=np.uint32) for c in csrs]
        cols = [np.ascontiguousarray(c[1], dtype=np.uint32) for c in csrs]
        vals = [np.ascontiguousarray(c[2], dtype=np.complex128) for c in csrs]
        N = np.array([r.size - 1 for r in rps], dtype=np.int64)
        Z = np.array([c.size for c in cols], dtype=np.int64)
        rp = np.concatenate(rps)
        col = np.concatenate(cols) if Z.sum() else np.zeros(1, np.uint32)
        val = np.concatenate(vals) if Z.sum() else np.zeros(1, np.complex128)
        sc = np.ascontiguousarray(scales, dtype=np.float64)
        _check(self._lib.tbtkx_multi_set(self._h, len(csrs), _p(N), _p(Z), _p(rp), _p(col),
                                         _p(val), _p(sc)))
        self.multi_N = N

    def multi_moments(self, sites, moments, flags: int = 0):
        """sites: list of per-model site arrays; moments: per-model M.
        Returns a list of (B_m, M_m) complex arrays."""
        ns = np.array([len(s) for s in sites], dtype=np.int32)
        st = np.ascontiguousarray(np.concatenate([np.asarray(s) for s in sites]), dtype=np.int64)
        Ms = np.ascontiguousarray(moments, dtype=np.int32)
        total = int((ns.astype(np.int64) * Ms).sum())
        out = np.zeros(total, dtype=np.complex128)
        _check(self._lib.tbtkx_multi_moments(self._h, _p(ns), _p(st), _p(Ms), flags, _p(out)))
        res, o = [], 0
        for b, m in zip(ns, Ms):
            res.append(out[o:o + b * m].reshape(b, m))
            o += b * m
        return res

    def last_timing(self):
        ms = ctypes.c_double(0)
        n = ctypes.c_int64(0)
        _check(self._lib.tbtkx_last_timing(self._h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value
