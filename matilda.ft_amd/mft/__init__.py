"""Python binding of libmft.so (the C ABI declared in include/mft.h).

The engine is native: every computation below happens in hand-written gfx950
HIP kernels inside libmft.so.  This module only moves host arrays and calls the
C entry points; there is no CPU fallback -- if the library or a GPU is missing
the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libmft.so")
HEADER = os.path.join(os.path.dirname(os.path.dirname(_HERE)), "include", "mft.h")

_lib = None

_p = C.c_void_p
_i = C.c_int
_d = C.c_double
_f = C.c_float

MFT_OK, MFT_ERR_INVALID, MFT_ERR_OOM, MFT_ERR_NONFINITE, MFT_ERR_DEVICE, MFT_ERR_UNSUPPORTED = range(6)
POT = {"gaussian": 0, "erf2": 1, "erfG": 2, "charges": 3}
INTEGRATOR = {"GJF": 0, "EM": 1, "none": 2}
TK_NAMES = ["bin", "scan", "scatter", "density", "fft_x_fwd", "fft_y_fwd", "conv_z", "fft_y_inv",
            "fft_x_inv", "gather", "integrate"]
FTS_POT = {"Helfand": 0, "Flory": 1, "Incompress": 2}


class MftError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mft error {code}: {msg}")
        self.code = code


def lib():
    """Load libmft.so (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} missing: build with `make -C matilda.ft_amd`")
        _lib = C.CDLL(LIB_PATH)
        _lib.mft_last_error.restype = C.c_char_p
        for name in ("mft_tild_stream", "mft_fts_stream"):
            if hasattr(_lib, name):
                getattr(_lib, name).restype = _p
    return _lib


def check(rc):
    if rc != MFT_OK:
        raise MftError(rc, lib().mft_last_error().decode())
    return rc


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_p)


def device_count():
    n = _i(0)
    check(lib().mft_device_count(C.byref(n)))
    return n.value


# ----------------------------------------------------------- buffers -----
class DeviceBuffer:
    """Raw device allocation through mft_malloc (no framework needed)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = _p()
        check(lib().mft_malloc(C.byref(self.ptr), C.c_size_t(self.nbytes)))

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        check(lib().mft_memcpy_h2d(b.ptr, _ptr(a), C.c_size_t(a.nbytes)))
        return b

    def to_array(self, dtype, shape):
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        check(lib().mft_synchronize())
        check(lib().mft_memcpy_d2h(_ptr(out), self.ptr, C.c_size_t(out.nbytes)))
        return out

    def zero(self):
        check(lib().mft_memset(self.ptr, 0, C.c_size_t(self.nbytes)))

    def free(self):
        if self.ptr:
            lib().mft_free(self.ptr)
            self.ptr = _p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    check(lib().mft_synchronize())


# --------------------------------------------------------------- RNG -----
def philox_raw(nblocks, seed, c1, c2, c3):
    b = DeviceBuffer(16 * nblocks)
    check(lib().mft_philox_raw(b.ptr, C.c_int64(nblocks), C.c_uint64(seed), C.c_uint32(c1), C.c_uint32(c2),
                               C.c_uint32(c3), None))
    return b.to_array(np.uint32, (nblocks, 4))


def philox_normal(nblocks, seed, c1, c2, c3):
    b = DeviceBuffer(16 * nblocks)
    check(lib().mft_philox_normal(b.ptr, C.c_int64(nblocks), C.c_uint64(seed), C.c_uint32(c1), C.c_uint32(c2),
                                  C.c_uint32(c3), None))
    return b.to_array(np.float32, (nblocks, 4))


# --------------------------------------------------------------- FFT -----
def fft3d(a, direction):
    """Box-convention transform of a complex array shaped (batch?, nz, ny, nx):
    direction 1 forward scaled by 1/M, -1 inverse unscaled (float32 math)."""
    a = np.ascontiguousarray(a, np.complex64)
    shp = a.shape
    nz, ny, nx = shp[-3:]
    batch = int(np.prod(shp[:-3])) if len(shp) > 3 else 1
    plan = _p()
    check(lib().mft_fft_plan_create(C.byref(plan), nx, ny, nz, batch))
    try:
        buf = DeviceBuffer.from_array(a)
        check(lib().mft_fft_exec(plan, buf.ptr, buf.ptr, direction, None))
        return buf.to_array(np.complex64, shp)
    finally:
        lib().mft_fft_plan_destroy(plan)


def fft_supported(n):
    return bool(lib().mft_fft_supported(n))


# -------------------------------------------------------------- TILD -----
class TildPotentialC(C.Structure):
    _fields_ = [("kind", _i), ("maskI", C.c_uint32), ("maskJ", C.c_uint32), ("A", _f), ("sigma", _f),
                ("Rp", _f), ("ramp", _i), ("A_final", _f)]


class TildDescC(C.Structure):
    _fields_ = [("nx", _i), ("ny", _i), ("nz", _i), ("lx", _f), ("ly", _f), ("lz", _f), ("pmeorder", _i),
                ("nparticles", _i), ("nspecies", _i), ("maxbonds", _i), ("nbondtypes", _i),
                ("bond_k", _p), ("bond_r0", _p), ("bond_style", _p), ("npotentials", _i),
                ("potentials", _p), ("integrator", _i), ("delt", _f), ("mass", _p), ("mobility", _p),
                ("seed", C.c_uint64), ("has_charges", _i)]


def grid_weights(x, Nx, dx, pmeorder):
    """mft_tild_grid_weights on host data; returns (inds, W)."""
    x = np.ascontiguousarray(x, np.float32).reshape(-1)
    ns = x.size // 3
    G = (pmeorder + 1) ** 3
    dxb = DeviceBuffer.from_array(x)
    W = DeviceBuffer(4 * ns * G)
    I = DeviceBuffer(4 * ns * G)
    check(lib().mft_tild_grid_weights(W.ptr, I.ptr, dxb.ptr, ns, _ptr(np.asarray(Nx, np.int32)),
                                      _ptr(np.asarray(dx, np.float32)), pmeorder, None))
    return I.to_array(np.int32, ns * G), W.to_array(np.float32, ns * G)


def fill_density(M, sites, inds, W, G, q=None):
    sites = np.ascontiguousarray(sites, np.int32)
    rho = DeviceBuffer(4 * M)
    rho.zero()
    rhoq = None
    qb = None
    if q is not None:
        rhoq = DeviceBuffer(4 * M)
        rhoq.zero()
        qb = DeviceBuffer.from_array(np.asarray(q, np.float32))
    sb = DeviceBuffer.from_array(sites)
    ib = DeviceBuffer.from_array(np.asarray(inds, np.int32))
    wb = DeviceBuffer.from_array(np.asarray(W, np.float32))
    ns = np.asarray(inds).size // G
    check(lib().mft_tild_fill_density(rho.ptr, rhoq.ptr if rhoq else None, qb.ptr if qb else None, sb.ptr,
                                      sites.size, ns, ib.ptr, wb.ptr, G, M, None))
    return rho.to_array(np.float32, M), (rhoq.to_array(np.float32, M) if rhoq else None)


def map_forces(ns, gridF, sites, inds, W, G, gvol, q=None):
    sites = np.ascontiguousarray(sites, np.int32)
    f = DeviceBuffer(12 * ns)
    f.zero()
    gb = DeviceBuffer.from_array(np.asarray(gridF, np.float32))
    sb = DeviceBuffer.from_array(sites)
    ib = DeviceBuffer.from_array(np.asarray(inds, np.int32))
    wb = DeviceBuffer.from_array(np.asarray(W, np.float32))
    qb = DeviceBuffer.from_array(np.asarray(q, np.float32)) if q is not None else None
    M = np.asarray(gridF).size // 3
    check(lib().mft_tild_map_forces(f.ptr, gb.ptr, qb.ptr if qb else None, sb.ptr, sites.size, ns, ib.ptr, wb.ptr, G,
                                    _f(gvol), M, None))
    return f.to_array(np.float32, 3 * ns)


def bonds(x, topo, bond_k, bond_r0, bond_style, L, maxbonds=2):
    x = np.ascontiguousarray(x, np.float32)
    ns = x.size // 3
    f = DeviceBuffer(12 * ns)
    f.zero()
    xb = DeviceBuffer.from_array(x)
    nb = DeviceBuffer.from_array(np.asarray(topo["nBonds"], np.int32))
    bt = DeviceBuffer.from_array(np.asarray(topo["bondedTo"], np.int32))
    ty = DeviceBuffer.from_array(np.asarray(topo["bondType"], np.int32))
    k = np.asarray(bond_k, np.float32)
    r0 = np.asarray(bond_r0, np.float32)
    st = np.asarray(bond_style, np.int32)
    check(lib().mft_tild_bonds(f.ptr, nb.ptr, bt.ptr, ty.ptr, _ptr(r0), _ptr(k), _ptr(st), k.size, xb.ptr,
                               _ptr(np.asarray(L, np.float32)), ns, maxbonds, None))
    return f.to_array(np.float32, 3 * ns)


class TildEngine:
    """Device-resident TILD system (mft_tild_*)."""

    def __init__(self, Nx, L, pmeorder, nspecies, nparticles, potentials, bond_k=(), bond_r0=(),
                 bond_style=(), integrator="GJF", delt=0.005, mass=None, mobility=None, seed=1234,
                 has_charges=False, maxbonds=2):
        self._keep = []
        pots = (TildPotentialC * max(1, len(potentials)))()
        for i, p in enumerate(potentials):
            pots[i] = TildPotentialC(POT[p["kind"]], p.get("maskI", 0), p.get("maskJ", 0), p["A"],
                                     p.get("sigma", 0.0), p.get("Rp", 0.0), int(p.get("ramp", False)),
                                     p.get("A_final", 0.0))
        bk = np.asarray(bond_k, np.float32)
        br = np.asarray(bond_r0, np.float32)
        bs = np.asarray(bond_style, np.int32)
        ma = np.asarray(mass if mass is not None else [1.0] * nspecies, np.float32)
        mo = np.asarray(mobility if mobility is not None else [1.0] * nspecies, np.float32)
        self._keep += [pots, bk, br, bs, ma, mo]
        d = TildDescC(Nx[0], Nx[1], Nx[2], L[0], L[1], L[2], pmeorder, nparticles, nspecies, maxbonds,
                      bk.size, _ptr(bk), _ptr(br), _ptr(bs), len(potentials), C.cast(pots, _p),
                      INTEGRATOR[integrator], delt, _ptr(ma), _ptr(mo), seed, int(has_charges))
        self.h = _p()
        check(lib().mft_tild_create(C.byref(self.h), C.byref(d)))
        self.N = nparticles
        self.M = int(np.prod(Nx))
        self.nspecies = nspecies
        self.npot = len(potentials)
        self.maxbonds = maxbonds

    def close(self):
        if self.h:
            lib().mft_tild_destroy(self.h)
            self.h = _p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_topology(self, type_, charge=None, nBonds=None, bondedTo=None, bondType=None):
        t = np.ascontiguousarray(type_, np.int32)
        q = np.ascontiguousarray(charge, np.float32) if charge is not None else None
        nb = np.ascontiguousarray(nBonds, np.int32) if nBonds is not None else None
        bt = np.ascontiguousarray(bondedTo, np.int32) if bondedTo is not None else None
        ty = np.ascontiguousarray(bondType, np.int32) if bondType is not None else None
        check(lib().mft_tild_set_topology(self.h, _ptr(t), _ptr(q), _ptr(nb), _ptr(bt), _ptr(ty)))

    def set_state(self, x, xo=None, noise_old=None):
        x = np.ascontiguousarray(x, np.float32)
        xo = np.ascontiguousarray(xo, np.float32) if xo is not None else None
        no = np.ascontiguousarray(noise_old, np.float32) if noise_old is not None else None
        check(lib().mft_tild_set_state(self.h, _ptr(x), _ptr(xo), _ptr(no)))

    def get_state(self):
        x = np.empty(3 * self.N, np.float32)
        xo = np.empty(3 * self.N, np.float32)
        no = np.empty(3 * self.N, np.float32)
        check(lib().mft_tild_get_state(self.h, _ptr(x), _ptr(xo), _ptr(no)))
        return x, xo, no

    def compute_forces(self):
        f = np.empty(3 * self.N, np.float32)
        check(lib().mft_tild_compute_forces(self.h, _ptr(f)))
        return f

    def densities(self, charge=False):
        rho = np.empty(self.nspecies * self.M, np.float32)
        rhoq = np.empty(self.M, np.float32) if charge else None
        check(lib().mft_tild_get_densities(self.h, _ptr(rho), _ptr(rhoq)))
        return rho.reshape(self.nspecies, self.M), rhoq

    def grid_force(self, s):
        F = np.empty(3 * self.M, np.float32)
        check(lib().mft_tild_get_grid_force(self.h, s, _ptr(F)))
        return F

    def step(self, n=1):
        check(lib().mft_tild_step(self.h, n))

    def synchronize(self):
        check(lib().mft_tild_synchronize(self.h))

    def energies(self):
        pe = np.zeros(max(1, self.npot), np.float64)
        be = _d(0)
        check(lib().mft_tild_energies(self.h, _ptr(pe), C.byref(be)))
        return pe[: self.npot], be.value

    def timing(self, enable):
        ms = np.zeros(len(TK_NAMES), np.float64)
        n = np.zeros(len(TK_NAMES), np.int64)
        check(lib().mft_tild_timing(self.h, int(enable), _ptr(ms), _ptr(n)))
        return dict(zip(TK_NAMES, ms)), dict(zip(TK_NAMES, n))

    def kernel_bytes(self):
        b = np.zeros(len(TK_NAMES), np.float64)
        check(lib().mft_tild_kernel_bytes(self.h, _ptr(b)))
        return dict(zip(TK_NAMES, b))
