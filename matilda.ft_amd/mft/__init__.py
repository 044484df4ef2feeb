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
POT = {"gaussian": 0, "erf2": 1, "erfG": 2, "charges": 3, "bias": 4, "wall": 5, "langevin": 6, "dpd": 7}
BIAS_PHASE = {"L": 0, "H": 1, "G": 2}
INTEGRATOR = {"GJF": 0, "EM": 1, "none": 2, "VV": 3}
TK_NAMES = ["bin", "scan", "scatter", "density", "fft_x_fwd", "fft_y_fwd", "conv_z", "fft_y_inv",
            "fft_x_inv", "gather", "integrate"]
FTS_POT = {"Helfand": 0, "Flory": 1, "Incompress": 2, "Edwards": 3}


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
        _lib.mft_server_error.restype = C.c_char_p
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


def mem_info():
    """(free, total) device memory in bytes."""
    f, t = C.c_size_t(), C.c_size_t()
    check(lib().mft_mem_info(C.byref(f), C.byref(t)))
    return f.value, t.value


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
                ("Rp", _f), ("ramp", _i), ("A_final", _f), ("phase", _i), ("dir", _i), ("n_periods", _i),
                ("wall_pos", _f), ("wall_sign", _i)]


class TildDescC(C.Structure):
    _fields_ = [("nx", _i), ("ny", _i), ("nz", _i), ("lx", _f), ("ly", _f), ("lz", _f), ("pmeorder", _i),
                ("nparticles", _i), ("nspecies", _i), ("maxbonds", _i), ("nbondtypes", _i),
                ("bond_k", _p), ("bond_r0", _p), ("bond_style", _p), ("npotentials", _i),
                ("potentials", _p), ("integrator", _i), ("delt", _f), ("mass", _p), ("mobility", _p),
                ("seed", C.c_uint64), ("has_charges", _i), ("maxangles", _i), ("nangletypes", _i),
                ("angle_k", _p), ("angle_theta", _p), ("angle_style", _p)]


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


def _i3(v):
    return _ptr(np.asarray(v, np.int32))


def _f3(v):
    return _ptr(np.asarray(v, np.float32))


def kspace_table(kind, A, sigma, Rp, Nx, L):
    """mft_tild_kspace_table: (u(k) [M], f(k) [M, 3]) complex64, x fastest."""
    M = int(np.prod(Nx))
    u = DeviceBuffer(8 * M)
    f = DeviceBuffer(24 * M)
    check(lib().mft_tild_kspace_table(u.ptr, f.ptr, POT[kind], _f(A), _f(sigma), _f(Rp), _i3(Nx), _f3(L), None))
    return u.to_array(np.complex64, M), f.to_array(np.complex64, (M, 3))


def pair_force(gridF, rho, fk, Nx, factor=1.0):
    """mft_tild_pair_force: gridF [M*3] + factor Re IFFT[fk_j FFT[rho]]."""
    M = int(np.prod(Nx))
    g = DeviceBuffer.from_array(np.asarray(gridF, np.float32))
    r = DeviceBuffer.from_array(np.asarray(rho, np.float32))
    fb = DeviceBuffer.from_array(np.asarray(fk, np.complex64))
    work = DeviceBuffer(16 * M)
    check(lib().mft_tild_pair_force(g.ptr, r.ptr, fb.ptr, _i3(Nx), _f(factor), work.ptr, None))
    return g.to_array(np.float32, 3 * M)


def potential_energy(rhoI, rhoJ, uk, Nx, gvol):
    """mft_tild_potential_energy (FP64 sum)."""
    M = int(np.prod(Nx))
    a = DeviceBuffer.from_array(np.asarray(rhoI, np.float32))
    b = DeviceBuffer.from_array(np.asarray(rhoJ, np.float32))
    u = DeviceBuffer.from_array(np.asarray(uk, np.complex64))
    work = DeviceBuffer(8 * M + 8 * ((M + 255) // 256))
    e = _d(0.0)
    check(lib().mft_tild_potential_energy(C.byref(e), a.ptr, b.ptr, u.ptr, _i3(Nx), _d(gvol), work.ptr, None))
    return e.value


def gjf_step(x, xo, noise_old, f, types, mass, mobility, delt, L, seed, step):
    """mft_tild_gjf_step on host arrays; returns (x, xo, noise_old)."""
    n = np.asarray(types).size
    bufs = [DeviceBuffer.from_array(np.asarray(a, np.float32)) for a in (x, xo, noise_old, f)]
    tb = DeviceBuffer.from_array(np.asarray(types, np.int32))
    m = np.asarray(mass, np.float32)
    D = np.asarray(mobility, np.float32)
    check(lib().mft_tild_gjf_step(bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr, tb.ptr, n, m.size, _ptr(m),
                                  _ptr(D), _f(delt), _f3(L), C.c_uint64(seed), C.c_int64(step), None))
    return tuple(b.to_array(np.float32, 3 * n) for b in bufs[:3])


def em_step(x, f, delt, L, seed, step):
    """mft_tild_em_step on host arrays; returns x."""
    x = np.asarray(x, np.float32)
    n = x.size // 3
    xb = DeviceBuffer.from_array(x)
    fb = DeviceBuffer.from_array(np.asarray(f, np.float32))
    check(lib().mft_tild_em_step(xb.ptr, fb.ptr, n, _f(delt), _f3(L), C.c_uint64(seed), C.c_int64(step), None))
    return xb.to_array(np.float32, 3 * n)


def tild_desc(Nx, L, pmeorder, nspecies, nparticles, potentials, bond_k=(), bond_r0=(), bond_style=(),
              integrator="GJF", delt=0.005, mass=None, mobility=None, seed=1234, has_charges=False, maxbonds=2,
              angle_k=(), angle_theta=(), angle_style=(), maxangles=0):
    """mft_tild_desc and the arrays it points to (keep them alive)."""
    pots = (TildPotentialC * max(1, len(potentials)))()
    for i, p in enumerate(potentials):
        pots[i] = TildPotentialC(POT[p["kind"]], p.get("maskI", 0), p.get("maskJ", 0), p["A"],
                                 p.get("sigma", 0.0), p.get("Rp", 0.0), int(p.get("ramp", False)),
                                 p.get("A_final", 0.0), BIAS_PHASE.get(p.get("phase", "L"), -1), p.get("dir", 0),
                                 p.get("n_periods", 1), p.get("wall_pos", 0.0), p.get("wall_sign", 1))
    bk = np.asarray(bond_k, np.float32)
    br = np.asarray(bond_r0, np.float32)
    bs = np.asarray(bond_style, np.int32)
    ma = np.asarray(mass if mass is not None else [1.0] * nspecies, np.float32)
    mo = np.asarray(mobility if mobility is not None else [1.0] * nspecies, np.float32)
    d = TildDescC(Nx[0], Nx[1], Nx[2], L[0], L[1], L[2], pmeorder, nparticles, nspecies, maxbonds,
                  bk.size, _ptr(bk), _ptr(br), _ptr(bs), len(potentials), C.cast(pots, _p),
                  INTEGRATOR[integrator], delt, _ptr(ma), _ptr(mo), seed, int(has_charges))
    ak = np.asarray(angle_k, np.float32)
    at = np.asarray(angle_theta, np.float32)
    ast = np.asarray(angle_style, np.int32)
    d.maxangles = maxangles
    d.nangletypes = ak.size
    d.angle_k, d.angle_theta, d.angle_style = _ptr(ak), _ptr(at), _ptr(ast)
    return d, [pots, bk, br, bs, ma, mo, ak, at, ast]


class TildEngine:
    """Device-resident TILD system (mft_tild_*)."""

    def __init__(self, Nx, L, pmeorder, nspecies, nparticles, potentials, bond_k=(), bond_r0=(),
                 bond_style=(), integrator="GJF", delt=0.005, mass=None, mobility=None, seed=1234,
                 has_charges=False, maxbonds=2, angle_k=(), angle_theta=(), angle_style=(), maxangles=0):
        d, self._keep = tild_desc(Nx, L, pmeorder, nspecies, nparticles, potentials, bond_k, bond_r0, bond_style,
                                  integrator, delt, mass, mobility, seed, has_charges, maxbonds, angle_k,
                                  angle_theta, angle_style, maxangles)
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

    def set_angles(self, nAngles, angleType, angleGroup):
        """Per-particle angle lists (before set_state): nAngles[N], angleType[N, maxangles],
        angleGroup[N, maxangles, 3] particle ids i, j (vertex), k."""
        na = np.ascontiguousarray(nAngles, np.int32)
        at = np.ascontiguousarray(angleType, np.int32)
        ag = np.ascontiguousarray(angleGroup, np.int32)
        check(lib().mft_tild_set_angles(self.h, _ptr(na), _ptr(at), _ptr(ag)))

    def angle_energy(self):
        e = _d(0)
        check(lib().mft_tild_angle_energy(self.h, C.byref(e)))
        return e.value

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

    def set_run(self, max_steps):
        """Declare a run of max_steps steps: potential ramps follow the step
        index within the run across step() calls."""
        check(lib().mft_tild_set_run(self.h, C.c_int64(max_steps)))

    def synchronize(self):
        check(lib().mft_tild_synchronize(self.h))

    def structure_factor(self, species_mask):
        """mft_tild_structure_factor: |FFT[rho_group]|^2 [M] at the last density pass."""
        M = self.M
        sk = DeviceBuffer(4 * M)
        work = DeviceBuffer(8 * M)
        check(lib().mft_tild_structure_factor(self.h, C.c_uint32(species_mask), sk.ptr, 0, work.ptr))
        return sk.to_array(np.float32, M)

    def energies(self):
        pe = np.zeros(max(1, self.npot), np.float64)
        be = _d(0)
        check(lib().mft_tild_energies(self.h, _ptr(pe), C.byref(be)))
        return pe[: self.npot], be.value

    def step_energies(self, n=1):
        """n steps, then the energies a log after the last one reports (pair
        energies of that step's density pass, bonds at the new positions)."""
        pe = np.zeros(max(1, self.npot), np.float64)
        be = _d(0)
        check(lib().mft_tild_step_energies(self.h, n, _ptr(pe), C.byref(be)))
        return pe[: self.npot], be.value

    def timing(self, enable):
        ms = np.zeros(len(TK_NAMES), np.float64)
        n = np.zeros(len(TK_NAMES), np.int64)
        check(lib().mft_tild_timing(self.h, int(enable), _ptr(ms), _ptr(n)))
        return dict(zip(TK_NAMES, ms)), dict(zip(TK_NAMES, n))

    def timing_select(self, names=None):
        """Restrict timing events to the named kernel groups (None: all)."""
        mask = 0xFFFFFFFF if names is None else sum(1 << TK_NAMES.index(n) for n in names)
        check(lib().mft_tild_timing_select(self.h, C.c_uint32(mask)))

    def energies_device(self):
        """mft_tild_energies_device into a device buffer, read back afterwards:
        (pair energies [npotentials], bond energy)."""
        buf = DeviceBuffer(8 * (self.npot + 1))
        check(lib().mft_tild_energies_device(self.h, buf.ptr))
        v = buf.to_array(np.float64, (self.npot + 1,))
        buf.free()
        return v[:self.npot], float(v[self.npot])

    def bond_virial(self):
        """mft_tild_bond_virial: [N, 6] per-particle bond virial terms."""
        v = np.empty(6 * self.N, np.float32)
        check(lib().mft_tild_bond_virial(self.h, _ptr(v)))
        return v.reshape(-1, 6)

    def set_gather(self, variant):
        check(lib().mft_tild_set_gather(self.h, int(variant)))

    def set_grid_passes(self, variant):
        check(lib().mft_tild_set_grid_passes(self.h, int(variant)))

    def autotune(self, reps=5):
        c = _i(0)
        check(lib().mft_tild_autotune(self.h, reps, C.byref(c)))
        return c.value

    def autotune_timings(self):
        """mft_tild_autotune_timings: {variant: ms} of the last autotune."""
        ms = np.zeros(8)
        check(lib().mft_tild_autotune_timings(self.h, _ptr(ms), 8))
        names = ["gather_direct", "gather_lds_brick_species", "gather_lds_brick", "gather_lds_pipelined",
                 "xy_per_axis_pass", "xy_per_plane_pass", "x_fwd_direct_pass", "x_fwd_staged_pass"]
        return {k: float(v) for k, v in zip(names, ms) if v >= 0}

    def kernel_bytes(self):
        b = np.zeros(len(TK_NAMES), np.float64)
        check(lib().mft_tild_kernel_bytes(self.h, _ptr(b)))
        return dict(zip(TK_NAMES, b))


# --------------------------------------------------------------- FTS -----
class FtsPotentialC(C.Structure):
    _fields_ = [("kind", _i), ("I", _i), ("J", _i), ("chiN", _d), ("kappaN", _d), ("delt_pl", _d), ("delt_mi", _d),
                ("scheme", _i), ("zero_mean", _i)]


FTS_SCHEME = {"EM": 0, "EMPC": 1, "1S": 2}


class FtsMoleculeC(C.Structure):
    _fields_ = [("nblocks", _i), ("N", _i * 16), ("species", _i * 16), ("phi", _d), ("activity", _d)]


class FtsDescC(C.Structure):
    _fields_ = [("nx", _i), ("ny", _i), ("nz", _i), ("lx", _f), ("ly", _f), ("lz", _f), ("Nr", _d), ("C", _d),
                ("nspecies", _i), ("npotentials", _i), ("potentials", FtsPotentialC * 8), ("nmolecules", _i),
                ("molecules", FtsMoleculeC * 4), ("cl", _i), ("seed", C.c_uint64)]


def fts_desc(spec):
    """C description from a dict: Nx, L, Nr, C, nspecies, potentials
    [dict(kind, I, J, chiN, kappaN (B for Edwards), dt_pl, dt_mi)], molecules
    [dict(blocks=[(N, species)], phi or activity)], cl, seed."""
    d = FtsDescC()
    d.nx, d.ny, d.nz = spec["Nx"]
    d.lx, d.ly, d.lz = spec["L"]
    d.Nr = spec["Nr"]
    d.C = spec["C"]
    d.nspecies = spec["nspecies"]
    d.npotentials = len(spec["potentials"])
    for k, p in enumerate(spec["potentials"]):
        d.potentials[k] = _fts_pot(p)
    d.nmolecules = len(spec["molecules"])
    for k, m in enumerate(spec["molecules"]):
        mc = FtsMoleculeC()
        mc.nblocks = len(m["blocks"])
        for j, (n, s) in enumerate(m["blocks"]):
            mc.N[j] = n
            mc.species[j] = s
        mc.phi = m.get("phi", 1.0)
        mc.activity = m.get("activity", 0.0)
        d.molecules[k] = mc
    d.cl = int(spec.get("cl", True))
    d.seed = spec.get("seed", 1234)
    return d


def _c64(a):
    return np.ascontiguousarray(a, np.complex64)


def _fts_pot(p):
    return FtsPotentialC(FTS_POT[p["kind"]], p.get("I", -1), p.get("J", -1), p.get("chiN", 0.0),
                         p.get("kappaN", p.get("B", 0.0)), p.get("dt_pl", 0.0), p.get("dt_mi", 0.0),
                         FTS_SCHEME[p.get("scheme", "EM")], int(p.get("zero_mean", False)))


def _fts_mol(blocks):
    mc = FtsMoleculeC()
    mc.nblocks = len(blocks)
    for j, (n, s) in enumerate(blocks):
        mc.N[j] = n
        mc.species[j] = s
    mc.phi = 1.0
    return mc


def _ptr_array(bufs):
    arr = (_p * max(1, len(bufs)))()
    for k, b in enumerate(bufs):
        arr[k] = b.ptr if b is not None else None
    return arr


def fts_bond_factor(Nx, L, Nr):
    """mft_fts_bond_factor: exp(-k^2/(Nr-1)) as complex64 [M]."""
    M = int(np.prod(Nx))
    g = DeviceBuffer(8 * M)
    check(lib().mft_fts_bond_factor(g.ptr, _i3(Nx), _f3(L), _d(Nr), None))
    return g.to_array(np.complex64, M)


def fts_species_fields(pots, wpl, wmi, nspecies, Nr):
    """mft_fts_species_fields; pots: dicts as fts_desc, wpl/wmi: [npots] complex
    arrays (wmi entries None for non-Flory).  Returns W [nspecies, M]."""
    M = np.asarray(wpl[0]).size
    pc = (FtsPotentialC * max(1, len(pots)))(*[_fts_pot(p) for p in pots])
    wp = [DeviceBuffer.from_array(_c64(w)) for w in wpl]
    wm = [DeviceBuffer.from_array(_c64(w)) if w is not None else None for w in wmi]
    W = DeviceBuffer(8 * M * nspecies)
    check(lib().mft_fts_species_fields(W.ptr, nspecies, len(pots), pc, _ptr_array(wp), _ptr_array(wm), _d(Nr),
                                       C.c_long(M), None))
    return W.to_array(np.complex64, (nspecies, M))


def fts_update_fields(pot, wpl, wmi, rho, C_, Nr, Nx, L, cl, seed, it, noise_col):
    """mft_fts_update_fields; returns (wpl, wmi)."""
    M = int(np.prod(Nx))
    pc = _fts_pot(pot)
    a = DeviceBuffer.from_array(_c64(wpl))
    b = DeviceBuffer.from_array(_c64(wmi)) if wmi is not None else None
    rb = [DeviceBuffer.from_array(_c64(r)) for r in rho]
    check(lib().mft_fts_update_fields(a.ptr, b.ptr if b else None, C.byref(pc), _ptr_array(rb), len(rb), _d(C_),
                                      _d(Nr), _i3(Nx), _f3(L), int(cl), C.c_uint64(seed), C.c_int64(it), noise_col,
                                      None))
    return a.to_array(np.complex64, M), (b.to_array(np.complex64, M) if b else None)


def fts_propagate(W, g, blocks, Nx):
    """mft_fts_propagate: (q, qdag) [Ntot, M] from W [nspecies, M] and g [M]."""
    M = int(np.prod(Nx))
    Ntot = sum(n for n, _ in blocks)
    mc = _fts_mol(blocks)
    Wb = DeviceBuffer.from_array(_c64(W))
    gb = DeviceBuffer.from_array(_c64(g))
    q = DeviceBuffer(8 * M * Ntot)
    qd = DeviceBuffer(8 * M * Ntot)
    check(lib().mft_fts_propagate(q.ptr, qd.ptr, Wb.ptr, gb.ptr, C.byref(mc), _i3(Nx), None))
    return q.to_array(np.complex64, (Ntot, M)), qd.to_array(np.complex64, (Ntot, M))


def fts_density(rho, q, qdag, W, blocks, factor):
    """mft_fts_density: rho [nspecies, M] + the molecule's block densities."""
    rho = _c64(rho)
    M = rho.shape[-1]
    mc = _fts_mol(blocks)
    bufs = [DeviceBuffer.from_array(_c64(a)) for a in (rho, q, qdag, W)]
    check(lib().mft_fts_density(bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr, C.byref(mc), _d(factor.real),
                                _d(factor.imag), C.c_long(M), None))
    return bufs[0].to_array(np.complex64, rho.shape)


def fts_partition_function(q_last, gvol, V):
    """mft_fts_partition_function: complex Q."""
    q_last = _c64(q_last)
    M = q_last.size
    qb = DeviceBuffer.from_array(q_last)
    work = DeviceBuffer(32 * ((M + 255) // 256))
    out = (_d * 2)()
    check(lib().mft_fts_partition_function(out, qb.ptr, _d(gvol), _d(V), C.c_long(M), work.ptr, None))
    return complex(out[0], out[1])


def fts_partition_function_noalloc(q_last, gvol, V, M=None):
    """mft_fts_partition_function_noalloc: complex Q of the first M points of
    q_last (a DeviceBuffer or an array), without a work buffer."""
    if isinstance(q_last, DeviceBuffer):
        qb = q_last
    else:
        q_last = _c64(q_last)
        M = q_last.size if M is None else M
        qb = DeviceBuffer.from_array(q_last)
    out = (_d * 2)()
    check(lib().mft_fts_partition_function_noalloc(out, qb.ptr, _d(gvol), _d(V), C.c_long(M), None))
    return complex(out[0], out[1])


def fts_hamiltonian_term(pot, wpl, wmi, C_, gvol):
    """mft_fts_hamiltonian_term: complex H of one potential."""
    wpl = _c64(wpl)
    M = wpl.size
    pc = _fts_pot(pot)
    a = DeviceBuffer.from_array(wpl)
    b = DeviceBuffer.from_array(_c64(wmi)) if wmi is not None else None
    work = DeviceBuffer(32 * ((M + 255) // 256))
    out = (_d * 2)()
    check(lib().mft_fts_hamiltonian_term(out, C.byref(pc), a.ptr, b.ptr if b else None, _d(C_), _d(gvol),
                                         C.c_long(M), work.ptr, None))
    return complex(out[0], out[1])


class FtsSystem:
    """One CL-FTS system on the device (mft_fts_*)."""

    def __init__(self, spec):
        self.spec = spec
        self.h = _p()
        self._d = fts_desc(spec)
        check(lib().mft_fts_create(C.byref(self.h), C.byref(self._d)))
        self.M = int(np.prod(spec["Nx"]))
        self.npot = len(spec["potentials"])
        self.nmol = len(spec["molecules"])
        self.nspecies = spec["nspecies"]

    def close(self):
        if self.h:
            lib().mft_fts_destroy(self.h)
            self.h = _p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_fields(self, pot, wpl=None, wmi=None):
        a = _c64(wpl) if wpl is not None else None
        b = _c64(wmi) if wmi is not None else None
        check(lib().mft_fts_set_fields(self.h, pot, _ptr(a), _ptr(b)))

    def get_fields(self, pot, minus=False):
        a = np.empty(self.M, np.complex64)
        b = np.empty(self.M, np.complex64) if minus else None
        check(lib().mft_fts_get_fields(self.h, pot, _ptr(a), _ptr(b)))
        return (a, b) if minus else a

    def init(self):
        check(lib().mft_fts_init(self.h))

    def step(self, n=1):
        check(lib().mft_fts_step(self.h, n))

    def synchronize(self):
        check(lib().mft_fts_synchronize(self.h))

    def densities(self):
        r = np.empty(self.nspecies * self.M, np.complex64)
        check(lib().mft_fts_get_densities(self.h, _ptr(r)))
        return r.reshape(self.nspecies, self.M)

    def species_fields(self):
        r = np.empty(self.nspecies * self.M, np.complex64)
        check(lib().mft_fts_get_species_fields(self.h, _ptr(r)))
        return r.reshape(self.nspecies, self.M)

    def propagators(self, m):
        Nt = sum(n for n, _ in self.spec["molecules"][m]["blocks"])
        q = np.empty(Nt * self.M, np.complex64)
        qd = np.empty(Nt * self.M, np.complex64)
        check(lib().mft_fts_get_propagators(self.h, m, _ptr(q), _ptr(qd)))
        return q, qd

    def observables(self):
        Q = np.zeros(2 * self.nmol)
        H = np.zeros(2)
        parts = np.zeros(2 * (self.npot + self.nmol))
        check(lib().mft_fts_observables(self.h, _ptr(Q), _ptr(H), _ptr(parts)))
        return Q.view(np.complex128), complex(H[0], H[1]), parts.view(np.complex128)

    def chemical_potentials(self):
        mu = np.zeros(2 * self.nmol)
        check(lib().mft_fts_chemical_potentials(self.h, _ptr(mu)))
        return mu.view(np.complex128)

    def snapshot(self, slot):
        check(lib().mft_fts_snapshot(self.h, slot, 0))

    def restore(self, slot):
        check(lib().mft_fts_snapshot(self.h, slot, 1))

    def set_activity(self, m, activity):
        check(lib().mft_fts_set_activity(self.h, m, _d(activity)))

    def nsites(self):
        n = np.zeros(self.nmol)
        check(lib().mft_fts_nsites(self.h, _ptr(n)))
        return n


class TildGroup:
    """Replica group (mft_tild_group_*): engines advanced by one HIP graph of
    eight steps each; the engines stay owned by the caller."""

    def __init__(self, engines):
        self.engines = list(engines)
        arr = (_p * len(self.engines))(*[e.h for e in self.engines])
        self.h = _p()
        check(lib().mft_tild_group_create(C.byref(self.h), arr, len(self.engines)))

    def step(self, n=1):
        check(lib().mft_tild_group_step(self.h, int(n)))

    def synchronize(self):
        check(lib().mft_tild_group_synchronize(self.h))

    def close(self):
        if self.h:
            lib().mft_tild_group_destroy(self.h)
            self.h = _p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FtsBatch:
    """Many CL-FTS replicas advanced together (mft_fts_batch_*)."""

    def __init__(self, specs):
        self.specs = specs
        arr = (FtsDescC * len(specs))()
        for k, s in enumerate(specs):
            arr[k] = fts_desc(s)
        self._arr = arr
        self.h = _p()
        check(lib().mft_fts_batch_create(C.byref(self.h), arr, len(specs)))

    def close(self):
        if self.h:
            lib().mft_fts_batch_destroy(self.h)
            self.h = _p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_fields(self, rep, pot, wpl=None, wmi=None):
        a = _c64(wpl) if wpl is not None else None
        b = _c64(wmi) if wmi is not None else None
        check(lib().mft_fts_batch_set_fields(self.h, rep, pot, _ptr(a), _ptr(b)))

    def init(self):
        check(lib().mft_fts_batch_init(self.h))

    def set_chunk_bytes(self, nbytes):
        check(lib().mft_fts_batch_set_chunk_bytes(self.h, _d(nbytes)))

    def autotune(self, reps=2):
        ms = _d(0.0)
        check(lib().mft_fts_batch_autotune(self.h, int(reps), C.byref(ms)))
        return ms.value

    def step(self, n=1):
        check(lib().mft_fts_batch_step(self.h, n))

    def synchronize(self):
        check(lib().mft_fts_batch_synchronize(self.h))

    def observables(self, rep):
        s = self.specs[rep]
        nmol, npot = len(s["molecules"]), len(s["potentials"])
        Q = np.zeros(2 * nmol)
        H = np.zeros(2)
        parts = np.zeros(2 * (npot + nmol))
        check(lib().mft_fts_batch_observables(self.h, rep, _ptr(Q), _ptr(H), _ptr(parts)))
        return Q.view(np.complex128), complex(H[0], H[1]), parts.view(np.complex128)

    def densities(self, rep):
        s = self.specs[rep]
        M = int(np.prod(s["Nx"]))
        r = np.empty(s["nspecies"] * M, np.complex64)
        check(lib().mft_fts_batch_get_densities(self.h, rep, _ptr(r)))
        return r.reshape(s["nspecies"], M)

    def get_fields(self, rep, pot, minus=False):
        M = int(np.prod(self.specs[rep]["Nx"]))
        a = np.empty(M, np.complex64)
        b = np.empty(M, np.complex64) if minus else None
        check(lib().mft_fts_batch_get_fields(self.h, rep, pot, _ptr(a), _ptr(b)))
        return (a, b) if minus else a

    def species_fields(self, rep):
        s = self.specs[rep]
        M = int(np.prod(s["Nx"]))
        r = np.empty(s["nspecies"] * M, np.complex64)
        check(lib().mft_fts_batch_get_species_fields(self.h, rep, _ptr(r)))
        return r.reshape(s["nspecies"], M)

    def chemical_potentials(self, rep):
        mu = np.zeros(2 * len(self.specs[rep]["molecules"]))
        check(lib().mft_fts_batch_chemical_potentials(self.h, rep, _ptr(mu)))
        return mu.view(np.complex128)

    def replica(self, rep):
        """A view of replica rep with the single-system accessors."""
        return BatchReplica(self, rep)


class BatchReplica:
    """Accessors of one replica of an FtsBatch, named like FtsSystem's."""

    def __init__(self, batch, rep):
        self.b, self.rep = batch, rep

    def observables(self):
        return self.b.observables(self.rep)

    def densities(self):
        return self.b.densities(self.rep)

    def species_fields(self):
        return self.b.species_fields(self.rep)

    def get_fields(self, pot, minus=False):
        return self.b.get_fields(self.rep, pot, minus)

    def chemical_potentials(self):
        return self.b.chemical_potentials(self.rep)


REQ_STATE = {0: "queued", 1: "running", 2: "done", 3: "failed"}


class Server:
    """Request loop with continuous batching (mft_server_*): CL-FTS solves join
    one device-resident pool, TILD runs get their own engines; requests are
    admitted against a device-memory budget."""

    def __init__(self, hbm_budget_bytes, max_fts_replicas=256):
        self.h = _p()
        check(lib().mft_server_create(C.byref(self.h), C.c_double(hbm_budget_bytes), max_fts_replicas))
        self.meta = {}

    def close(self):
        if self.h:
            lib().mft_server_destroy(self.h)
            self.h = _p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit_fts(self, spec, fields, niter):
        d = fts_desc(spec)
        npot = len(spec["potentials"])
        M = int(np.prod(spec["Nx"]))
        keep = []
        wpl = (_p * max(1, npot))()
        wmi = (_p * max(1, npot))()
        for p in range(npot):
            for arr, key in ((wpl, "pl"), (wmi, "mi")):
                f = fields.get((p, key))
                if f is not None:
                    a = _c64(f)
                    assert a.size == M
                    keep.append(a)
                    arr[p] = a.ctypes.data
        rid = C.c_int64()
        check(lib().mft_server_submit_fts(self.h, C.byref(d), wpl, wmi, int(niter), C.byref(rid)))
        self.meta[rid.value] = ("fts", spec)
        return rid.value

    def submit_tild(self, cfg, nsteps, integrator="GJF"):
        topo = cfg["topo"]
        n = topo["type"].size
        d, keep = tild_desc(cfg["Nx"], cfg["L"], cfg["pmeorder"], len(cfg["species"]), n, cfg["potentials"],
                            cfg["bond_k"], cfg["bond_r0"], cfg["bond_style"], integrator, cfg["delt"], cfg["mass"],
                            cfg["mobility"], cfg["seed"], cfg["has_charges"])
        t = np.ascontiguousarray(topo["type"], np.int32)
        q = np.ascontiguousarray(topo["charge"], np.float32) if cfg["has_charges"] else None
        nb = np.ascontiguousarray(topo["nBonds"], np.int32)
        bt = np.ascontiguousarray(topo["bondedTo"], np.int32)
        ty = np.ascontiguousarray(topo["bondType"], np.int32)
        x = np.ascontiguousarray(cfg["x"], np.float32)
        rid = C.c_int64()
        check(lib().mft_server_submit_tild(self.h, C.byref(d), _ptr(t), _ptr(q), _ptr(nb), _ptr(bt), _ptr(ty),
                                           _ptr(x), int(nsteps), C.byref(rid)))
        self.meta[rid.value] = ("tild", cfg)
        return rid.value

    def pump(self, max_iters=50):
        nr, nq = C.c_int(), C.c_int()
        check(lib().mft_server_pump(self.h, int(max_iters), C.byref(nr), C.byref(nq)))
        return nr.value, nq.value

    def status(self, rid):
        st = C.c_int()
        check(lib().mft_server_status(self.h, C.c_int64(rid), C.byref(st)))
        return REQ_STATE[st.value]

    def error(self, rid):
        return lib().mft_server_error(self.h, C.c_int64(rid)).decode()

    def fts_result(self, rid):
        spec = self.meta[rid][1]
        nmol, npot = len(spec["molecules"]), len(spec["potentials"])
        M = int(np.prod(spec["Nx"]))
        Q = np.zeros(2 * nmol)
        H = np.zeros(2)
        parts = np.zeros(2 * (npot + nmol))
        rho = np.empty(spec["nspecies"] * M, np.complex64)
        check(lib().mft_server_fts_result(self.h, C.c_int64(rid), _ptr(Q), _ptr(H), _ptr(parts), _ptr(rho)))
        return Q.view(np.complex128), complex(H[0], H[1]), parts.view(np.complex128), rho.reshape(-1, M)

    def fts_final_state(self, rid):
        """Final fields {(pot, 'pl'|'mi'): array} and chemical potentials."""
        spec = self.meta[rid][1]
        M = int(np.prod(spec["Nx"]))
        fields, mu = {}, np.zeros(2 * len(spec["molecules"]))
        for p, pot in enumerate(spec["potentials"]):
            a = np.empty(M, np.complex64)
            b = np.empty(M, np.complex64) if pot["kind"] == "Flory" else None
            check(lib().mft_server_fts_final_state(self.h, C.c_int64(rid), p, _ptr(a), _ptr(b), _ptr(mu)))
            fields[(p, "pl")] = a
            if b is not None:
                fields[(p, "mi")] = b
        return fields, mu.view(np.complex128)

    def tild_result(self, rid):
        cfg = self.meta[rid][1]
        n = cfg["topo"]["type"].size
        x = np.empty(3 * n, np.float32)
        pe = np.zeros(max(1, len(cfg["potentials"])))
        be = _d(0)
        check(lib().mft_server_tild_result(self.h, C.c_int64(rid), _ptr(x), _ptr(pe), C.byref(be)))
        return x, pe[:len(cfg["potentials"])], be.value

    def release(self, rid):
        check(lib().mft_server_release(self.h, C.c_int64(rid)))
        self.meta.pop(rid, None)
