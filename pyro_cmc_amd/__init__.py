"""Python bindings (ctypes) of libpyro_cmc.so, the MI355X sweep-and-measure
library.  The C ABI is declared in include/pyro_cmc.h.

The product path is the HIP library: if libpyro_cmc.so is missing this
module raises on first use instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path
from typing import Optional, Sequence

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libpyro_cmc.so"

FP32, FP64 = 32, 64
METROPOLIS, LIFTED, LIFTED_ROT, OVERRELAX = 0, 1, 2, 3
STATUS = {0: "OK", 1: "INVALID", 2: "STATE", 3: "LOGIC", 4: "HIP", 5: "UNSUPPORTED"}
FAMILIES = ("sweep", "energy", "exchange", "dft", "contract", "munu", "munu_flush")

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_i8p = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")
_vp = C.c_void_p

# name -> argtypes (all return int status)
_SIGS = {
    "pyro_create": [C.c_int, _i32p, _i32p, _i32p, _i32p, C.c_int, C.c_int, C.c_uint64, C.POINTER(_vp)],
    "pyro_destroy": [_vp],
    "pyro_define_coupling": [_vp, C.c_char_p, C.c_int, _i32p, _i32p, _f64p],
    "pyro_define_J1J2J3": [_vp, C.c_double, C.c_double, C.c_double, C.c_double],
    "pyro_setup_lattice": [_vp],
    "pyro_pyrochlore_shell": [C.c_int, _i32p, _vp, C.c_int],
    "pyro_info": [_vp, _i64p],
    "pyro_site_positions": [_vp, _i32p],
    "pyro_neighbours": [_vp, _vp, _vp],
    "pyro_colouring": [_vp, _i32p],
    "pyro_set_field": [_vp, C.c_int, _f64p],
    "pyro_set_T_ref": [_vp, C.c_double],
    "pyro_set_temperatures": [_vp, _f64p],
    "pyro_set_spins": [_vp, C.c_int, _f64p],
    "pyro_get_spins": [_vp, C.c_int, _f64p],
    "pyro_set_lifted_dir": [_vp, C.c_int, _i8p],
    "pyro_get_lifted_dir": [_vp, C.c_int, _i8p],
    "pyro_randomize_spins": [_vp, C.c_uint64],
    "pyro_pass_count": [_vp, C.POINTER(C.c_uint64)],
    "pyro_sweep": [_vp, C.c_int, C.c_int, C.c_int, _vp],
    "pyro_overrelax": [_vp, C.c_int],
    "pyro_local_fields": [_vp, C.c_int, _f64p],
    "pyro_energy_per_cell": [_vp, _f64p],
    "pyro_energy_per_cell_noalloc": [_vp, C.c_int, _i32p, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                     _f64p, _vp, _vp],
    "pyro_energy_sample": [_vp],
    "pyro_pt_init": [_vp, _f64p, C.c_uint64],
    "pyro_pt_rounds": [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int],
    "pyro_pt_stats": [_vp, _vp, _vp, _vp, _vp, _vp],
    "pyro_pt_permutation": [_vp, _i32p],
    "pyro_pt_reset_stats": [_vp],
    "pyro_pt_set_exchange_interval": [_vp, C.c_int],
    "pyro_set_replica_couplings": [_vp, _vp],
    "pyro_anneal_schedule": [_vp, _f64p, C.c_int],
    "pyro_anneal_steps": [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int],
    "pyro_anneal_history": [_vp, _f64p],
    "pyro_ssf_configure": [_vp, C.c_int, _i32p, _i32p, C.c_int, C.c_int, C.c_int, C.c_int],
    "pyro_ssf_sample": [_vp, _vp],
    "pyro_pt_ssf_map": [_vp, _vp],
    "pyro_ssf_read": [_vp, C.c_int, _vp, _vp, C.POINTER(C.c_uint64)],
    "pyro_ssf_read_munu": [_vp, C.c_int, C.c_int, _f64p],
    "pyro_ssf_reset": [_vp, C.c_int],
    "pyro_spin_ft": [_vp, C.c_int, _f64p],
    "pyro_sweep_order": [_vp, C.c_int, _i32p, _i32p],
    "pyro_synchronize": [_vp],
    "pyro_set_strategy": [_vp, C.c_int],
    "pyro_get_strategy": [_vp, C.POINTER(C.c_int)],
    "pyro_set_tile_shape": [_vp, _vp],
    "pyro_set_fused_passes": [_vp, C.c_int],
    "pyro_set_wait_timeout": [_vp, C.c_double],
    "pyro_get_tile_shape": [_vp, _i32p],
    "pyro_profile_enable": [_vp, C.c_int],
    "pyro_profile_read": [_vp, _f64p, _u64p],
    "pyro_mark": [_vp, C.c_int],
    "pyro_elapsed_ms": [_vp, C.POINTER(C.c_double)],
    "pyro_lt_spectrum": [_i32p, C.c_double, C.c_double, C.c_double, C.c_double, _f64p, _i32p, _f64p],
    "pyro_batch_create": [_vp, C.c_int, C.POINTER(_vp)],
    "pyro_batch_anneal_steps": [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int],
    "pyro_batch_info": [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "pyro_batch_destroy": [_vp],
}

_lib: Optional[C.CDLL] = None


class PyroError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


def lib() -> C.CDLL:
    """Load libpyro_cmc.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build it with `python -m pyro_cmc_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        L.pyro_last_error.argtypes = []
        L.pyro_last_error.restype = C.c_char_p
        L.pyro_version.restype = C.c_int
        L.pyro_device_count.restype = C.c_int
        _lib = L
    return _lib


def exported_symbols() -> list[str]:
    return ["pyro_last_error", "pyro_version", "pyro_device_count", *_SIGS.keys()]


def _check(code: int) -> None:
    if code != 0:
        raise PyroError(code, lib().pyro_last_error().decode())


def device_count() -> int:
    return int(lib().pyro_device_count())


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# Pyrochlore cubic cell (16 sites, sl = 4 f + p), edge 8.
_FCC = np.array([[0, 0, 0], [0, 4, 4], [4, 0, 4], [4, 4, 0]], np.int32)
_LINK = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], np.int32)


def seed_key(seed_index: int) -> int:
    """64-bit RNG key of a seed index: the SplitMix64 output for state
    seed_index (the host drivers' --seed), used as context seed and
    randomize key."""
    m = (1 << 64) - 1
    z = (int(seed_index) + 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)


def host_tool(name: str):
    """Path of a built host driver (anneal, temper, spiral_byhand)."""
    from pathlib import Path

    return Path(__file__).resolve().parent / name


def lt_spectrum(L, J1, J2=0.0, J3=0.0, Delta=1.0):
    """Luttinger-Tisza spectrum of the pyrochlore J1-J2-J3 XXZ model on the
    L x L x L (or L[0] x L[1] x L[2]) grid of k (luttinger_tisza.cpp:147-181).

    Returns (lambda_min[n_k], k_star[3], eigvec[48] complex): the lowest
    eigenvalue of M(k) per k (flat order (K0 L1 + K1) L2 + K2), the first k
    of the strict minimum and the eigenvector there.  The LT energy per
    spin is lambda_min.min() / 2."""
    Ls = np.array([L] * 3 if np.isscalar(L) else list(L), np.int32)
    lam = np.empty(int(np.prod(Ls.astype(np.int64))), np.float64)
    k_star = np.zeros(3, np.int32)
    ev = np.zeros((48, 2), np.float64)
    _check(lib().pyro_lt_spectrum(Ls, float(J1), float(J2), float(J3), float(Delta), lam, k_star, ev))
    return lam, k_star, ev[:, 0] + 1j * ev[:, 1]


def energy_per_cell_noalloc(spins_ptr: int, precision: int, L, n_replicas: int, J1: float, J2: float = 0.0,
                            J3: float = 0.0, Delta: float = 1.0, B=(0.0, 0.0, 0.0), e_ptr: int = 0,
                            stream: int = 0) -> None:
    """pyro_energy_per_cell_noalloc: energy per cell of n_replicas replicas
    whose spins ([R][16 L0 L1 L2][3], FP32 or FP64) are already in device
    memory at spins_ptr, into the double[R] device buffer at e_ptr, enqueued
    on `stream` (a hipStream_t handle, 0 = the null stream); allocates
    nothing."""
    Ls = np.array([L] * 3 if np.isscalar(L) else list(L), np.int32)
    _check(lib().pyro_energy_per_cell_noalloc(_vp(spins_ptr), int(precision), Ls, int(n_replicas), float(J1),
                                              float(J2), float(J3), float(Delta), np.asarray(B, np.float64),
                                              _vp(e_ptr), _vp(stream)))


def pyrochlore_cell():
    pos = np.array([_FCC[f] + _LINK[p] for f in range(4) for p in range(4)], np.int32)
    kind = np.array([p for f in range(4) for p in range(4)], np.int32)
    return pos, kind, np.array([8, 8, 8], np.int32)


class Context:
    """R replicas of one lattice + coupling set on the GPU."""

    def __init__(self, sl_pos, sl_kind, cell, L, n_replicas=1, precision=FP64, seed=0):
        self._h = _vp()
        self.sl_pos = np.ascontiguousarray(sl_pos, np.int32).reshape(-1, 3)
        self.sl_kind = np.ascontiguousarray(sl_kind, np.int32)
        self.cell = np.ascontiguousarray(cell, np.int32)
        self.Ls = np.ascontiguousarray(L, np.int32)
        _check(lib().pyro_create(len(self.sl_kind), self.sl_pos.ravel(), self.sl_kind, self.cell,
                                 self.Ls, int(n_replicas), int(precision), int(seed), C.byref(self._h)))
        self.R = int(n_replicas)
        self.precision = int(precision)
        self.n_sl = len(self.sl_kind)
        self.n_cells = int(np.prod(self.Ls))
        self.N = self.n_sl * self.n_cells
        self.n_corr = 0

    @classmethod
    def pyrochlore(cls, L, n_replicas=1, precision=FP64, seed=0):
        pos, kind, cell = pyrochlore_cell()
        Ls = [L, L, L] if np.isscalar(L) else list(L)
        return cls(pos, kind, cell, Ls, n_replicas, precision, seed)

    def close(self):
        if self._h:
            _check(lib().pyro_destroy(self._h))
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # construction
    def define_coupling(self, name: str, rel: Sequence[Sequence[Sequence[int]]], J):
        counts = np.array([len(v) for v in rel], np.int32)
        flat = np.array([x for v in rel for d in v for x in d], np.int32) if counts.sum() else np.zeros(3, np.int32)
        _check(lib().pyro_define_coupling(self._h, name.encode(), len(rel), counts, flat,
                                          np.ascontiguousarray(J, np.float64).ravel()))

    def define_J1J2J3(self, J1, J2=0.0, J3=0.0, Delta=1.0):
        _check(lib().pyro_define_J1J2J3(self._h, J1, J2, J3, Delta))

    def setup(self):
        _check(lib().pyro_setup_lattice(self._h))
        return self

    def info(self):
        a = np.zeros(8, np.int64)
        _check(lib().pyro_info(self._h, a))
        keys = ("N", "n_cells", "n_sl", "R", "precision", "n_slot", "n_colours", "isotropic")
        return dict(zip(keys, a.tolist()))

    def positions(self):
        a = np.zeros((self.N, 3), np.int32)
        _check(lib().pyro_site_positions(self._h, a.ravel()))
        return a

    def neighbours(self):
        ns = self.info()["n_slot"]
        nbr = np.zeros((self.N, ns), np.int32)
        J = np.zeros((self.N, ns, 9), np.float64)
        _check(lib().pyro_neighbours(self._h, _ptr(nbr), _ptr(J)))
        return nbr, J.reshape(self.N, ns, 3, 3)

    def colouring(self):
        a = np.zeros(self.N, np.int32)
        _check(lib().pyro_colouring(self._h, a))
        return a

    # state
    def set_field(self, B, replica=-1):
        _check(lib().pyro_set_field(self._h, replica, np.ascontiguousarray(B, np.float64)))

    def set_T_ref(self, T_ref):
        _check(lib().pyro_set_T_ref(self._h, float(T_ref)))

    def set_temperatures(self, T):
        T = np.ascontiguousarray(np.broadcast_to(np.asarray(T, np.float64), (self.R,)))
        _check(lib().pyro_set_temperatures(self._h, T))

    def set_spins(self, S, replica=0):
        S = np.ascontiguousarray(S, np.float64).reshape(self.N, 3)
        _check(lib().pyro_set_spins(self._h, replica, S.ravel()))

    def get_spins(self, replica=0):
        S = np.zeros((self.N, 3), np.float64)
        _check(lib().pyro_get_spins(self._h, replica, S.ravel()))
        return S

    def set_lifted_dir(self, d, replica=0):
        _check(lib().pyro_set_lifted_dir(self._h, replica, np.ascontiguousarray(d, np.int8)))

    def get_lifted_dir(self, replica=0):
        d = np.zeros(self.N, np.int8)
        _check(lib().pyro_get_lifted_dir(self._h, replica, d))
        return d

    def randomize(self, seed):
        _check(lib().pyro_randomize_spins(self._h, int(seed)))

    def pass_count(self):
        v = C.c_uint64()
        _check(lib().pyro_pass_count(self._h, C.byref(v)))
        return int(v.value)

    # sweeps
    def sweep(self, kind=METROPOLIS, n_overrelax=1, n_sweeps=1, count=True):
        """Sweeps; returns the accepted counts per replica (None when
        count=False, which skips the counting)."""
        acc = np.zeros(self.R, np.uint64) if count else None
        _check(lib().pyro_sweep(self._h, kind, n_overrelax, n_sweeps, _ptr(acc)))
        return acc

    def overrelax(self, n_passes=1):
        _check(lib().pyro_overrelax(self._h, n_passes))

    # observables
    def local_fields(self, replica=0):
        h = np.zeros((self.N, 3), np.float64)
        _check(lib().pyro_local_fields(self._h, replica, h.ravel()))
        return h

    def energy_per_cell(self):
        e = np.zeros(self.R, np.float64)
        _check(lib().pyro_energy_per_cell(self._h, e))
        return e

    def energy_sample(self):
        _check(lib().pyro_energy_sample(self._h))

    # tempering
    def pt_init(self, T_ladder, swap_seed=0):
        _check(lib().pyro_pt_init(self._h, np.ascontiguousarray(T_ladder, np.float64), int(swap_seed)))

    def pt_rounds(self, kind=METROPOLIS, n_overrelax=1, n_sweep=1, sample_ssf=False, n_rounds=1):
        """sample_ssf: False / True (one SSF sample per round, after its
        sweeps) or 2 (one after each sweep of the round)."""
        _check(lib().pyro_pt_rounds(self._h, kind, n_overrelax, n_sweep, int(sample_ssf), n_rounds))

    def pt_stats(self):
        R = self.R
        att = np.zeros(max(R - 1, 1), np.uint64)
        acc = np.zeros(max(R - 1, 1), np.uint64)
        E = np.zeros(R)
        E2 = np.zeros(R)
        n = np.zeros(R, np.uint64)
        _check(lib().pyro_pt_stats(self._h, _ptr(att), _ptr(acc), _ptr(E), _ptr(E2), _ptr(n)))
        return dict(attempts=att[: R - 1], accepts=acc[: R - 1], E=E, E2=E2, n=n)

    def pt_permutation(self):
        a = np.zeros(self.R, np.int32)
        _check(lib().pyro_pt_permutation(self._h, a))
        return a

    def pt_reset_stats(self):
        _check(lib().pyro_pt_reset_stats(self._h))

    def pt_exchange_interval(self, every):
        """Exchange after every `every`-th tempering round only (default 1)."""
        _check(lib().pyro_pt_set_exchange_interval(self._h, int(every)))

    # batches of independent anneals
    def set_replica_couplings(self, J1J2J3Delta):
        """Per-replica (J1, J2, J3, Delta), shape [R][4]; None restores the
        context's couplings."""
        arr = None if J1J2J3Delta is None else np.ascontiguousarray(J1J2J3Delta, np.float64).reshape(self.R, 4)
        _check(lib().pyro_set_replica_couplings(self._h, _ptr(arr)))

    def anneal_schedule(self, T):
        T = np.ascontiguousarray(T, np.float64)
        self.n_sched = len(T)
        _check(lib().pyro_anneal_schedule(self._h, T, len(T)))

    def anneal_steps(self, kind=METROPOLIS, n_overrelax=1, n_sweep=1, sample_ssf=False, n_steps=1):
        _check(lib().pyro_anneal_steps(self._h, kind, n_overrelax, n_sweep, int(bool(sample_ssf)), n_steps))

    def anneal_history(self):
        E = np.zeros((self.R, self.n_sched))
        _check(lib().pyro_anneal_history(self._h, E.ravel()))
        return E

    # structure factor
    def ssf_configure(self, correlators=("xx", "yy", "zz"), n_slots=1, store_error_term=True,
                      munu=False, munu_buffer=16):
        ax = {"x": 0, "y": 1, "z": 2}
        al = np.array([ax[c[0]] for c in correlators], np.int32)
        be = np.array([ax[c[1]] for c in correlators], np.int32)
        self.n_corr = len(correlators)
        self.n_slots = n_slots
        _check(lib().pyro_ssf_configure(self._h, len(al), al, be, n_slots, int(store_error_term),
                                        int(bool(munu)), int(munu_buffer)))

    def pt_ssf_map(self, mapping=None):
        arr = None if mapping is None else np.ascontiguousarray(mapping, np.int32)
        _check(lib().pyro_pt_ssf_map(self._h, _ptr(arr)))

    def ssf_sample(self, slot_of_replica=None):
        arr = None if slot_of_replica is None else np.ascontiguousarray(slot_of_replica, np.int32)
        _check(lib().pyro_ssf_sample(self._h, _ptr(arr)))

    def ssf_read(self, slot=0, with_sq=True):
        S = np.zeros((self.n_corr, self.n_cells, 2))
        S2 = np.zeros((self.n_corr, self.n_cells, 2)) if with_sq else None
        n = C.c_uint64()
        _check(lib().pyro_ssf_read(self._h, slot, _ptr(S), _ptr(S2), C.byref(n)))
        return S, S2, int(n.value)

    def ssf_read_munu(self, slot, c):
        a = np.zeros((self.n_cells, self.n_sl, self.n_sl, 2))
        _check(lib().pyro_ssf_read_munu(self._h, slot, c, a.ravel()))
        return a

    def ssf_reset(self, slot=-1):
        _check(lib().pyro_ssf_reset(self._h, slot))

    def spin_ft(self, replica=0):
        a = np.zeros((3, self.n_sl, self.n_cells, 2))
        _check(lib().pyro_spin_ft(self._h, replica, a.ravel()))
        return a

    def sweep_order(self, n_overrelax=1):
        n = self.N * (1 + n_overrelax)
        site = np.zeros(n, np.int32)
        koff = np.zeros(n, np.int32)
        _check(lib().pyro_sweep_order(self._h, n_overrelax, site, koff))
        return site, koff

    # execution control
    def synchronize(self):
        _check(lib().pyro_synchronize(self._h))

    def set_strategy(self, s):
        _check(lib().pyro_set_strategy(self._h, s))

    def set_tile_shape(self, T=None):
        """Tile shape (cells per axis) of the tiled sweep; None = automatic."""
        t = None if T is None else np.ascontiguousarray(T, np.int32)
        _check(lib().pyro_set_tile_shape(self._h, _ptr(t)))

    def set_wait_timeout(self, seconds):
        """Bound on each stage wait of the XCD-resident sweep (strategy 6)
        (pyro_set_wait_timeout)."""
        _check(lib().pyro_set_wait_timeout(self._h, float(seconds)))

    def set_fused_passes(self, on=True):
        """Tiled sweeps: every pass of a site back to back within its colour
        stage, from one field evaluation (pyro_set_fused_passes)."""
        _check(lib().pyro_set_fused_passes(self._h, int(bool(on))))

    def tile_shape(self):
        t = np.zeros(3, np.int32)
        _check(lib().pyro_get_tile_shape(self._h, t))
        return tuple(int(v) for v in t)

    def strategy(self):
        v = C.c_int()
        _check(lib().pyro_get_strategy(self._h, C.byref(v)))
        return v.value

    def profile(self, on=True):
        _check(lib().pyro_profile_enable(self._h, int(on)))

    def profile_read(self):
        ms = np.zeros(len(FAMILIES))
        n = np.zeros(len(FAMILIES), np.uint64)
        _check(lib().pyro_profile_read(self._h, ms, n))
        return {f: (float(ms[i]), int(n[i])) for i, f in enumerate(FAMILIES)}

    def mark(self, which):
        _check(lib().pyro_mark(self._h, which))

    def elapsed_ms(self):
        v = C.c_double()
        _check(lib().pyro_elapsed_ms(self._h, C.byref(v)))
        return float(v.value)


class Batch:
    """Contexts stepped together (pyro_batch_*): one graph per anneal step,
    the stencil-strategy contexts sweeping in one persistent launch.  The
    contexts stay open while the batch is."""

    def __init__(self, contexts):
        self.contexts = list(contexts)
        hs = (_vp * len(self.contexts))(*[c._h for c in self.contexts])
        self._h = _vp()
        _check(lib().pyro_batch_create(hs, len(self.contexts), C.byref(self._h)))

    def anneal_steps(self, kind=METROPOLIS, n_overrelax=1, n_sweep=1, sample_ssf=False, n_steps=1):
        _check(lib().pyro_batch_anneal_steps(self._h, kind, n_overrelax, n_sweep, int(bool(sample_ssf)), n_steps))

    def info(self):
        """(contexts in the persistent launch, cell blocks per workgroup and
        colour stage, workgroups of the launch)."""
        v = [C.c_int() for _ in range(3)]
        _check(lib().pyro_batch_info(self._h, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def close(self):
        if self._h:
            _check(lib().pyro_batch_destroy(self._h))
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
