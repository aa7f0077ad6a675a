"""The allocation-free energy entry point (pyro_energy_per_cell_noalloc)
against pyro_energy_per_cell on contexts holding the same spins, couplings
and field: bit for bit, on two calls in a row (the second with a smaller
lattice and fewer replicas, so the scratch array holds the first call's
partials when the second starts), through both energy kernels (tiles in
LDS, and the global-memory stencil on a lattice no tile divides), and a
size beyond the scratch array rejected before any launch.  The test's
device buffers come from hipMalloc of the HIP runtime the library itself
loads (through ctypes); the entry point allocates none."""
import ctypes as C

import numpy as np
import pytest

from helpers import random_spins

pytestmark = pytest.mark.gpu


class _Hip:
    """hipMalloc / hipMemcpy / hipFree / hipDeviceSynchronize of the HIP
    runtime libpyro_cmc.so is linked against (test plumbing only)."""

    def __init__(self, pyro):
        pyro.lib()  # loads libamdhip64 as the library's dependency
        self.h = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)  # the soname the library loaded
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize"):
            getattr(self.h, f).restype = C.c_int

    def _ok(self, rc, what):
        assert rc == 0, f"{what} failed with {rc}"

    def alloc(self, nbytes):
        p = C.c_void_p()
        self._ok(self.h.hipMalloc(C.byref(p), nbytes), "hipMalloc")
        return p.value

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        self._ok(self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        return p

    def download(self, p, shape, dtype):
        out = np.empty(shape, dtype)
        self._ok(self.h.hipDeviceSynchronize(), "hipDeviceSynchronize")
        self._ok(self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def free(self, p):
        self._ok(self.h.hipFree(p), "hipFree")

MODEL = dict(J1=1.0, J2=-0.05, J3=0.05, Delta=0.8)
B = (0.1, -0.2, 0.3)


def _reference(pyro, S, prec, L):
    R = len(S)
    ctx = pyro.Context.pyrochlore(L, n_replicas=R, precision=prec, seed=1)
    ctx.define_J1J2J3(MODEL["J1"], MODEL["J2"], MODEL["J3"], MODEL["Delta"])
    ctx.setup()
    ctx.set_field(np.array(B))
    for r in range(R):
        ctx.set_spins(S[r], r)
    e = ctx.energy_per_cell()
    ctx.close()
    return e


def _noalloc(pyro, hip, S, prec, L):
    R = len(S)
    d = hip.upload(np.asarray(S, np.float32 if prec == 32 else np.float64))
    e = hip.upload(np.full(R, np.nan))
    pyro.energy_per_cell_noalloc(d, prec, L, R, B=B, e_ptr=e, stream=0, **MODEL)
    out = hip.download(e, (R,), np.float64)
    hip.free(d)
    hip.free(e)
    return out


@pytest.mark.parametrize("prec", [32, 64])
def test_noalloc_energy_matches_context_energy(gpu, prec):
    hip = _Hip(gpu)
    rng = np.random.default_rng(prec)
    for L, R in (((12, 12, 12), 8), ((3, 3, 3), 3)):  # tiled energy kernel, then the stencil kernel
        N = 16 * int(np.prod(L))
        S = np.array([random_spins(N, rng) for _ in range(R)])
        got = _noalloc(gpu, hip, S, prec, L)
        want = _reference(gpu, S, prec, L)
        print(f"L={L} R={R} FP{prec}: {got[:3]}")
        assert np.all(np.isfinite(got))
        assert np.array_equal(got, want), (got, want)


def test_noalloc_rejects_sizes_beyond_its_scratch(gpu):
    hip = _Hip(gpu)
    L, R = (47, 47, 47), 21  # stencil kernel (no tile divides 47): 21 x 25956 partials > 2^19
    d = hip.alloc(R * 16 * 47 ** 3 * 3 * 4)
    e = hip.alloc(R * 8)
    with pytest.raises(gpu.PyroError) as err:
        gpu.energy_per_cell_noalloc(d, 32, L, R, B=B, e_ptr=e, **MODEL)
    assert err.value.code == 1 and "scratch" in str(err.value)
    hip.free(d)
    hip.free(e)
    # a size that fits still works after the rejection
    got = _noalloc(gpu, hip, np.array([random_spins(16 * 27, np.random.default_rng(1))]), 32, (3, 3, 3))
    assert np.isfinite(got[0])
