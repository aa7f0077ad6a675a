// Device-side single-site update math shared by the generic and the
// pyrochlore-specialised sweep kernels.  Included by .hip sources only.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include "kernels.hpp"
#include "rng.hpp"

namespace pyro {

// FP64 uses the correctly rounded library functions (the reference's
// arithmetic).  FP32 uses the hardware approximations: v_log_f32 / v_exp_f32
// (base 2, ~1 ulp), v_sin_f32 / v_cos_f32, which take their argument in
// revolutions, so sin(2 pi u) is one instruction, v_sqrt_f32, v_rsq_f32 and
// v_rcp_f32 (about 1 ulp each, one instruction each instead of the
// correctly rounded expansions).
template <typename Real> __device__ __forceinline__ Real dsqrt(Real x);
template <> __device__ __forceinline__ float dsqrt<float>(float x) { return __builtin_amdgcn_sqrtf(x); }
template <> __device__ __forceinline__ double dsqrt<double>(double x) { return sqrt(x); }
template <typename Real> __device__ __forceinline__ Real dexp(Real x);
template <> __device__ __forceinline__ float dexp<float>(float x) { return __expf(x); }
template <> __device__ __forceinline__ double dexp<double>(double x) { return exp(x); }
template <typename Real> __device__ __forceinline__ Real dlog(Real x);
template <> __device__ __forceinline__ float dlog<float>(float x) { return __logf(x); }
template <> __device__ __forceinline__ double dlog<double>(double x) { return log(x); }
// sin and cos of an angle in radians
template <typename Real> __device__ __forceinline__ void dsincos(Real x, Real* s, Real* c);
template <> __device__ __forceinline__ void dsincos<float>(float x, float* s, float* c) { __sincosf(x, s, c); }
template <> __device__ __forceinline__ void dsincos<double>(double x, double* s, double* c) { sincos(x, s, c); }
// sin and cos of 2 pi u for a uniform u in (0, 1)
template <typename Real> __device__ __forceinline__ void dsincos_rev(Real u, Real* s, Real* c);
template <> __device__ __forceinline__ void dsincos_rev<float>(float u, float* s, float* c) {
    *s = __builtin_amdgcn_sinf(u);
    *c = __builtin_amdgcn_cosf(u);
}
template <> __device__ __forceinline__ void dsincos_rev<double>(double u, double* s, double* c) {
    sincos(6.28318530717958647692 * u, s, c);
}
// a / b: reciprocal and multiply in FP32, a true division in FP64
template <typename Real> __device__ __forceinline__ Real ddiv(Real a, Real b);
template <> __device__ __forceinline__ float ddiv<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
template <> __device__ __forceinline__ double ddiv<double>(double a, double b) { return a / b; }
// v <- v / |v|: v * rsq(|v|^2) in FP32, division by the square root in FP64
template <typename Real>
__device__ __forceinline__ void normalize3(Real& x, Real& y, Real& z) {
    // explicit fma: the contraction is the same in every kernel that
    // inlines this, whatever code surrounds it
    const Real n2 = fma(z, z, fma(y, y, x * x));
    if constexpr (sizeof(Real) == 4) {
        const Real inv = __builtin_amdgcn_rsqf(n2);
        x *= inv;
        y *= inv;
        z *= inv;
    } else {
        const Real nrm = sqrt(n2);
        x /= nrm;
        y /= nrm;
        z /= nrm;
    }
}
// a . b and a x b with explicit fma: every kernel that inlines an update
// rounds its dot and cross products alike, whatever code surrounds it (under
// -ffp-contract=fast the compiler's choice of which product to fuse follows
// the code around it)
template <typename Real>
__device__ __forceinline__ Real dot3(Real ax, Real ay, Real az, Real bx, Real by, Real bz) {
    return fma(az, bz, fma(ay, by, ax * bx));
}
template <typename Real>
__device__ __forceinline__ void cross3(Real ax, Real ay, Real az, Real bx, Real by, Real bz, Real& cx, Real& cy,
                                       Real& cz) {
    cx = fma(ay, bz, -(az * by));
    cy = fma(az, bx, -(ax * bz));
    cz = fma(ax, by, -(ay * bx));
}
// 1 / sqrt(x)
template <typename Real> __device__ __forceinline__ Real dinv_sqrt(Real x);
template <> __device__ __forceinline__ float dinv_sqrt<float>(float x) { return __builtin_amdgcn_rsqf(x); }
template <> __device__ __forceinline__ double dinv_sqrt<double>(double x) { return 1.0 / sqrt(x); }

// Load base[idx] through a 32-bit unsigned byte offset from a wave-uniform
// base, so the compiler can use SGPR-base global loads instead of a 64-bit
// address computation per load.
template <typename Real>
__device__ __forceinline__ Real ld32(const Real* __restrict__ base, unsigned idx) {
    return *reinterpret_cast<const Real*>(reinterpret_cast<const char*>(base) +
                                          static_cast<size_t>(idx * unsigned(sizeof(Real))));
}

// One site's spin (x, y, z), the unit of the spin layout S[r][i][3].
template <typename Real>
struct Vec3 {
    Real x, y, z;
};
// Spin i of a replica's block Sr (= S + r * 3 N) in one 12-byte (FP32) or
// 24-byte (FP64) access, through a 32-bit byte offset from a wave-uniform
// base as ld32 does.
template <typename Real>
__device__ __forceinline__ Vec3<Real> ld3(const Real* __restrict__ Sr, unsigned i) {
    return *reinterpret_cast<const Vec3<Real>*>(reinterpret_cast<const char*>(Sr) +
                                                static_cast<size_t>(i * unsigned(3 * sizeof(Real))));
}
template <typename Real>
__device__ __forceinline__ void st3(Real* __restrict__ Sr, unsigned i, Real x, Real y, Real z) {
    *reinterpret_cast<Vec3<Real>*>(reinterpret_cast<char*>(Sr) + static_cast<size_t>(i * unsigned(3 * sizeof(Real)))) =
        Vec3<Real>{x, y, z};
}

// FP32 spins are renormalised after every accepted change (drift control);
// FP64 follows the reference exactly (no renormalisation after OR).
template <typename Real> struct Renorm { static constexpr bool value = false; };
template <> struct Renorm<float> { static constexpr bool value = true; };

// The five 24-bit uniforms of (site, replica, pass): one Philox4x32-10
// block.  They do not depend on the field.
template <typename Real>
__device__ __forceinline__ Uniform5 site_uniforms(const DeviceState<Real>& st, int site, int r,
                                                  unsigned long long pass) {
    return split_uniform5(philox4x32_10(
        Words4{uint32_t(site), uint32_t(r), uint32_t(pass), uint32_t(pass >> 32)}, st.k0, st.k1));
}

// What an update needs of its replica: the temperature of the slot the
// replica occupies, the proposal width sqrt(T / T_ref) (computed in FP64 as
// the reference does) and the global field.  Constant within a launch, so
// kernels that update many sites of one replica compute it once.
template <typename Real>
struct ReplicaParams {
    Real T, sig, Bx, By, Bz;
};
template <typename Real>
__device__ __forceinline__ ReplicaParams<Real> replica_params(const DeviceState<Real>& st, int r) {
    const double Td = st.T_slot[st.slot_of[r]];
    return ReplicaParams<Real>{Real(Td), Real(sqrt(Td / st.T_ref)), st.B[3 * r + 0], st.B[3 * r + 1],
                               st.B[3 * r + 2]};
}

// update_spin with the replica's parameters and the site's uniforms already
// drawn (replica_params, site_uniforms).
template <typename Real, int KIND>
__device__ __forceinline__ bool update_spin_u(const ReplicaParams<Real>& p, Real& px, Real& py, Real& pz,
                                              bool& changed, const Uniform5& u, Real hx, Real hy, Real hz,
                                              int8_t* dp) {
    changed = false;
    hx -= p.Bx;  // h_loc = local_field - global_field
    hy -= p.By;
    hz -= p.Bz;

    const Real sx = px, sy = py, sz = pz;
    const Real T = p.T;

    if constexpr (KIND == 0) {
        // Gaussian-kick Metropolis (MC.cpp:111-130).
        const Real sig = p.sig;
        Real s1, c1, s2, c2;
        const Real r1 = dsqrt<Real>(Real(-2) * dlog<Real>(u24<Real>(u.m[0])));
        dsincos_rev<Real>(u24<Real>(u.m[1]), &s1, &c1);
        const Real r2 = dsqrt<Real>(Real(-2) * dlog<Real>(u24<Real>(u.m[2])));
        dsincos_rev<Real>(u24<Real>(u.m[3]), &s2, &c2);
        Real nx = fma(sig, r1 * c1, sx);
        Real ny = fma(sig, r1 * s1, sy);
        Real nz = fma(sig, r2 * c2, sz);
        normalize3<Real>(nx, ny, nz);
        const Real curE = dot3(sx, sy, sz, hx, hy, hz);
        const Real newE = dot3(nx, ny, nz, hx, hy, hz);
        const Real dE = newE - curE;
        if (dE < Real(0) || u24<Real>(u.m[4]) < dexp<Real>(ddiv<Real>(-dE, T))) {
            px = nx;
            py = ny;
            pz = nz;
            changed = true;
            return true;
        }
        return false;
    } else if constexpr (KIND == 3) {
        // Overrelaxation: rotation by 2*pi*U about h_loc (MC.cpp:102-109, 145-153).
        // Products and sums as explicit fma, so that every kernel inlining
        // this (with or without other work around it) rounds alike.
        const Real n2 = fma(hz, hz, fma(hy, hy, hx * hx));
        if (!(n2 > Real(0))) return false;  // h_loc = 0: no axis, spin kept
        Real s, c;
        dsincos_rev<Real>(u24<Real>(u.m[0]), &s, &c);
        const Real d = fma(hz, sz, fma(hy, sy, hx * sx));
        const Real f = ddiv<Real>(Real(1) - c, n2) * d;
        const Real g = ddiv<Real>(s, dsqrt<Real>(n2));
        Real cx, cy, cz;
        cross3(hx, hy, hz, sx, sy, sz, cx, cy, cz);
        Real nx = fma(g, cx, fma(c, sx, f * hx));
        Real ny = fma(g, cy, fma(c, sy, f * hy));
        Real nz = fma(g, cz, fma(c, sz, f * hz));
        if constexpr (Renorm<Real>::value) {
            normalize3<Real>(nx, ny, nz);
        }
        px = nx;
        py = ny;
        pz = nz;
        changed = true;
        return false;
    } else if constexpr (KIND == 1) {
        // Lifted Metropolis, displacement form (MC.cpp:249-281).
        const int8_t d = *dp;
        const Real hn = dsqrt<Real>(dot3(hx, hy, hz, hx, hy, hz));
        if (!(hn > Real(0))) {  // no field direction: bounce
            *dp = int8_t(-d);
            return false;
        }
        const Real ex = ddiv<Real>(hx, hn), ey = ddiv<Real>(hy, hn), ez = ddiv<Real>(hz, hn);
        const Real uu = dot3(ex, ey, ez, sx, sy, sz);
        const Real du = (Real(d) * p.sig) * (-dlog<Real>(u24<Real>(u.m[0])));
        const Real nu = uu + du;
        const Real one_uu = fma(-uu, uu, Real(1));
        if (nu > Real(1) || nu < Real(-1) || fabs(one_uu) < Real(1e-16)) {
            *dp = int8_t(-d);
            return false;
        }
        const Real f = dsqrt<Real>(ddiv<Real>(fma(-nu, nu, Real(1)), one_uu));
        Real nx = fma(nu, ex, f * fma(-uu, ex, sx));
        Real ny = fma(nu, ey, f * fma(-uu, ey, sy));
        Real nz = fma(nu, ez, f * fma(-uu, ez, sz));
        normalize3<Real>(nx, ny, nz);
        const Real curE = dot3(sx, sy, sz, hx, hy, hz);
        const Real dE = dot3(nx, ny, nz, hx, hy, hz) - curE;
        if (u24<Real>(u.m[1]) < dexp<Real>(ddiv<Real>(-dE, T))) {
            px = nx;
            py = ny;
            pz = nz;
            changed = true;
            return true;
        }
        *dp = int8_t(-d);
        return false;
    } else {
        // Lifted rotation Metropolis with reflecting poles (MC.cpp:194-244).
        const int8_t d = *dp;
        const Real curE = dot3(sx, sy, sz, hx, hy, hz);
        Real ax, ay, az;
        cross3(sx, sy, sz, hx, hy, hz, ax, ay, az);
        const Real a2 = dot3(ax, ay, az, ax, ay, az);
        if (a2 <= Real(1e-20)) {
            *dp = int8_t(-d);
            return false;
        }
        const Real th = Real(d) * p.sig;
        Real s, c;
        dsincos<Real>(th, &s, &c);
        const Real dd = dot3(ax, ay, az, sx, sy, sz);
        const Real f = ddiv<Real>(Real(1) - c, a2) * dd;
        const Real g = ddiv<Real>(s, dsqrt<Real>(a2));
        Real cx, cy, cz;
        cross3(ax, ay, az, sx, sy, sz, cx, cy, cz);
        Real nx = fma(g, cx, fma(f, ax, c * sx));
        Real ny = fma(g, cy, fma(f, ay, c * sy));
        Real nz = fma(g, cz, fma(f, az, c * sz));
        normalize3<Real>(nx, ny, nz);
        Real bx, by, bz;
        cross3(nx, ny, nz, hx, hy, hz, bx, by, bz);
        const Real b2 = dot3(bx, by, bz, bx, by, bz);
        if (dot3(ax, ay, az, bx, by, bz) < Real(0)) {
            *dp = int8_t(-d);
            return false;
        }
        const Real dE = dot3(nx, ny, nz, hx, hy, hz) - curE;
        const Real ratio = dsqrt<Real>(ddiv<Real>(b2, a2)) * dexp<Real>(ddiv<Real>(-dE, T));
        if (u24<Real>(u.m[0]) < ratio) {
            px = nx;
            py = ny;
            pz = nz;
            changed = true;
            return true;
        }
        *dp = int8_t(-d);
        return false;
    }
}

// One single-site update of kind KIND of the spin (px, py, pz) at `site` of
// replica r (parameters p = replica_params(st, r)), pass index `pass`, given
// the local field (hx, hy, hz) without the global field.  Returns whether
// the proposal was accepted (OR: always false); `changed` tells whether the
// spin was modified.  dp points at the site's lifted direction (read and
// flipped by the lifted kinds only).
template <typename Real, int KIND>
__device__ __forceinline__ bool update_spin_p(const ReplicaParams<Real>& p, const DeviceState<Real>& st, Real& px,
                                              Real& py, Real& pz, bool& changed, int site, int r,
                                              unsigned long long pass, Real hx, Real hy, Real hz, int8_t* dp) {
    return update_spin_u<Real, KIND>(p, px, py, pz, changed, site_uniforms(st, site, r, pass), hx, hy, hz, dp);
}

// update_spin_p on the spin stored in global memory (Sr = S + r * 3 N,
// site i at Sr[3 i .. 3 i + 2]).
template <typename Real, int KIND>
__device__ __forceinline__ bool apply_update_p(const ReplicaParams<Real>& p, const DeviceState<Real>& st,
                                               Real* __restrict__ Sr, int site, int r, unsigned long long pass,
                                               Real hx, Real hy, Real hz) {
    const Vec3<Real> v = ld3(Sr, unsigned(site));
    Real x = v.x, y = v.y, z = v.z;
    bool changed;
    const bool acc = update_spin_p<Real, KIND>(p, st, x, y, z, changed, site, r, pass, hx, hy, hz,
                                               st.dir + size_t(r) * size_t(st.N) + site);
    if (changed) st3(Sr, unsigned(site), x, y, z);
    return acc;
}
template <typename Real, int KIND>
__device__ __forceinline__ bool apply_update(const DeviceState<Real>& st, Real* __restrict__ Sr, int site, int r,
                                             unsigned long long pass, Real hx, Real hy, Real hz) {
    return apply_update_p<Real, KIND>(replica_params(st, r), st, Sr, site, r, pass, hx, hy, hz);
}

// Wave-level count of accepted updates added to acc[r].
// Workgroup g of a 1-D grid of gx * R workgroups -> (block bx within the
// replica, replica r).  The dispatcher hands workgroups to the 8 XCDs
// round-robin by linear id, so when R is a multiple of 8 this places replica
// r wholly on XCD r % 8 and each XCD's L2 caches only its own replicas'
// spins (the neighbour gathers of a colour pass then hit in L2).
__device__ __forceinline__ void xcd_split(unsigned g, unsigned gx, unsigned R, unsigned& bx,
                                          unsigned& r) {
    if ((R & 7u) == 0u) {
        const unsigned s = g >> 3, j = s / gx;
        bx = s - j * gx;
        r = (g & 7u) + 8u * j;
    } else {
        r = g / gx;
        bx = g - r * gx;
    }
}

// Per-colour kernels: wave w of block bx owns counter slot bx * (block
// size / 64) + w of its replica.  Must be reached by every lane of the wave.
template <typename Real>
__device__ __forceinline__ void count_accepts(const DeviceState<Real>& st, int r, unsigned bx, bool a) {
    if (st.acc_wave == nullptr) return;
    const unsigned long long mask = __ballot(a);
    if ((threadIdx.x & 63) == 0 && mask) {
        unsigned int* p = st.acc_wave + size_t(r) * st.acc_waves + bx * (blockDim.x >> 6) + (threadIdx.x >> 6);
        *p += unsigned(__popcll(mask));
    }
}

}  // namespace pyro
