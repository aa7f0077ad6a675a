// Sweep kernels specialised to the pyrochlore cubic cell with the
// J1-J2-J3 couplings of build_J1J2J3_h (cli_bits.hpp:144-207).
//
// The neighbour stencil is a compile-time table (pyro_stencil.hpp), so a
// site's 30 neighbour addresses are its cell's wrapped coordinates plus
// constants: no neighbour-index loads, no dependent load chain.  The field
// is summed by shell,
//   h_i = J1 sum_nn1 S_j + J2 sum_nn2 S_j + J3 sum_nn3 S_j
//         + J1 (Delta - 1)/3 l_i sum_nn1 (l_j . S_j),
// where l_p is the [111] link of pyro_sl p (z_p = l_p / sqrt 3), which is
// h_i = sum_j J_ij S_j for J_ij = J1 [I + (Delta - 1) z_i z_j^T] on nn1 bonds
// (cli_bits.hpp:168-197) and isotropic J2, J3.
//
// Two launch shapes visit sites in the same colour order with the same
// per-site counters and the same arithmetic, so they give bit-identical
// results:
//   resident: one workgroup per replica runs whole sweeps, colours separated
//             by workgroup barriers; a replica's spins (12 B x 16 L^3) stay
//             in the L2 of the XCD its workgroup runs on;
//   colour:   one launch per colour over all replicas (large lattices).

#pragma once

#include <hip/hip_runtime.h>

#include <utility>

#include "device_update.hpp"
#include "kernels.hpp"
#include "pyro_stencil.hpp"

namespace pyro {

namespace {

using pyrostencil::kStencil;

static_assert(kStencil.n[0][29].sl >= 0, "stencil must hold 30 neighbours");

template <int... I, typename F>
__device__ __forceinline__ void static_for_impl(std::integer_sequence<int, I...>, F&& f) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    static_for_impl(std::make_integer_sequence<int, N>{}, f);
}

// Wrapped row offsets of the three cells along each axis around n.
struct CellAxes {
    int a0[3], a1[3], a2[3];
};

// n / d and n % d for 0 <= n < 2^24 through a float reciprocal and one
// correction step (integer division is a long instruction sequence).
__device__ __forceinline__ int fast_div(int n, int d, int& rem) {
    int q = int(float(n) * __frcp_rn(float(d)));
    int r = n - q * d;
    if (r >= d) { ++q; r -= d; }
    if (r < 0) { --q; r += d; }
    rem = r;
    return q;
}

__device__ __forceinline__ CellAxes cell_axes_of(int cell, int L0, int L1, int L2);

__device__ __forceinline__ CellAxes cell_axes(int n0, int n1, int n2, int L0, int L1, int L2) {
    CellAxes c;
    const int L12 = L1 * L2;
    c.a0[0] = (n0 == 0 ? L0 - 1 : n0 - 1) * L12;
    c.a0[1] = n0 * L12;
    c.a0[2] = (n0 == L0 - 1 ? 0 : n0 + 1) * L12;
    c.a1[0] = (n1 == 0 ? L1 - 1 : n1 - 1) * L2;
    c.a1[1] = n1 * L2;
    c.a1[2] = (n1 == L1 - 1 ? 0 : n1 + 1) * L2;
    c.a2[0] = (n2 == 0 ? L2 - 1 : n2 - 1);
    c.a2[1] = n2;
    c.a2[2] = (n2 == L2 - 1 ? 0 : n2 + 1);
    return c;
}

__device__ __forceinline__ CellAxes cell_axes_of(int cell, int L0, int L1, int L2) {
    int n2, n1;
    const int t = fast_div(cell, L2, n2);
    const int n0 = fast_div(t, L1, n1);
    return cell_axes(n0, n1, n2, L0, L1, L2);
}

// Neighbour sums over the bonds to sites of earlier colours: per shell,
// and the J1 neighbours again per target pyro_sl (for the XXZ term).
template <typename Real>
struct LowSums {
    Real s[3][3];
    Real v[4][3];
};

// Field of a site of sublattice SL; fetch(n, x, y, z) loads the spin of
// stencil neighbour n (compile-time target sublattice and cell offset).
// FWD: only the neighbours on higher sublattices (no bond joins equal
// sublattices), so summing S_i . h_i over all sites counts every bond once.
// SPLIT = NCOL > 0: also the sums over the neighbours of earlier colours
// (colour = sublattice mod NCOL) into *low, from the same loads.  They are
// sums of the loaded values only (no products), so the full field's
// arithmetic is that of SPLIT = 0.
// Whether stencil neighbour k of sublattice SL is the first one its shell
// sums (the shell's sum starts from its value rather than from 0 + value).
template <int SL, bool FWD>
constexpr bool first_in_shell(int k) {
    const int shell = k < 6 ? 0 : (k < 18 ? 1 : 2);
    const int begin = shell == 0 ? 0 : (shell == 1 ? 6 : 18);
    for (int j = begin; j < k; ++j)
        if (!FWD || kStencil.n[SL][j].sl > SL) return false;
    return true;
}

template <typename Real, int SHELLS, int SL, bool FWD = false, int SPLIT = 0, typename Fetch>
__device__ __forceinline__ void pyro_field_with(Fetch&& fetch, const PyroCouplings<Real>& J, Real& hx,
                                                Real& hy, Real& hz, LowSums<Real>* low = nullptr) {
    constexpr int NS = pyrostencil::slots_for(SHELLS);
    Real s[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // per shell x, y, z
    Real zs = Real(0);
    if constexpr (SPLIT > 0) *low = LowSums<Real>{};
    static_for<NS>([&](auto kk) __attribute__((always_inline)) {
        constexpr int k = decltype(kk)::value;
        constexpr pyrostencil::Nbr n = kStencil.n[SL][k];
        constexpr int shell = k < 6 ? 0 : (k < 18 ? 1 : 2);
        static_assert(n.sl != SL, "no bond joins a sublattice to itself");
        if constexpr (!FWD || n.sl > SL) {
            Real x, y, z;
            fetch(n, x, y, z);
            // a shell's first term starts its sum: 0 + x would cost an add
            // per component (it cannot be folded away, 0 + -0 being +0)
            if constexpr (first_in_shell<SL, FWD>(k)) {
                s[shell][0] = x;
                s[shell][1] = y;
                s[shell][2] = z;
            } else {
                s[shell][0] += x;
                s[shell][1] += y;
                s[shell][2] += z;
            }
            constexpr int tp = n.sl % 4;
            if constexpr (shell == 0)
                zs += Real(pyrostencil::kLink[tp][0]) * x + Real(pyrostencil::kLink[tp][1]) * y +
                      Real(pyrostencil::kLink[tp][2]) * z;
            if constexpr (SPLIT > 0 && n.sl % SPLIT < SL % SPLIT) {
                low->s[shell][0] += x;
                low->s[shell][1] += y;
                low->s[shell][2] += z;
                if constexpr (shell == 0) {
                    low->v[tp][0] += x;
                    low->v[tp][1] += y;
                    low->v[tp][2] += z;
                }
            }
        }
    });
    hx = J.J1 * s[0][0];
    hy = J.J1 * s[0][1];
    hz = J.J1 * s[0][2];
    if constexpr (SHELLS >= 2) {
        hx += J.J2 * s[1][0];
        hy += J.J2 * s[1][1];
        hz += J.J2 * s[1][2];
    }
    if constexpr (SHELLS >= 3) {
        hx += J.J3 * s[2][0];
        hy += J.J3 * s[2][1];
        hz += J.J3 * s[2][2];
    }
    if (J.xxz) {
        constexpr int p = SL % 4;
        const Real t = J.Jx * zs;
        hx += Real(pyrostencil::kLink[p][0]) * t;
        hy += Real(pyrostencil::kLink[p][1]) * t;
        hz += Real(pyrostencil::kLink[p][2]) * t;
    }
}

// S . h_low for the spin S = (x, y, z) of a site of pyro_sl p and the sums
// of LowSums: J_shell S . s[shell], plus for XXZ Jx (z_p . S) sum_tp
// z_tp . v[tp], written as sum_tp ((z_p . S) z_tp) . v[tp] so that none of
// its products is one the field computes.  Explicit fma throughout, so every
// kernel that inlines it (per-colour or resident) rounds alike.
template <typename Real, int SHELLS>
__device__ __forceinline__ Real low_bond_energy(const LowSums<Real>& lo, const PyroCouplings<Real>& J, int p,
                                                Real x, Real y, Real z) {
    Real e = J.J1 * dot3(x, y, z, lo.s[0][0], lo.s[0][1], lo.s[0][2]);
    if constexpr (SHELLS >= 2) e = fma(J.J2, dot3(x, y, z, lo.s[1][0], lo.s[1][1], lo.s[1][2]), e);
    if constexpr (SHELLS >= 3) e = fma(J.J3, dot3(x, y, z, lo.s[2][0], lo.s[2][1], lo.s[2][2]), e);
    if (J.xxz) {
        const Real a = dot3(x, y, z, Real(pyrostencil::kLink[p][0]), Real(pyrostencil::kLink[p][1]),
                            Real(pyrostencil::kLink[p][2]));
        Real w = Real(0);
#pragma unroll
        for (int tp = 0; tp < 4; ++tp)
            w += dot3(a * Real(pyrostencil::kLink[tp][0]), a * Real(pyrostencil::kLink[tp][1]),
                      a * Real(pyrostencil::kLink[tp][2]), lo.v[tp][0], lo.v[tp][1], lo.v[tp][2]);
        e = fma(J.Jx, w, e);
    }
    return e;
}

// Energy term of a site updated in a pass that also sums the energy:
// S . h_low - B . S in FP64 (see pyro_colour_kernel).
template <typename Real, int SHELLS>
__device__ __forceinline__ double low_site_energy(const LowSums<Real>& lo, const PyroCouplings<Real>& J, int p,
                                                  const ReplicaParams<Real>& rp, Real x, Real y, Real z) {
    const Real bond = low_bond_energy<Real, SHELLS>(lo, J, p, x, y, z);
    const Real bs = dot3(x, y, z, rp.Bx, rp.By, rp.Bz);
    return double(bond) - double(bs);
}

// A replica's spins in global memory (Sr = S + r * 3 N) read through the
// caches as usual: load(i) is site i's record.
template <typename Real>
struct GlobalSpins {
    const Real* __restrict__ Sr;
    __device__ __forceinline__ Vec3<Real> operator()(unsigned i) const { return ld3(Sr, i); }
};

// Field from spins in global memory read by `load` (GlobalSpins, or a
// loader that bypasses the L1), neighbour cells from the wrapped axes, one
// 3-component load per neighbour.
template <typename Real, int SHELLS, int SL, int SPLIT = 0, typename Load>
__device__ __forceinline__ void pyro_field_l(const Load& load, int nc, const CellAxes& c,
                                             const PyroCouplings<Real>& J, Real& hx, Real& hy, Real& hz,
                                             LowSums<Real>* low = nullptr) {
    pyro_field_with<Real, SHELLS, SL, false, SPLIT>(
        [&](const pyrostencil::Nbr& n, Real& x, Real& y, Real& z) __attribute__((always_inline)) {
            const unsigned j = unsigned(n.sl * nc + c.a0[n.d[0] + 1] + c.a1[n.d[1] + 1] + c.a2[n.d[2] + 1]);
            const Vec3<Real> v = load(j);
            x = v.x;
            y = v.y;
            z = v.z;
        },
        J, hx, hy, hz, low);
}

// Field from a tile in LDS laid out [sl][box cell] with sublattice stride
// LS, every sublattice's box B0 x B1 x B2 = (T0 + 1) x (T1 + 1) x (T2 + 1)
// cells shifted down one cell on the axes where its halo is on the lower
// side (pyrostencil::halo_lo); lc = (i0 * B1 + i1) * B2 + i2 for the site's
// tile cell (i0, i1, i2).
template <typename Real, int SHELLS, int SL, int B1, int B2, int LS, bool FWD = false>
__device__ __forceinline__ void pyro_field_lds(const Real* lx, const Real* ly, const Real* lz, int lc,
                                               const PyroCouplings<Real>& J, Real& hx, Real& hy, Real& hz) {
    pyro_field_with<Real, SHELLS, SL, FWD>(
        [&](const pyrostencil::Nbr& n, Real& x, Real& y, Real& z) __attribute__((always_inline)) {
            using pyrostencil::halo_lo;
            const int j = lc + (n.sl * LS + (n.d[0] + halo_lo(n.sl, 0)) * B1 * B2 +
                                (n.d[1] + halo_lo(n.sl, 1)) * B2 + (n.d[2] + halo_lo(n.sl, 2)));
            x = lx[j];
            y = ly[j];
            z = lz[j];
        },
        J, hx, hy, hz);
}

// Runtime sublattice -> compile-time stencil, spins read by `load`.
template <typename Real, int SHELLS, int SPLIT = 0, typename Load>
__device__ __forceinline__ void pyro_field_rt_l(int sl, const Load& load, int nc, const CellAxes& c,
                                                const PyroCouplings<Real>& J, Real& hx, Real& hy, Real& hz,
                                                LowSums<Real>* low = nullptr) {
    switch (sl) {
#define PYRO_SL_CASE(X) \
    case X: pyro_field_l<Real, SHELLS, X, SPLIT>(load, nc, c, J, hx, hy, hz, low); break;
        PYRO_SL_CASE(0) PYRO_SL_CASE(1) PYRO_SL_CASE(2) PYRO_SL_CASE(3)
        PYRO_SL_CASE(4) PYRO_SL_CASE(5) PYRO_SL_CASE(6) PYRO_SL_CASE(7)
        PYRO_SL_CASE(8) PYRO_SL_CASE(9) PYRO_SL_CASE(10) PYRO_SL_CASE(11)
        PYRO_SL_CASE(12) PYRO_SL_CASE(13) PYRO_SL_CASE(14) PYRO_SL_CASE(15)
#undef PYRO_SL_CASE
        default: hx = hy = hz = Real(0);
    }
}
// pyro_field_rt_l with every neighbour record loaded before the sums, so
// that all 30 loads are in flight at once whatever the scheduler would make
// of the interleaved loads and sums.
template <typename Real, int SHELLS, int SPLIT = 0, typename Load>
__device__ __forceinline__ void pyro_field_rt_gathered(int sl, const Load& load, int nc, const CellAxes& c,
                                                       const PyroCouplings<Real>& J, Real& hx, Real& hy, Real& hz,
                                                       LowSums<Real>* low = nullptr) {
    constexpr int NS = pyrostencil::slots_for(SHELLS);
    switch (sl) {
#define PYRO_SL_CASE(X)                                                                                          \
    case X: {                                                                                                    \
        Vec3<Real> v[NS];                                                                                        \
        static_for<NS>([&](auto kk) __attribute__((always_inline)) {                                            \
            constexpr pyrostencil::Nbr n = kStencil.n[X][decltype(kk)::value];                                  \
            v[decltype(kk)::value] =                                                                             \
                load(unsigned(n.sl * nc + c.a0[n.d[0] + 1] + c.a1[n.d[1] + 1] + c.a2[n.d[2] + 1]));             \
        });                                                                                                      \
        int k = 0;                                                                                               \
        pyro_field_with<Real, SHELLS, X, false, SPLIT>(                                                         \
            [&](const pyrostencil::Nbr&, Real& x, Real& y, Real& z) __attribute__((always_inline)) {            \
                x = v[k].x;                                                                                      \
                y = v[k].y;                                                                                      \
                z = v[k].z;                                                                                      \
                ++k;                                                                                             \
            },                                                                                                   \
            J, hx, hy, hz, low);                                                                                 \
        break;                                                                                                   \
    }
        PYRO_SL_CASE(0) PYRO_SL_CASE(1) PYRO_SL_CASE(2) PYRO_SL_CASE(3)
        PYRO_SL_CASE(4) PYRO_SL_CASE(5) PYRO_SL_CASE(6) PYRO_SL_CASE(7)
        PYRO_SL_CASE(8) PYRO_SL_CASE(9) PYRO_SL_CASE(10) PYRO_SL_CASE(11)
        PYRO_SL_CASE(12) PYRO_SL_CASE(13) PYRO_SL_CASE(14) PYRO_SL_CASE(15)
#undef PYRO_SL_CASE
        default: hx = hy = hz = Real(0);
    }
}
// pyro_field_rt_gathered with its loads split around wait(), for a colour
// stage that must wait for the stage before it: the neighbours outside the
// previous stage's colour (final since an earlier stage) are loaded first,
// then wait() runs (false: give up, and so does this), then the neighbours
// of that colour.  The sums then follow in the stencil's order, so the field
// equals pyro_field_rt_gathered's bit for bit.  Stages run the colours in
// order, so the previous stage's colour is colour(sl) - 1 (mod NCOL), with
// colour(s) = s for 16 colours and s % 4 for 4.
// The SPLIT sums of pyro_field_with for sublattice SL (its neighbours of
// earlier colours, per shell and for shell 0 per target pyro_sl) from the
// neighbours' values v[k] in stencil order: the same additions in the same
// order, from zero.
template <typename Real, int SHELLS, int SL, int SPLIT>
__device__ __forceinline__ void low_sums(const Vec3<Real>* v, LowSums<Real>& low) {
    constexpr int NS = pyrostencil::slots_for(SHELLS);
    low = LowSums<Real>{};
    static_for<NS>([&](auto kk) __attribute__((always_inline)) {
        constexpr int k = decltype(kk)::value;
        constexpr pyrostencil::Nbr n = kStencil.n[SL][k];
        constexpr int shell = k < 6 ? 0 : (k < 18 ? 1 : 2);
        if constexpr (n.sl % SPLIT < SL % SPLIT) {
            low.s[shell][0] += v[k].x;
            low.s[shell][1] += v[k].y;
            low.s[shell][2] += v[k].z;
            if constexpr (shell == 0) {
                constexpr int tp = n.sl % 4;
                low.v[tp][0] += v[k].x;
                low.v[tp][1] += v[k].y;
                low.v[tp][2] += v[k].z;
            }
        }
    });
}

template <typename Real, int SHELLS, int SPLIT = 0, typename Load, typename Wait>
__device__ __forceinline__ bool pyro_field_rt_staged(int sl, const Load& load, int nc, const CellAxes& c,
                                                     const PyroCouplings<Real>& J, Wait&& wait, Real& hx, Real& hy,
                                                     Real& hz, LowSums<Real>* low = nullptr) {
    constexpr int NS = pyrostencil::slots_for(SHELLS);
    constexpr int NCOL = pyrostencil::colours_for(SHELLS);
    static_assert(NS <= 32, "one late-load bit per neighbour");
    // the neighbours' site indices and which of them have the previous
    // stage's colour (the stencil of sublattice sl, looked up once)
    unsigned idx[NS];
    unsigned late = 0;
    switch (sl) {
#define PYRO_SL_CASE(X)                                                                                          \
    case X:                                                                                                      \
        static_for<NS>([&](auto kk) __attribute__((always_inline)) {                                            \
            constexpr pyrostencil::Nbr n = kStencil.n[X][decltype(kk)::value];                                  \
            idx[decltype(kk)::value] = unsigned(n.sl * nc + c.a0[n.d[0] + 1] + c.a1[n.d[1] + 1] + c.a2[n.d[2] + 1]); \
            if constexpr (n.sl % NCOL == (X % NCOL + NCOL - 1) % NCOL) late |= 1u << decltype(kk)::value;        \
        });                                                                                                      \
        break;
        PYRO_SL_CASE(0) PYRO_SL_CASE(1) PYRO_SL_CASE(2) PYRO_SL_CASE(3)
        PYRO_SL_CASE(4) PYRO_SL_CASE(5) PYRO_SL_CASE(6) PYRO_SL_CASE(7)
        PYRO_SL_CASE(8) PYRO_SL_CASE(9) PYRO_SL_CASE(10) PYRO_SL_CASE(11)
        PYRO_SL_CASE(12) PYRO_SL_CASE(13) PYRO_SL_CASE(14) PYRO_SL_CASE(15)
#undef PYRO_SL_CASE
        default:
            static_for<NS>([&](auto kk) __attribute__((always_inline)) { idx[decltype(kk)::value] = 0u; });
    }
    Vec3<Real> v[NS];
    static_for<NS>([&](auto kk) __attribute__((always_inline)) {
        if (!((late >> decltype(kk)::value) & 1u)) v[decltype(kk)::value] = load(idx[decltype(kk)::value]);
    });
    // the early loads stay ahead of the wait: the compiler may not move
    // memory accesses across this, and nothing here waits for them
    asm volatile("" ::: "memory");
    if (!wait()) return false;
    static_for<NS>([&](auto kk) __attribute__((always_inline)) {
        if ((late >> decltype(kk)::value) & 1u) v[decltype(kk)::value] = load(idx[decltype(kk)::value]);
    });
    // the field, then (SPLIT) the sums over the earlier colours' neighbours
    // in a pass of their own over the loaded values: the same additions in
    // the same order as pyro_field_with's SPLIT sums, kept apart from the
    // field's so the field and the update do not wait for them
    switch (sl) {
#define PYRO_SL_CASE(X)                                                                                          \
    case X: {                                                                                                    \
        int k = 0;                                                                                               \
        pyro_field_with<Real, SHELLS, X, false, 0>(                                                             \
            [&](const pyrostencil::Nbr&, Real& x, Real& y, Real& z) __attribute__((always_inline)) {            \
                x = v[k].x;                                                                                      \
                y = v[k].y;                                                                                      \
                z = v[k].z;                                                                                      \
                ++k;                                                                                             \
            },                                                                                                   \
            J, hx, hy, hz);                                                                                      \
        if constexpr (SPLIT > 0) low_sums<Real, SHELLS, X, SPLIT>(v, *low);                                      \
        break;                                                                                                   \
    }
        PYRO_SL_CASE(0) PYRO_SL_CASE(1) PYRO_SL_CASE(2) PYRO_SL_CASE(3)
        PYRO_SL_CASE(4) PYRO_SL_CASE(5) PYRO_SL_CASE(6) PYRO_SL_CASE(7)
        PYRO_SL_CASE(8) PYRO_SL_CASE(9) PYRO_SL_CASE(10) PYRO_SL_CASE(11)
        PYRO_SL_CASE(12) PYRO_SL_CASE(13) PYRO_SL_CASE(14) PYRO_SL_CASE(15)
#undef PYRO_SL_CASE
        default: hx = hy = hz = Real(0);
    }
    return true;
}

// The same with the spins of replica block Sr read through the caches.
template <typename Real, int SHELLS, int SPLIT = 0>
__device__ __forceinline__ void pyro_field_rt(int sl, const Real* Sr, int nc, const CellAxes& c,
                                              const PyroCouplings<Real>& J, Real& hx, Real& hy, Real& hz,
                                              LowSums<Real>* low = nullptr) {
    pyro_field_rt_l<Real, SHELLS, SPLIT>(sl, GlobalSpins<Real>{Sr}, nc, c, J, hx, hy, hz, low);
}

// Sum of v over the 64 lanes of a wave (every lane active), the same in
// every lane, in a fixed order: an inclusive scan within each row of 16
// lanes by DPP row shifts (lanes shifted in from outside the row read 0),
// then the four row totals (lanes 15, 31, 47, 63) added in order.  Short
// latency next to a shuffle tree, which makes each step an LDS-crossbar
// round trip.
__device__ __forceinline__ double wave_sum_f64(double v) {
    auto shr = [](double x, auto k) __attribute__((always_inline)) {
        constexpr int ctrl = 0x110 + decltype(k)::value;  // DPP row_shr:k
        const long long b = __double_as_longlong(x);
        const int lo = __builtin_amdgcn_update_dpp(0, int(b), ctrl, 0xF, 0xF, true);
        const int hi = __builtin_amdgcn_update_dpp(0, int(b >> 32), ctrl, 0xF, 0xF, true);
        return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
    };
    v += shr(v, std::integral_constant<int, 1>{});
    v += shr(v, std::integral_constant<int, 2>{});
    v += shr(v, std::integral_constant<int, 4>{});
    v += shr(v, std::integral_constant<int, 8>{});
    const long long b = __double_as_longlong(v);
    const int lo = int(b), hi = int(b >> 32);
    auto row = [&](int lane) __attribute__((always_inline)) {
        const int l = __builtin_amdgcn_readlane(lo, lane), h = __builtin_amdgcn_readlane(hi, lane);
        return __longlong_as_double((static_cast<long long>(h) << 32) | static_cast<unsigned int>(l));
    };
    return ((row(15) + row(31)) + row(47)) + row(63);
}

// One launch per colour; grid.x = (sublattices of the colour) x (cell
// blocks), grid.y = replicas.  PER_REP: per-replica couplings (a batch of
// models); the single-model instance carries no branch on them, which
// would otherwise sit ahead of the neighbour gathers.
// EN (the last pass of a tempering round or anneal step): also the energy terms (MC.cpp:305-312) of
// the sites it updates, S_i . h_i over the bonds to sites of earlier colours
// (already final in this pass) minus B . S_i, with S_i the updated spin.
// Over the pass's colours this counts every bond once, when its later
// endpoint is visited, so the partials sum to the energy of the state the
// pass leaves.  One partial per (one-wave) block: epart[r * e_stride + bx].
// FUSED (the fused pass order): each site takes its KIND update and then
// its first OR update back to back, from one evaluation of its field (no
// bonded site changes in between: a colour class holds no bond), drawing
// the OR's uniform as uniform kUniformsOf[KIND] of its sweep (component
// j % 5 of the block with pass index pass + j / 5).  ucomp: an OR pass of
// the fused order that runs on its own takes component ucomp of its block.
constexpr int kPyroColourThreads = 64;

template <typename Real, int SHELLS, int KIND, bool PER_REP, int BS, bool EN, bool FUSED>
__global__ __launch_bounds__(BS) void pyro_colour_kernel(DeviceState<Real> st,
                                                          PyroCouplings<Real> J, int colour,
                                                          unsigned long long pass_index, int ucomp,
                                                          double* __restrict__ epart, int e_stride) {
    constexpr int NCOL = pyrostencil::colours_for(SHELLS);
    static_assert(!EN || BS == 64, "one energy partial per wave");
    static_assert(!FUSED || KIND != 3, "the fused launch starts with the sweep's first kind");
    const int nc = st.n_cells;
    const int bps = (nc + BS - 1) / BS;
    unsigned bx, ru;
    xcd_split(blockIdx.x, gridDim.x / unsigned(st.R), unsigned(st.R), bx, ru);
    const int q = int(bx) / bps;
    const int cell = (int(bx) % bps) * BS + threadIdx.x;
    const int sl = NCOL == 16 ? colour : colour + 4 * q;
    const int r = int(ru);
    const unsigned long long pass = pass_index + (st.pass_base ? *st.pass_base : 0ull);
    bool accepted = false;
    double e = 0.0;
    if (cell < nc) {
        const size_t N = size_t(st.N);
        Real* __restrict__ Sr = st.S + size_t(r) * 3 * N;
        const CellAxes c = cell_axes_of(cell, st.L0, st.L1, st.L2);
        const PyroCouplings<Real> Jr = PER_REP ? J.at(r) : J;
        Real hx, hy, hz;
        LowSums<Real> lo;
        if constexpr (EN) pyro_field_rt<Real, SHELLS, NCOL>(sl, Sr, nc, c, Jr, hx, hy, hz, &lo);
        else pyro_field_rt<Real, SHELLS>(sl, Sr, nc, c, Jr, hx, hy, hz);
        const ReplicaParams<Real> p = replica_params(st, r);
        const int site = sl * nc + cell;
        const Vec3<Real> v = ld3(Sr, unsigned(site));
        Real x = v.x, y = v.y, z = v.z;
        Uniform5 ub = site_uniforms(st, site, r, pass);
        if constexpr (KIND == 3) ub.m[0] = uniform_at(ub, ucomp);
        bool changed;
        accepted = update_spin_u<Real, KIND>(p, x, y, z, changed, ub, hx, hy, hz, st.dir + size_t(r) * N + site);
        if constexpr (FUSED) {
            constexpr int j = kUniformsOf[KIND];  // the sweep's uniform of its first OR update
            if constexpr (j % 5 == 0) ub = site_uniforms(st, site, r, pass + unsigned(j / 5));
            Uniform5 uo;
            uo.m[0] = ub.m[j % 5];
            bool ch;
            update_spin_u<Real, 3>(p, x, y, z, ch, uo, hx, hy, hz, nullptr);
            changed = changed || ch;
        }
        if (changed) st3(Sr, unsigned(site), x, y, z);
        if constexpr (EN) e = low_site_energy<Real, SHELLS>(lo, Jr, sl % 4, p, x, y, z);
    }
    if constexpr (KIND != 3) count_accepts(st, r, bx, accepted);
    if constexpr (EN) {
        const double t = wave_sum_f64(e);
        if (threadIdx.x == 0) epart[size_t(r) * unsigned(e_stride) + bx] = t;
    }
}

// Energy terms per site (MC.cpp:305-312) with the stencil: S_i . h_i over
// the bonds to higher sublattices (each bond once, the same sum as
// 1/2 sum_i S_i . h_i with half the gathers) minus B . S_i; one partial sum
// per 64-site (one-wave) block and replica, partial[r][block].  1-D grid of
// blocks x R with XCD-aware replica placement (xcd_split), so a replica's
// neighbour gathers are served by one L2.
template <typename Real, int SHELLS>
__global__ __launch_bounds__(64) void pyro_energy_kernel(DeviceState<Real> st, PyroCouplings<Real> J,
                                                         double* __restrict__ partial) {
    const unsigned gx = gridDim.x / unsigned(st.R);
    unsigned bx, ru;
    xcd_split(blockIdx.x, gx, unsigned(st.R), bx, ru);
    const int i = int(bx) * 64 + threadIdx.x;
    const int r = int(ru);
    double e = 0.0;
    if (i < st.N) {
        const int nc = st.n_cells;
        const size_t N = size_t(st.N);
        const Real* Sr = st.S + size_t(r) * 3 * N;
        const int sl = i / nc;
        const CellAxes c = cell_axes_of(i - sl * nc, st.L0, st.L1, st.L2);
        const PyroCouplings<Real> Jr = J.at(r);
        auto fetch = [&](const pyrostencil::Nbr& n, Real& x, Real& y, Real& z) __attribute__((always_inline)) {
            const unsigned j = unsigned(n.sl * nc + c.a0[n.d[0] + 1] + c.a1[n.d[1] + 1] + c.a2[n.d[2] + 1]);
            const Vec3<Real> v = ld3(Sr, j);
            x = v.x;
            y = v.y;
            z = v.z;
        };
        Real hx = Real(0), hy = Real(0), hz = Real(0);
        switch (sl) {
#define PYRO_SL_CASE(X) \
    case X: pyro_field_with<Real, SHELLS, X, true>(fetch, Jr, hx, hy, hz); break;
            PYRO_SL_CASE(0) PYRO_SL_CASE(1) PYRO_SL_CASE(2) PYRO_SL_CASE(3)
            PYRO_SL_CASE(4) PYRO_SL_CASE(5) PYRO_SL_CASE(6) PYRO_SL_CASE(7)
            PYRO_SL_CASE(8) PYRO_SL_CASE(9) PYRO_SL_CASE(10) PYRO_SL_CASE(11)
            PYRO_SL_CASE(12) PYRO_SL_CASE(13) PYRO_SL_CASE(14) PYRO_SL_CASE(15)
#undef PYRO_SL_CASE
            default: break;
        }
        const Vec3<Real> si = ld3(Sr, unsigned(i));
        const Real sx = si.x, sy = si.y, sz = si.z;
        const Real bond = sx * hx + sy * hy + sz * hz;
        const Real bs = sx * st.B[3 * r] + sy * st.B[3 * r + 1] + sz * st.B[3 * r + 2];
        e = double(bond) - double(bs);
    }
    for (int off = 32; off > 0; off >>= 1) e += __shfl_down(e, off, 64);
    if (threadIdx.x == 0) partial[size_t(r) * gx + bx] = e;
}

// Local field of every site of one replica (MC.cpp:87-96), out[3][N].
template <typename Real, int SHELLS>
__global__ __launch_bounds__(256) void pyro_fields_kernel(DeviceState<Real> st, PyroCouplings<Real> J,
                                                          int r, Real* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= st.N) return;
    const int nc = st.n_cells;
    const size_t N = size_t(st.N);
    const Real* Sr = st.S + size_t(r) * 3 * N;
    const int sl = i / nc;
    const CellAxes c = cell_axes_of(i - sl * nc, st.L0, st.L1, st.L2);
    Real hx, hy, hz;
    pyro_field_rt<Real, SHELLS>(sl, Sr, nc, c, J.at(r), hx, hy, hz);
    out[i] = hx;
    out[N + i] = hy;
    out[2 * N + i] = hz;
}

// The update of one site by a resident workgroup, with the draws and
// arithmetic of pyro_colour_kernel: KIND from block `pass` (an OR pass
// alone: component ucomp of it) and with FUSED (n_or >= 1) the first OR
// update after it from the sweep's uniform kUniformsOf[KIND]; the spin
// (x, y, z) is updated and stored when it changed.  Returns whether the
// KIND proposal was accepted.
template <typename Real, int KIND, bool FUSED>
__device__ __forceinline__ bool resident_update(const ReplicaParams<Real>& p, const DeviceState<Real>& st, Real& x,
                                                Real& y, Real& z, int site, int r, unsigned long long pass, int ucomp,
                                                int n_or, Real hx, Real hy, Real hz, Real* __restrict__ Sr,
                                                size_t N) {
    Uniform5 ub = site_uniforms(st, site, r, pass);
    if constexpr (KIND == 3) ub.m[0] = uniform_at(ub, ucomp);
    bool changed;
    const bool acc = update_spin_u<Real, KIND>(p, x, y, z, changed, ub, hx, hy, hz, st.dir + size_t(r) * N + site);
    if constexpr (FUSED && KIND != 3) {
        if (n_or >= 1) {
            constexpr int j = kUniformsOf[KIND];
            if constexpr (j % 5 == 0) ub = site_uniforms(st, site, r, pass + unsigned(j / 5));
            Uniform5 uo;
            uo.m[0] = ub.m[j % 5];
            bool ch;
            update_spin_u<Real, 3>(p, x, y, z, ch, uo, hx, hy, hz, nullptr);
            changed = changed || ch;
        }
    }
    if (changed) st3(Sr, unsigned(site), x, y, z);
    return acc;
}

// One workgroup per replica; thread t owns cells t, t + BS, ... (CPT cells).
// EN: the last pass of the last sweep also leaves the energy partials of
// pyro_colour_kernel<..., EN>, at the same indices and with the same
// per-site arithmetic and wave sums: wave w of the workgroup covers cells
// 64 w .. 64 w + 63 + j BS, which is block j (BS / 64) + w of the per-colour
// launch, so the two strategies give bit-identical energies.
// FUSED: the fused pass order of pyro_colour_kernel<..., FUSED> (stage
// group 0: each site's FIRST update and its first OR update; group g >= 1:
// OR pass g + 1, drawing uniform kUniformsOf[FIRST] + g of the sweep), with
// the same draws and arithmetic, so it too is bit-identical to the
// per-colour launches.
template <typename Real, int SHELLS, int FIRST, int BS, int CPT, bool EN, bool FUSED>
__global__ __launch_bounds__(BS) void pyro_resident_kernel(DeviceState<Real> st,
                                                           PyroCouplings<Real> J, int n_or,
                                                           int n_sweeps,
                                                           unsigned long long pass_index,
                                                           double* __restrict__ epart) {
    constexpr int NCOL = pyrostencil::colours_for(SHELLS);
    constexpr int SPC = 16 / NCOL;
    const int r = blockIdx.x;
    const PyroCouplings<Real> Jr = J.at(r);
    const int nc = st.n_cells;
    const size_t N = size_t(st.N);
    Real* __restrict__ Sr = st.S + size_t(r) * 3 * N;
    const unsigned long long pass_start = pass_index + (st.pass_base ? *st.pass_base : 0ull);
    const ReplicaParams<Real> rp = replica_params(st, r);
    unsigned int my_acc = 0;
    static_assert(CPT == 2, "the resident kernel owns two cells per thread");
    const int cell_b = threadIdx.x + BS < nc ? threadIdx.x + BS : 0;
    const CellAxes ax_a = cell_axes_of(threadIdx.x < nc ? threadIdx.x : 0, st.L0, st.L1, st.L2);
    const CellAxes ax_b = cell_axes_of(cell_b, st.L0, st.L1, st.L2);
    const int bps = (nc + 63) / 64;
    const size_t es = size_t(pyro_colour_energy_blocks(nc));
    // stage groups per sweep: every pass alone, or (fused) the first kind
    // with the first OR update, then the further OR passes
    const int n_groups = FUSED ? (n_or > 1 ? n_or : 1) : n_or + 1;
    for (int sw = 0; sw < n_sweeps; ++sw) {
        const unsigned long long p0 = pass_start + (unsigned long long)sw * unsigned(1 + n_or);
        for (int grp = 0; grp < n_groups; ++grp) {
            const bool en_pass = EN && sw == n_sweeps - 1 && grp == n_groups - 1;  // uniform
            // this group's block and, for an OR pass alone, the component
            // of it that the pass draws
            int ucomp = 0;
            unsigned long long pass = p0 + unsigned(grp);
            if (FUSED && grp > 0) {
                const int j = kUniformsOf[FIRST] + grp;
                pass = p0 + unsigned(j / 5);
                ucomp = j % 5;
            }
            for (int col = 0; col < NCOL; ++col) {
                for (int q = 0; q < SPC; ++q) {
                    const int sl = NCOL == 16 ? col : col + 4 * q;
                    // one cell at a time: the 16 waves of the workgroup hide
                    // the gather latency, and registers stay within 128
#pragma unroll 1
                    for (int j = 0; j < CPT; ++j) {
                        const int cell = threadIdx.x + j * BS;
                        CellAxes c;
#pragma unroll
                        for (int a = 0; a < 3; ++a) {
                            c.a0[a] = j ? ax_b.a0[a] : ax_a.a0[a];
                            c.a1[a] = j ? ax_b.a1[a] : ax_a.a1[a];
                            c.a2[a] = j ? ax_b.a2[a] : ax_a.a2[a];
                        }
                        if constexpr (EN) {
                            if (en_pass) {
                                // every lane reaches the wave sum below
                                double e = 0.0;
                                if (cell < nc) {
                                    LowSums<Real> lo;
                                    Real hx, hy, hz;
                                    pyro_field_rt<Real, SHELLS, NCOL>(sl, Sr, nc, c, Jr, hx, hy, hz, &lo);
                                    const int site = sl * nc + cell;
                                    const Vec3<Real> v = ld3(Sr, unsigned(site));
                                    Real x = v.x, y = v.y, z = v.z;
                                    if (grp == 0 && FIRST != 3)
                                        my_acc += resident_update<Real, FIRST, FUSED>(rp, st, x, y, z, site, r, pass,
                                                                                     0, n_or, hx, hy, hz, Sr, N)
                                                      ? 1u : 0u;
                                    else
                                        resident_update<Real, 3, false>(rp, st, x, y, z, site, r, pass, ucomp, n_or, hx,
                                                                        hy, hz, Sr, N);
                                    e = low_site_energy<Real, SHELLS>(lo, Jr, sl % 4, rp, x, y, z);
                                }
                                const double t = wave_sum_f64(e);
                                const int cb = j * (BS / 64) + int(threadIdx.x >> 6);
                                if ((threadIdx.x & 63) == 0 && cb < bps)
                                    epart[size_t(r) * es + size_t(col * SPC + q) * unsigned(bps) + unsigned(cb)] = t;
                                continue;
                            }
                        }
                        if (cell >= nc) break;
                        Real hx, hy, hz;
                        pyro_field_rt<Real, SHELLS>(sl, Sr, nc, c, Jr, hx, hy, hz);
                        const int site = sl * nc + cell;
                        const Vec3<Real> v = ld3(Sr, unsigned(site));
                        Real x = v.x, y = v.y, z = v.z;
                        if (grp == 0 && FIRST != 3)
                            my_acc += resident_update<Real, FIRST, FUSED>(rp, st, x, y, z, site, r, pass, 0, n_or, hx,
                                                                         hy, hz, Sr, N)
                                          ? 1u : 0u;
                        else
                            resident_update<Real, 3, false>(rp, st, x, y, z, site, r, pass, ucomp, n_or, hx, hy, hz,
                                                            Sr, N);
                    }
                }
                __syncthreads();
            }
        }
    }
    if (st.acc == nullptr) return;  // uniform: counts not requested
    __shared__ unsigned int red[BS / 64];
    unsigned int v = my_acc;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int i = 0; i < BS / 64; ++i) t += red[i];
        atomicAdd(st.acc + r, t);
    }
}

// Persistent per-colour sweeps held on one XCD per replica (strategy 6).
//
// Every launch of pyro_colour_kernel begins with the L2 writeback and
// invalidation that keep the eight XCDs' L2 caches coherent, so each of the
// 16 colour stages of a pass fetches the replica set again from beyond L2.
// Here one launch runs whole sweeps: the workgroups of replica r all run on
// XCD r % 8 (each workgroup reads its XCD from the XCC_ID register and
// takes a ticket among that XCD's workgroups) and wait for one another
// between colour stages on a per-replica counter, so a replica's spins stay
// in its XCD's L2 across the stages.
//
// Coherence between the CUs of an XCD goes through its L2, the point of
// coherence of those CUs.  A workgroup counts a finished stage with an
// atomic add after its spin stores have reached the L2 (a release, which
// waits for them; the unit is built with -mtgsplit, whose workgroup-scope
// release waits for stores to complete), and polls the count with an atomic
// OR of 0: atomics execute in the L2 of the issuing XCD, never in an L1.
// Every read of the spins in this kernel is a device-scope buffer load
// (cache policy sc1), which misses the CU's L1 and is served by the L2.
// Of a stage's loads only those of the previous stage's colour wait for
// that stage: the site's own spin and its other neighbours are final
// already, so they are issued before the wait and arrive during it
// (pyro_field_rt_staged).
// Measured alternatives: workgroup-scope loads of the counters or spins (sc0)
// and an L1 invalidation by a workgroup-scope acquire read stale values;
// a device-scope acquire (which also invalidates the L2) was about four
// times slower than the sc1 loads.
//
// Workgroup with ticket t on XCD x (one wave) -> replica x + 8 (t / bpr),
// block b = t % bpr of each colour stage, the same cells and block
// index as pyro_colour_kernel's block b (bpr = its blocks per replica), and
// the same per-site counters, draws and arithmetic (the pass-by-pass or the
// fused pass order of pyro_resident_kernel), so the results equal strategy
// 3's bit for bit.  Every wait is bounded: when the workgroup sees its
// replica's count later than `timeout_ticks` of the device's wall clock
// after it counted its own previous stage, it sets *err and returns, and
// the host reports the error instead of the results.
constexpr int kHwRegXccId = 20;  // hwreg(HW_REG_XCC_ID), bits 3:0 = the XCD

__device__ __forceinline__ int xcc_id() {
    return int(__builtin_amdgcn_s_getreg((31 << 11) | kHwRegXccId) & 0xFu);
}

// A replica's spins read with device-scope buffer loads (cache policy sc1:
// past the L1, from the L2).  The byte offset of site i is a 24-bit
// multiply (full rate, where a 32-bit one is quarter rate): the launches
// that use it hold every workgroup resident at once, which keeps a
// replica far below 2^24 sites (and the launchers check it).
constexpr int kBufferWord3 = 0x00020000;  // raw-buffer descriptor word 3 of gfx9, 32-bit data
constexpr int kPolicySc1 = 16;            // the sc1 bit of a buffer load's cache policy
constexpr int kXcdPad = 32;               // counters 128 bytes apart: one cache line each
template <typename Real>
struct L2Spins;
template <>
struct L2Spins<float> {
    __amdgpu_buffer_rsrc_t rs;
    __device__ __forceinline__ Vec3<float> operator()(unsigned i) const {
        const auto v = __builtin_amdgcn_raw_buffer_load_b96(rs, __umul24(i, 12u), 0, kPolicySc1);
        return Vec3<float>{__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2])};
    }
};
template <>
struct L2Spins<double> {
    __amdgpu_buffer_rsrc_t rs;
    __device__ __forceinline__ Vec3<double> operator()(unsigned i) const {
        const unsigned o = __umul24(i, 24u);
        const auto a = __builtin_amdgcn_raw_buffer_load_b64(rs, o, 0, kPolicySc1);
        const auto b = __builtin_amdgcn_raw_buffer_load_b64(rs, o + 8u, 0, kPolicySc1);
        const auto c = __builtin_amdgcn_raw_buffer_load_b64(rs, o + 16u, 0, kPolicySc1);
        return Vec3<double>{__hiloint2double(int(a[1]), int(a[0])), __hiloint2double(int(b[1]), int(b[0])),
                            __hiloint2double(int(c[1]), int(c[0]))};
    }
};

// Stage counter of replica r: wait until it reaches `target` (device-scope
// reads, past the L1); false (and *err set) when the count is seen only
// after `timeout` ticks of the wall clock from `arrived`, the time this
// workgroup counted its previous stage.
__device__ __forceinline__ bool xcd_wait(unsigned int* ctr, unsigned int target, int* err,
                                         unsigned long long arrived, unsigned long long timeout,
                                         unsigned int zero) {
    for (;;) {
        // an atomic read-modify-write executes in the L2 (never the L1);
        // `zero` (0, a kernel argument) keeps the compiler from folding an
        // OR with 0 into a plain load, which the L1 may serve stale
        const unsigned int v = __hip_atomic_fetch_or(ctr, zero, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (wall_clock64() - arrived > timeout) {
            atomicOr(err, 1);
            return false;
        }
        if (v >= target) return true;
        __builtin_amdgcn_s_sleep(1);
    }
}

// resident_update with the uniforms drawn beforehand: u0 the block of the
// KIND update (for an OR pass alone, its component already in m[0]) and u1
// the block holding the fused OR update's uniform.
template <typename Real, int KIND, bool FUSED>
__device__ __forceinline__ bool drawn_update(const ReplicaParams<Real>& p, const DeviceState<Real>& st, Real& x,
                                             Real& y, Real& z, int site, int r, int n_or, const Uniform5& u0,
                                             const Uniform5& u1, Real hx, Real hy, Real hz, Real* __restrict__ Sr,
                                             size_t N) {
    bool changed;
    const bool acc = update_spin_u<Real, KIND>(p, x, y, z, changed, u0, hx, hy, hz, st.dir + size_t(r) * N + site);
    if constexpr (FUSED && KIND != 3) {
        if (n_or >= 1) {
            constexpr int j = kUniformsOf[KIND];
            Uniform5 uo;
            uo.m[0] = (j % 5 == 0 ? u1 : u0).m[j % 5];
            bool ch;
            update_spin_u<Real, 3>(p, x, y, z, ch, uo, hx, hy, hz, nullptr);
            changed = changed || ch;
        }
    }
    if (changed) st3(Sr, unsigned(site), x, y, z);
    return acc;
}

__global__ void xcd_zero_kernel(unsigned int* __restrict__ p, int n) {
    for (int i = int(threadIdx.x); i < n; i += int(blockDim.x)) p[i] = 0u;
}

template <typename Real, int SHELLS, int FIRST, bool PER_REP, bool EN, bool FUSED>
__global__ __launch_bounds__(64) void pyro_xcd_kernel(DeviceState<Real> st, PyroCouplings<Real> J, int n_or,
                                                      int n_sweeps, unsigned long long pass_index,
                                                      double* __restrict__ epart, unsigned int* __restrict__ bar,
                                                      int* __restrict__ err, unsigned long long timeout_ticks,
                                                      unsigned int zero) {
    constexpr int NCOL = pyrostencil::colours_for(SHELLS);
    if (*err) return;  // an earlier launch failed: the host reports it at its next read
    // the XCD this workgroup runs on and its place among that XCD's
    // workgroups: a ticket from the XCD's own counter bar[R + x] (the
    // dispatcher hands consecutive workgroups to consecutive XCDs from a
    // starting XCD that varies from launch to launch, so each XCD receives
    // a whole share of the grid, but not always the same ids)
    const int x = xcc_id() & 7;
    unsigned int tk = 0;
    if (threadIdx.x == 0)
        tk = __hip_atomic_fetch_add(bar + size_t(st.R + x) * kXcdPad, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    const int t = int(__builtin_amdgcn_readfirstlane(tk));
    const int nc = st.n_cells;
    const int bps = (nc + 63) / 64;
    const int bpr = bps * (16 / NCOL);  // blocks per replica and colour stage
    const int r = x + 8 * (t / bpr);
    if (r >= st.R) return;  // no replica for this workgroup: it takes no part in any count
    const int b = t - (t / bpr) * bpr;
    const int q = b / bps;
    const int cell = (b - q * bps) * 64 + int(threadIdx.x);
    const bool live = cell < nc;
    const size_t N = size_t(st.N);
    Real* __restrict__ Sr = st.S + size_t(r) * 3 * N;
    const L2Spins<Real> ld{__builtin_amdgcn_make_buffer_rsrc(Sr, 0, int(3 * N * sizeof(Real)), kBufferWord3)};
    const CellAxes axes = cell_axes_of(live ? cell : 0, st.L0, st.L1, st.L2);
    const PyroCouplings<Real> Jr = PER_REP ? J.at(r) : J;
    const ReplicaParams<Real> rp = replica_params(st, r);
    const unsigned long long pass_start = pass_index + (st.pass_base ? *st.pass_base : 0ull);
    const int per_launch = bpr;
    const size_t es = size_t(pyro_colour_energy_blocks(nc));
    unsigned int* ctr = bar + size_t(r) * kXcdPad;
    const int n_groups = FUSED ? (n_or > 1 ? n_or : 1) : n_or + 1;
    unsigned int n_acc = 0;
    unsigned int stage = 0;
    unsigned long long arrived = 0;
    double part = 0.0;  // EN: lane col, the energy partial of colour col's last stage
    for (int sw = 0; sw < n_sweeps; ++sw) {
        const unsigned long long p0 = pass_start + (unsigned long long)sw * unsigned(1 + n_or);
        for (int grp = 0; grp < n_groups; ++grp) {
            const bool en_pass = EN && sw == n_sweeps - 1 && grp == n_groups - 1;  // uniform
            int ucomp = 0;
            unsigned long long pass = p0 + unsigned(grp);
            if (FUSED && grp > 0) {
                const int j = kUniformsOf[FIRST] + grp;
                pass = p0 + unsigned(j / 5);
                ucomp = j % 5;
            }
            for (int col = 0; col < NCOL; ++col, ++stage) {
                const int sl = NCOL == 16 ? col : col + 4 * q;
                const int site = sl * nc + cell;
                // this stage's uniforms, drawn before the wait: they depend
                // on the site and the pass alone, and the rounds then run
                // while the other workgroups finish the previous stage
                const bool first = grp == 0 && FIRST != 3;
                Uniform5 u0 = site_uniforms(st, site, r, pass), u1 = u0;
                if (first) {
                    if constexpr (FUSED && kUniformsOf[FIRST] % 5 == 0)
                        u1 = site_uniforms(st, site, r, pass + unsigned(kUniformsOf[FIRST] / 5));
                } else {
                    u0.m[0] = uniform_at(u0, ucomp);
                }
                asm volatile("" : "+v"(u0.m[0]), "+v"(u0.m[1]), "+v"(u0.m[2]), "+v"(u0.m[3]), "+v"(u0.m[4]),
                             "+v"(u1.m[0]));
                // the neighbour offsets from the cell's axes afresh in each
                // stage: hoisted out of the loop (16 stencils of 30) they
                // would take hundreds of registers; the empty asm hides the
                // axes' values from the compiler so it cannot hoist
                CellAxes c = axes;
#pragma unroll
                for (int k = 0; k < 3; ++k) asm volatile("" : "+v"(c.a0[k]), "+v"(c.a1[k]), "+v"(c.a2[k]));
                // the site's own spin and the neighbours outside the previous
                // stage's colour are final already: they are loaded before
                // the wait for that stage, the rest after it (lanes past the
                // last cell load the values of cell 0 and update nothing)
                const Vec3<Real> v = ld(unsigned(live ? site : sl * nc));
                auto wait = [&]() __attribute__((always_inline)) {
                    return stage == 0 || xcd_wait(ctr, stage * unsigned(bpr), err, arrived, timeout_ticks, zero);
                };
                // EN: every stage sums the earlier colours' neighbours and
                // the site's energy, used by the last group's stages only.
                // One code path for all stages: a path of their own would
                // run once per launch from a cold instruction cache, which
                // cost more (about 1.6 us per stage) than the sums
                Real hx, hy, hz;
                LowSums<Real> lo;
                const bool ok = pyro_field_rt_staged<Real, SHELLS, EN ? NCOL : 0>(sl, ld, nc, c, Jr, wait, hx, hy, hz,
                                                                                 EN ? &lo : nullptr);
                if (!ok) return;
                bool a = false;
                double e = 0.0;
                if (live) {
                    Real sx = v.x, sy = v.y, sz = v.z;
                    if (first)
                        a = drawn_update<Real, FIRST, FUSED>(rp, st, sx, sy, sz, site, r, n_or, u0, u1, hx, hy, hz, Sr,
                                                             N);
                    else
                        drawn_update<Real, 3, false>(rp, st, sx, sy, sz, site, r, n_or, u0, u1, hx, hy, hz, Sr, N);
                    if constexpr (EN) e = low_site_energy<Real, SHELLS>(lo, Jr, sl % 4, rp, sx, sy, sz);
                }
                if (grp == 0 && FIRST != 3 && st.acc_wave) n_acc += unsigned(__popcll(__ballot(a)));
                if (en_pass) {
                    const double tot = wave_sum_f64(e);
                    if (int(threadIdx.x) == col) part = tot;
                }
                // release: this wave's stores reach the L2 before its count
                if (threadIdx.x == 0)
                    __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                arrived = wall_clock64();
            }
        }
    }
    // the energy partials of the last group's stages, lane col holding
    // colour col's: stored here rather than in their stages, whose counts
    // would otherwise wait for these partial-line stores to complete
    if (EN && n_sweeps > 0 && int(threadIdx.x) < NCOL)
        epart[size_t(threadIdx.x) * unsigned(per_launch) + size_t(r) * es + unsigned(b)] = part;
    if (st.acc_wave && threadIdx.x == 0 && n_acc) st.acc_wave[size_t(r) * unsigned(st.acc_waves) + unsigned(b)] += n_acc;
}

// Whole sweeps of a batch of contexts (lattices of different sizes, each
// with its own replicas, temperatures and couplings) in one persistent
// launch: pyro_xcd_kernel's scheme over a work table instead of one
// context's replicas.  The table lists, per XCD, the workgroups of the
// replicas placed there: work[x * per_xcd + t] = (member, replica, worker,
// workers) for the workgroup holding ticket t on XCD x (member < 0: idle).
// Worker w of a replica's nw takes blocks w, w + nw, ... of each colour
// stage (the blocks of pyro_colour_kernel), and the replica's counter
// reaches stage * nw when the stage is done.  Each site sees the same
// counters, draws and arithmetic as in pyro_colour_kernel, so every
// member's results equal its own strategy-3 sweeps bit for bit.  A wait
// that runs past the member's bound sets the member's error flag.
template <typename Real, int SHELLS, int FIRST, bool PER_REP, bool EN, bool FUSED>
__global__ __launch_bounds__(64) void pyro_batch_kernel(const BatchMember<Real>* __restrict__ members,
                                                        const int4* __restrict__ work, int per_xcd, int n_items,
                                                        int n_or, int n_sweeps, unsigned int* __restrict__ bar,
                                                        unsigned int zero) {
    constexpr int NCOL = pyrostencil::colours_for(SHELLS);
    const int x = xcc_id() & 7;
    unsigned int tk = 0;
    if (threadIdx.x == 0)
        tk = __hip_atomic_fetch_add(bar + size_t(n_items + x) * kXcdPad, 1u, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_WORKGROUP);
    const int t = int(__builtin_amdgcn_readfirstlane(tk));
    if (t >= per_xcd) return;  // the dispatcher gives every XCD per_xcd workgroups: never taken
    const int4 wk = work[size_t(x) * unsigned(per_xcd) + unsigned(t)];
    if (wk.x < 0) return;  // an idle entry: no replica, no count
    const BatchMember<Real>& M = members[wk.x];
    int* const err = M.err;
    if (*err) return;  // an earlier launch of this member failed: its next read reports it
    const unsigned long long timeout_ticks = M.timeout;
    const DeviceState<Real> st = M.st;
    const int r = wk.y, w = wk.z, nw = wk.w;
    const int nc = st.n_cells;
    const int bps = (nc + 63) / 64;
    const int bpr = bps * (16 / NCOL);
    const size_t N = size_t(st.N);
    Real* __restrict__ Sr = st.S + size_t(r) * 3 * N;
    const L2Spins<Real> ld{__builtin_amdgcn_make_buffer_rsrc(Sr, 0, int(3 * N * sizeof(Real)), kBufferWord3)};
    const PyroCouplings<Real> Jr = PER_REP ? M.J.at(r) : M.J;
    const ReplicaParams<Real> rp = replica_params(st, r);
    const unsigned long long pass_start = M.pass_index + (st.pass_base ? *st.pass_base : 0ull);
    const size_t es = size_t(pyro_colour_energy_blocks(nc));
    unsigned int* ctr = bar + size_t(M.item0 + r) * kXcdPad;
    const int n_groups = FUSED ? (n_or > 1 ? n_or : 1) : n_or + 1;
    unsigned int n_acc = 0;
    unsigned int stage = 0;
    unsigned long long arrived = 0;
    double part = 0.0;  // EN: lane jb NCOL + col, block jb's energy partial of colour col
    for (int sw = 0; sw < n_sweeps; ++sw) {
        const unsigned long long p0 = pass_start + (unsigned long long)sw * unsigned(1 + n_or);
        for (int grp = 0; grp < n_groups; ++grp) {
            const bool en_pass = EN && sw == n_sweeps - 1 && grp == n_groups - 1;  // uniform
            int ucomp = 0;
            unsigned long long pass = p0 + unsigned(grp);
            if (FUSED && grp > 0) {
                const int j = kUniformsOf[FIRST] + grp;
                pass = p0 + unsigned(j / 5);
                ucomp = j % 5;
            }
            const bool first = grp == 0 && FIRST != 3;
            for (int col = 0; col < NCOL; ++col, ++stage) {
                if (stage > 0 && !xcd_wait(ctr, stage * unsigned(nw), err, arrived, timeout_ticks, zero)) return;
                for (int b = w, jb = 0; b < bpr; b += nw, ++jb) {
                    const int q = b / bps;
                    const int cell = (b - q * bps) * 64 + int(threadIdx.x);
                    const bool live = cell < nc;
                    const int sl = NCOL == 16 ? col : col + 4 * q;
                    const int site = sl * nc + cell;
                    bool a = false;
                    double e = 0.0;
                    if (live) {
                        Uniform5 u0 = site_uniforms(st, site, r, pass), u1 = u0;
                        if (first) {
                            if constexpr (FUSED && kUniformsOf[FIRST] % 5 == 0)
                                u1 = site_uniforms(st, site, r, pass + unsigned(kUniformsOf[FIRST] / 5));
                        } else {
                            u0.m[0] = uniform_at(u0, ucomp);
                        }
                        const CellAxes c = cell_axes_of(cell, st.L0, st.L1, st.L2);
                        Real hx, hy, hz;
                        LowSums<Real> lo;
                        // EN: the earlier colours' sums and the site energy in
                        // every stage, one code path (as in pyro_xcd_kernel)
                        pyro_field_rt_gathered<Real, SHELLS, EN ? NCOL : 0>(sl, ld, nc, c, Jr, hx, hy, hz,
                                                                            EN ? &lo : nullptr);
                        const Vec3<Real> v = ld(unsigned(site));
                        Real sx = v.x, sy = v.y, sz = v.z;
                        if (first)
                            a = drawn_update<Real, FIRST, FUSED>(rp, st, sx, sy, sz, site, r, n_or, u0, u1, hx, hy,
                                                                 hz, Sr, N);
                        else
                            drawn_update<Real, 3, false>(rp, st, sx, sy, sz, site, r, n_or, u0, u1, hx, hy, hz, Sr,
                                                         N);
                        if constexpr (EN) e = low_site_energy<Real, SHELLS>(lo, Jr, sl % 4, rp, sx, sy, sz);
                    }
                    if (first && st.acc_wave) n_acc += unsigned(__popcll(__ballot(a)));
                    if (en_pass) {
                        const double tot = wave_sum_f64(e);
                        const int slot = jb * NCOL + col;  // this block's partial of colour col
                        if (slot < 64) {
                            if (int(threadIdx.x) == slot) part = tot;
                        } else if (threadIdx.x == 0) {
                            M.epart[size_t(col) * unsigned(bpr) + size_t(r) * es + unsigned(b)] = tot;
                        }
                    }
                }
                // release: this wave's stores reach the L2 before its count
                if (threadIdx.x == 0) __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                arrived = wall_clock64();
            }
        }
    }
    // the energy partials held in lanes (block jb's colour col in lane
    // jb NCOL + col), stored after the last stage as in pyro_xcd_kernel
    if (EN && n_sweeps > 0) {
        const int jb = int(threadIdx.x) / NCOL, col = int(threadIdx.x) % NCOL;
        const int b = w + jb * nw;
        if (b < bpr) M.epart[size_t(col) * unsigned(bpr) + size_t(r) * es + unsigned(b)] = part;
    }
    if (st.acc_wave && threadIdx.x == 0 && n_acc) st.acc_wave[size_t(r) * unsigned(st.acc_waves) + unsigned(w)] += n_acc;
}

// Lane -> tile cell of the tiled sweep.  A colour stage reads, for every
// cell of the tile, the neighbour at the same constant offset from the
// cell's LDS word, so a wave's lanes conflict in an LDS bank when their
// cells' words agree modulo the bank count.  The halo padding of the box
// rows makes row-major lanes collide about 1.8-fold; here wave w's lane b
// takes an unassigned cell whose word is b modulo 64, so each wave reads 64
// distinct banks (and each half-wave 32 distinct banks of 32).  Cells left
// over (tiles with more than BS/64 cells per bank) fill idle lanes; idle
// lanes hold -1.
template <int T0, int T1, int T2, int BS>
struct TileLanes {
    short cell[BS];
    constexpr TileLanes() : cell{} {
        constexpr int NI = T0 * T1 * T2, B1 = T1 + 1, B2 = T2 + 1;
        bool used[NI] = {};
        for (int l = 0; l < BS; ++l) cell[l] = -1;
        for (int w = 0; w < BS / 64; ++w)
            for (int b = 0; b < 64; ++b)
                for (int m = 0; m < NI; ++m) {
                    const int i2 = m % T2, i1 = (m / T2) % T1, i0 = m / (T1 * T2);
                    if (!used[m] && ((i0 * B1 + i1) * B2 + i2) % 64 == b) {
                        cell[w * 64 + b] = short(m);
                        used[m] = true;
                        break;
                    }
                }
        for (int m = 0; m < NI; ++m)
            for (int l = 0; l < BS && !used[m]; ++l)
                if (cell[l] < 0) {
                    cell[l] = short(m);
                    used[m] = true;
                }
    }
};
template <int T0, int T1, int T2, int BS>
inline constexpr TileLanes<T0, T1, T2, BS> kTileLanes{};

// Copies the tile at cell (o0, o1, o2) with its one-sided halo into LDS,
// [sl][box cell] with sublattice stride LS: sublattice sl's box of
// (T0 + 1)(T1 + 1)(T2 + 1) cells is shifted down one cell on the axes where
// its halo is on the lower side (halo_lo(sl, a) = 1 iff R_fcc[sl / 4] has 4
// on axis a), coordinates wrapped periodically.
template <typename Real, int T0, int T1, int T2, int BS, int LS, int UNROLL = 1>
__device__ __forceinline__ void load_tile_box(Real* lx, Real* ly, Real* lz, const Real* __restrict__ Sr, int o0,
                                              int o1, int o2, int L0, int L1, int L2, int nc) {
    constexpr int B1 = T1 + 1, B2 = T2 + 1, BC = (T0 + 1) * B1 * B2;
    static_assert(LS >= BC, "the stride holds a sublattice's box");
    auto site = [&](int e) -> unsigned {
        const int sl = e / BC, bc = e % BC;
        const int a2 = bc % B2, a1 = (bc / B2) % B1, a0 = bc / (B1 * B2);
        const int f = sl >> 2;
        int g0 = o0 + a0 - (f == 2 || f == 3), g1 = o1 + a1 - (f == 1 || f == 3), g2 = o2 + a2 - (f == 1 || f == 2);
        g0 += g0 < 0 ? L0 : 0;
        g0 -= g0 >= L0 ? L0 : 0;
        g1 += g1 < 0 ? L1 : 0;
        g1 -= g1 >= L1 ? L1 : 0;
        g2 += g2 < 0 ? L2 : 0;
        g2 -= g2 >= L2 ? L2 : 0;
        return unsigned(sl * nc + (g0 * L1 + g1) * L2 + g2);
    };
    if constexpr (UNROLL == 0) {
        // all of the thread's loads (a wide workgroup's few) issued before
        // any value is stored, a fixed count of them with the last predicated
        constexpr int KB = (16 * BC + BS - 1) / BS;
        Vec3<Real> v[KB];
#pragma unroll
        for (int k = 0; k < KB; ++k) {
            const int e = int(threadIdx.x) + k * BS;
            if (e < 16 * BC) v[k] = ld3(Sr, site(e));
        }
#pragma unroll
        for (int k = 0; k < KB; ++k) {
            const int e = int(threadIdx.x) + k * BS;
            if (e < 16 * BC) {
                const int le = e + (e / BC) * (LS - BC);
                lx[le] = v[k].x;
                ly[le] = v[k].y;
                lz[le] = v[k].z;
            }
        }
        return;
    }
    // UNROLL iterations at a time, so their loads are in flight together
#pragma unroll UNROLL
    for (int e = threadIdx.x; e < 16 * BC; e += BS) {
        const Vec3<Real> v = ld3(Sr, site(e));
        const int le = e + (e / BC) * (LS - BC);
        lx[le] = v.x;
        ly[le] = v.y;
        lz[le] = v.z;
    }
}

// Tiled sweep: one workgroup per (tile, replica) of one checkerboard phase.
// The T0 x T1 x T2 tile and its one-sided one-cell halo (per sublattice the
// side its bonds reach, so (T0 + 1)(T1 + 1)(T2 + 1) cells of each of the 16
// sublattices) are loaded into LDS once; every pass of the sweep (FIRST, then n_or OR passes) and
// every colour runs on the LDS copy; the tile interior is written back once.
// Tiles of one phase are two tiles apart along each axis, so no two of them
// share a bond or write a cell another reads: the phase equals the tiles
// processed one after another.  A sweep is the 8 phases in order, and within
// a tile the passes in order, each pass visiting the colours in order, so
// every site gets each pass of the sweep exactly once (with the counter of
// that pass).  The lifted kinds also stage the interior's directions in LDS.
// Workgroups map to tiles XCD by XCD: the 8 XCDs take contiguous chunks of
// the (replica, tile) order, so consecutive tiles along n2, whose halo rows
// share cache lines, are fetched through the same L2.
// FUSED (the fused pass order): within a colour stage each site takes its
// first-kind update and then its first OR update back to back, from one
// evaluation of its field.  No bonded site changes in between (a colour
// class has no bond inside it), so the field is the one each pass would
// compute: this is the sequential order "per tile, per colour, per site,
// first kind then OR", with half the field evaluations and barriers of a
// lifted + OR sweep.  Further OR passes (n_or >= 2) follow as colour stages
// of their own: a second OR rotation about the same field would only
// continue the first (the sum of two uniform angles is one uniform angle).
// The site's uniforms for the sweep are drawn in order from its blocks:
// uniform j of the sweep is component j % 5 of the block with pass index
// pass0 + j / 5, the first kind taking the first kUniformsOf[FIRST] and OR
// pass k uniform kUniformsOf[FIRST] + k - 1 (a lifted + OR sweep draws one
// Philox block per site instead of two).
template <typename Real, int SHELLS, int FIRST, int T0, int T1, int T2, int BS, bool FUSED>
__global__ __launch_bounds__(BS) void pyro_tiled_kernel(DeviceState<Real> st, PyroCouplings<Real> J,
                                                         int phase, int n_or,
                                                         unsigned long long pass_index) {
    constexpr int B1 = T1 + 1, B2 = T2 + 1;
    constexpr int NI = T0 * T1 * T2;
    constexpr int NCOL = pyrostencil::colours_for(SHELLS);
    constexpr int SPC = 16 / NCOL;
    constexpr bool LIFTED = FIRST == 1 || FIRST == 2;
    constexpr int LS = pyro_tile_stride(T0, T1, T2, int(sizeof(Real)), LIFTED);  // LDS sublattice stride
    extern __shared__ __align__(16) unsigned char pyro_smem[];
    Real* lx = reinterpret_cast<Real*>(pyro_smem);
    Real* ly = lx + 16 * LS;
    Real* lz = ly + 16 * LS;
    // lifted directions of the interior, one bit per site (set: -1), W words
    // per sublattice; sites of one sublattice flip their own bits with LDS
    // atomics, so lanes sharing a word never overwrite each other
    constexpr int W = (NI + 31) / 32;
    unsigned int* ldw = reinterpret_cast<unsigned int*>(lz + 16 * LS);  // [16][W]

    const unsigned tiles = gridDim.x / unsigned(st.R);  // tiles per replica and phase
    unsigned g = blockIdx.x;
    if ((gridDim.x & 7u) == 0u) g = (g & 7u) * (gridDim.x >> 3) + (g >> 3);
    const int r = __builtin_amdgcn_readfirstlane(int(g / tiles));  // uniform: the replica's loads are scalar
    const int tile = int(g - unsigned(r) * tiles);
    const PyroCouplings<Real> Jr = J.at(r);
    const int L0 = st.L0, L1 = st.L1, L2 = st.L2, nc = st.n_cells;
    const int q1n = L1 / (2 * T1), q2n = L2 / (2 * T2);
    const int m2 = tile % q2n;
    const int t12 = tile / q2n;
    const int m1 = t12 % q1n, m0 = t12 / q1n;
    const int o0 = ((phase & 1) + 2 * m0) * T0;
    const int o1 = (((phase >> 1) & 1) + 2 * m1) * T1;
    const int o2 = (((phase >> 2) & 1) + 2 * m2) * T2;
    const size_t N = size_t(st.N);
    Real* __restrict__ Sr = st.S + size_t(r) * 3 * N;

    // the lane's cell (a constant table), the pass counter and the
    // replica's parameters read before the box, so their loads (two of them
    // dependent) complete under its loads instead of ahead of the first stage
    static_assert(NI <= BS, "one tile cell per thread");
    const int my_cell = kTileLanes<T0, T1, T2, BS>.cell[threadIdx.x];
    const unsigned long long pass0 = pass_index + (st.pass_base ? *st.pass_base : 0ull);
    const ReplicaParams<Real> rp = replica_params(st, r);
    int8_t* const dir_r = st.dir + size_t(r) * N;
    // the interior cell mi = threadIdx.x (< NI) whose 16 sites this thread
    // stages directions for and writes back, and its lattice cell
    const int mi = int(threadIdx.x);
    const bool has_mi = mi < NI;
    const int mi2 = mi % T2, mi1 = (mi / T2) % T1, mi0 = mi / (T1 * T2);
    const int cell_mi = ((o0 + mi0) * L1 + (o1 + mi1)) * L2 + (o2 + mi2);
    // the interior's directions are read together with the box (set: -1)
    bool neg[LIFTED ? 16 : 1] = {};
    if constexpr (LIFTED) {
        if (has_mi) {
#pragma unroll
            for (int sl = 0; sl < 16; ++sl) neg[sl] = dir_r[sl * nc + cell_mi] < 0;
        }
    }
    load_tile_box<Real, T0, T1, T2, BS, LS, 4>(lx, ly, lz, Sr, o0, o1, o2, L0, L1, L2, nc);
    if constexpr (LIFTED) {
        // wave v holds cells 64 v .. 64 v + 63, the bits of words 2 v and
        // 2 v + 1 of each sublattice: a ballot gives both words whole (cells
        // past the tile read 0), and BS >= NI means every word is written
        const int v = int(threadIdx.x) >> 6;
#pragma unroll
        for (int sl = 0; sl < 16; ++sl) {
            const unsigned long long b = __ballot(neg[sl]);
            if ((threadIdx.x & 63) == 0) {
                if (2 * v < W) ldw[sl * W + 2 * v] = unsigned(b);
                if (2 * v + 1 < W) ldw[sl * W + 2 * v + 1] = unsigned(b >> 32);
            }
        }
    }
    __syncthreads();

    // the cell this lane updates: its box index (before the sublattice's
    // shift) and lattice cell
    const int c2 = my_cell % T2, c1 = (my_cell / T2) % T1, c0 = my_cell / (T1 * T2);
    const int my_lc = (c0 * B1 + c1) * B2 + c2;
    const int my_gc = ((o0 + c0) * L1 + (o1 + c1)) * L2 + (o2 + c2);
    unsigned int my_acc = 0;
    if constexpr (FUSED) {
        // group 0: the first kind and the first OR update of every site of a
        // colour back to back; group g >= 1: OR pass g + 1 on its own
        const int n_groups = n_or > 1 ? n_or : 1;
        for (int grp = 0; grp < n_groups; ++grp) {
        for (int col = 0; col < NCOL; ++col) {
            for (int q = 0; q < SPC; ++q) {
                const int sl = NCOL == 16 ? col : col + 4 * q;
                const int f = sl >> 2;
                const int own = ((f == 2 || f == 3) * B1 + (f == 1 || f == 3)) * B2 + (f == 1 || f == 2);
                if (my_cell >= 0) {
                    const int m = my_cell;
                    const int lc = my_lc;
                    Real hx, hy, hz;
                    switch (sl) {
#define PYRO_SL_CASE(X) \
    case X: pyro_field_lds<Real, SHELLS, X, B1, B2, LS>(lx, ly, lz, lc, Jr, hx, hy, hz); break;
                        PYRO_SL_CASE(0) PYRO_SL_CASE(1) PYRO_SL_CASE(2) PYRO_SL_CASE(3)
                        PYRO_SL_CASE(4) PYRO_SL_CASE(5) PYRO_SL_CASE(6) PYRO_SL_CASE(7)
                        PYRO_SL_CASE(8) PYRO_SL_CASE(9) PYRO_SL_CASE(10) PYRO_SL_CASE(11)
                        PYRO_SL_CASE(12) PYRO_SL_CASE(13) PYRO_SL_CASE(14) PYRO_SL_CASE(15)
#undef PYRO_SL_CASE
                        default: hx = hy = hz = Real(0);
                    }
                    const int site = sl * nc + my_gc;
                    const int li = sl * LS + lc + own;
                    Real x = lx[li], y = ly[li], z = lz[li];
                    bool any = false, ch;
                    if (grp == 0) {
                        int8_t d = 1, d0 = 1;
                        unsigned int* wp = nullptr;
                        const unsigned int bit = 1u << (m & 31);
                        if constexpr (LIFTED) {
                            wp = ldw + sl * W + (m >> 5);
                            d0 = d = (*wp & bit) ? int8_t(-1) : int8_t(1);
                        }
                        Uniform5 ub = site_uniforms(st, site, r, pass0);
                        my_acc += update_spin_u<Real, FIRST>(rp, x, y, z, ch, ub, hx, hy, hz, LIFTED ? &d : nullptr)
                                      ? 1u : 0u;
                        any = ch;
                        if constexpr (LIFTED) {
                            if (d != d0) atomicXor(wp, bit);
                        }
                        if (n_or >= 1) {
                            constexpr int j = kUniformsOf[FIRST];  // the sweep's uniform of its first OR update
                            if constexpr (j % 5 == 0) ub = site_uniforms(st, site, r, pass0 + unsigned(j / 5));
                            Uniform5 uo;
                            uo.m[0] = ub.m[j % 5];
                            update_spin_u<Real, 3>(rp, x, y, z, ch, uo, hx, hy, hz, nullptr);
                            any = any || ch;
                        }
                    } else {
                        const int j = kUniformsOf[FIRST] + grp;  // the sweep's uniform of OR pass grp + 1
                        const Uniform5 ub = site_uniforms(st, site, r, pass0 + unsigned(j / 5));
                        Uniform5 uo;
                        uo.m[0] = uniform_at(ub, j % 5);
                        update_spin_u<Real, 3>(rp, x, y, z, ch, uo, hx, hy, hz, nullptr);
                        any = ch;
                    }
                    if (any) {
                        lx[li] = x;
                        ly[li] = y;
                        lz[li] = z;
                    }
                }
            }
            __syncthreads();
        }
        }
    } else
    for (int k = 0; k <= n_or; ++k) {
        for (int col = 0; col < NCOL; ++col) {
            for (int q = 0; q < SPC; ++q) {
                const int sl = NCOL == 16 ? col : col + 4 * q;
                const int f = sl >> 2;
                const int own = ((f == 2 || f == 3) * B1 + (f == 1 || f == 3)) * B2 + (f == 1 || f == 2);
                if (my_cell >= 0) {
                    const int m = my_cell;
                    const int lc = my_lc;
                    // the field in a block of its own: drawing the uniforms
                    // next to the LDS gathers raises the register count past
                    // what three workgroups per CU allow
                    Real hx, hy, hz;
                    switch (sl) {
#define PYRO_SL_CASE(X) \
    case X: pyro_field_lds<Real, SHELLS, X, B1, B2, LS>(lx, ly, lz, lc, Jr, hx, hy, hz); break;
                        PYRO_SL_CASE(0) PYRO_SL_CASE(1) PYRO_SL_CASE(2) PYRO_SL_CASE(3)
                        PYRO_SL_CASE(4) PYRO_SL_CASE(5) PYRO_SL_CASE(6) PYRO_SL_CASE(7)
                        PYRO_SL_CASE(8) PYRO_SL_CASE(9) PYRO_SL_CASE(10) PYRO_SL_CASE(11)
                        PYRO_SL_CASE(12) PYRO_SL_CASE(13) PYRO_SL_CASE(14) PYRO_SL_CASE(15)
#undef PYRO_SL_CASE
                        default: hx = hy = hz = Real(0);
                    }
                    const int site = sl * nc + my_gc;
                    const int li = sl * LS + lc + own;
                    Real x = lx[li], y = ly[li], z = lz[li];
                    bool ch;
                    if (k == 0) {
                        int8_t d = 1, d0 = 1;
                        unsigned int* wp = nullptr;
                        const unsigned int bit = 1u << (m & 31);
                        if constexpr (LIFTED) {
                            wp = ldw + sl * W + (m >> 5);
                            d0 = d = (*wp & bit) ? int8_t(-1) : int8_t(1);
                        }
                        my_acc += update_spin_p<Real, FIRST>(rp, st, x, y, z, ch, site, r, pass0, hx, hy, hz,
                                                             LIFTED ? &d : nullptr) ? 1u : 0u;
                        if constexpr (LIFTED) {
                            if (d != d0) atomicXor(wp, bit);
                        }
                    } else {
                        update_spin_p<Real, 3>(rp, st, x, y, z, ch, site, r, pass0 + k, hx, hy, hz, nullptr);
                    }
                    if (ch) {
                        lx[li] = x;
                        ly[li] = y;
                        lz[li] = z;
                    }
                }
            }
            __syncthreads();
        }
    }

    if (has_mi) {
#pragma unroll
        for (int sl = 0; sl < 16; ++sl) {
            const int f = sl >> 2;
            const int li = sl * LS + ((mi0 + (f == 2 || f == 3)) * B1 + (mi1 + (f == 1 || f == 3))) * B2 +
                           (mi2 + (f == 1 || f == 2));
            const int site = sl * nc + cell_mi;
            st3(Sr, unsigned(site), lx[li], ly[li], lz[li]);
            if constexpr (LIFTED)
                dir_r[site] = ((ldw[sl * W + (mi >> 5)] >> (mi & 31)) & 1u) ? int8_t(-1) : int8_t(1);
        }
    }

    if (st.acc == nullptr) return;  // uniform: counts not requested
    __shared__ unsigned int red[BS / 64];
    unsigned int v = my_acc;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int i = 0; i < BS / 64; ++i) t += red[i];
        atomicAdd(st.acc + r, t);
    }
}

// Energy partial sums of pyro_energy_kernel from tiles in LDS, for lattices
// the tiled sweep takes: one workgroup per (tile, replica) over all tiles
// (no phases: the kernel only reads), the tile and its one-sided halo read
// once.  The tile's sites are then taken as (sublattice, 64-cell chunk)
// items, item i by wave i mod (BS / 64), so a wave's lanes share one
// sublattice and the 16 sublattices are spread over the waves; each site
// adds S_i . h_i over the bonds to higher sublattices minus B . S_i, with
// the same per-site arithmetic as pyro_energy_kernel.  partial[r][tile].
template <typename Real, int SHELLS, int T0, int T1, int T2, int BS>
__global__ __launch_bounds__(BS) void pyro_energy_tiled_kernel(DeviceState<Real> st, PyroCouplings<Real> J,
                                                               double* __restrict__ partial) {
    constexpr int B1 = T1 + 1, B2 = T2 + 1;
    constexpr int LS = pyro_tile_stride(T0, T1, T2, int(sizeof(Real)), false);  // LDS sublattice stride
    extern __shared__ __align__(16) unsigned char pyro_smem[];
    Real* lx = reinterpret_cast<Real*>(pyro_smem);
    Real* ly = lx + 16 * LS;
    Real* lz = ly + 16 * LS;
    const unsigned tiles = gridDim.x / unsigned(st.R);  // tiles per replica
    unsigned g = blockIdx.x;
    if ((gridDim.x & 7u) == 0u) g = (g & 7u) * (gridDim.x >> 3) + (g >> 3);
    const int r = __builtin_amdgcn_readfirstlane(int(g / tiles));  // uniform: the replica's loads are scalar
    const int tile = int(g - unsigned(r) * tiles);
    const int q1n = st.L1 / T1, q2n = st.L2 / T2;
    const int m2 = tile % q2n;
    const int t12 = tile / q2n;
    const int m1 = t12 % q1n, m0 = t12 / q1n;
    const size_t N = size_t(st.N);
    const Real* __restrict__ Sr = st.S + size_t(r) * 3 * N;
    // the replica's couplings and field read before the box, so that their
    // loads complete under its loads
    const PyroCouplings<Real> Jr = J.at(r);
    const Real bx = st.B[3 * r], by = st.B[3 * r + 1], bz = st.B[3 * r + 2];
    load_tile_box<Real, T0, T1, T2, BS, LS, 0>(lx, ly, lz, Sr, m0 * T0, m1 * T1, m2 * T2, st.L0, st.L1, st.L2,
                                               st.n_cells);
    __syncthreads();
    double e = 0.0;
    constexpr int NI = T0 * T1 * T2, CH = (NI + 63) / 64, WAVES = BS / 64;
    for (int item = threadIdx.x >> 6; item < 16 * CH; item += WAVES) {
        const int sl = item / CH;  // uniform across the wave
        const int m = (item - sl * CH) * 64 + (threadIdx.x & 63);
        if (m < NI) {
            const int i2 = m % T2, i1 = (m / T2) % T1, i0 = m / (T1 * T2);
            const int lc = (i0 * B1 + i1) * B2 + i2;
            const int f = sl >> 2;
            const int own = ((f == 2 || f == 3) * B1 + (f == 1 || f == 3)) * B2 + (f == 1 || f == 2);
            Real hx = Real(0), hy = Real(0), hz = Real(0);
            switch (sl) {
#define PYRO_SL_CASE(X) \
    case X: pyro_field_lds<Real, SHELLS, X, B1, B2, LS, true>(lx, ly, lz, lc, Jr, hx, hy, hz); break;
                PYRO_SL_CASE(0) PYRO_SL_CASE(1) PYRO_SL_CASE(2) PYRO_SL_CASE(3)
                PYRO_SL_CASE(4) PYRO_SL_CASE(5) PYRO_SL_CASE(6) PYRO_SL_CASE(7)
                PYRO_SL_CASE(8) PYRO_SL_CASE(9) PYRO_SL_CASE(10) PYRO_SL_CASE(11)
                PYRO_SL_CASE(12) PYRO_SL_CASE(13) PYRO_SL_CASE(14) PYRO_SL_CASE(15)
#undef PYRO_SL_CASE
                default: break;
            }
            const int li = sl * LS + lc + own;
            const Real sx = lx[li], sy = ly[li], sz = lz[li];
            const Real bond = sx * hx + sy * hy + sz * hz;
            const Real bs = sx * bx + sy * by + sz * bz;
            e += double(bond) - double(bs);
        }
    }
    for (int off = 32; off > 0; off >>= 1) e += __shfl_down(e, off, 64);
    __shared__ double red[BS / 64];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < BS / 64; ++i) t += red[i];
        partial[size_t(r) * tiles + tile] = t;
    }
}

template <typename Real, typename F>
hipError_t dispatch_pyro(const PyroCouplings<Real>& J, F&& f) {
    switch (J.shells) {
        case 1: return f(std::integral_constant<int, 1>{});
        case 2: return f(std::integral_constant<int, 2>{});
        case 3: return f(std::integral_constant<int, 3>{});
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

template <typename Real>
hipError_t launch_pyro_colour(const DeviceState<Real>& st, const PyroCouplings<Real>& J, int colour,
                              int kind, unsigned long long pass_index, hipStream_t s, double* epart, bool fused,
                              int ucomp) {
    const int ncol = J.shells == 3 ? 16 : 4;
    if (fused && kind == 3) return hipErrorInvalidValue;
    if (ucomp < 0 || ucomp > 4 || (ucomp != 0 && kind != 3)) return hipErrorInvalidValue;
    // One wave per workgroup: measured on the L=12 x 64 tempering lattice,
    // 64-thread workgroups take 5.17 us per colour launch against 5.39 us
    // for 128, 256 or 512 threads (same sites, same results).
    constexpr int BS = kPyroColourThreads;
    const int bps = (st.n_cells + BS - 1) / BS;
    const int per_launch = bps * (16 / ncol);
    const dim3 grid(unsigned(per_launch) * unsigned(st.R));  // replica-major, XCD-aware (xcd_split)
    // energy partials of colour c at [r][c * per_launch + block]: 16 bps per replica
    double* ep = epart ? epart + size_t(colour) * size_t(per_launch) : nullptr;
    const int es = pyro_colour_energy_blocks(st.n_cells);
    auto go = [&](auto per_rep, auto en, auto fu) -> hipError_t {
        constexpr bool PR = decltype(per_rep)::value;
        constexpr bool EN = decltype(en)::value;
        constexpr bool FU = decltype(fu)::value;
        return dispatch_pyro(J, [&](auto shells) -> hipError_t {
            constexpr int SH = decltype(shells)::value;
            switch (kind) {
                case 0: hipLaunchKernelGGL((pyro_colour_kernel<Real, SH, 0, PR, BS, EN, FU>), grid, dim3(BS), 0, s, st, J, colour, pass_index, ucomp, ep, es); break;
                case 1: hipLaunchKernelGGL((pyro_colour_kernel<Real, SH, 1, PR, BS, EN, FU>), grid, dim3(BS), 0, s, st, J, colour, pass_index, ucomp, ep, es); break;
                case 2: hipLaunchKernelGGL((pyro_colour_kernel<Real, SH, 2, PR, BS, EN, FU>), grid, dim3(BS), 0, s, st, J, colour, pass_index, ucomp, ep, es); break;
                case 3:
                    if constexpr (!FU) {
                        hipLaunchKernelGGL((pyro_colour_kernel<Real, SH, 3, PR, BS, EN, false>), grid, dim3(BS), 0, s, st, J, colour, pass_index, ucomp, ep, es);
                        break;
                    }
                    return hipErrorInvalidValue;
                default: return hipErrorInvalidValue;
            }
            return hipGetLastError();
        });
    };
    auto by_fused = [&](auto per_rep, auto en) -> hipError_t {
        return fused ? go(per_rep, en, std::true_type{}) : go(per_rep, en, std::false_type{});
    };
    if (epart) return J.per_replica ? by_fused(std::true_type{}, std::true_type{}) : by_fused(std::false_type{}, std::true_type{});
    return J.per_replica ? by_fused(std::true_type{}, std::false_type{}) : by_fused(std::false_type{}, std::false_type{});
}

template <typename Real>
hipError_t launch_pyro_energy(const DeviceState<Real>& st, const PyroCouplings<Real>& J, double* partial,
                              hipStream_t s) {
    const dim3 grid(unsigned(pyro_energy_blocks(st.N)) * unsigned(st.R));  // XCD-aware (xcd_split)
    return dispatch_pyro(J, [&](auto shells) -> hipError_t {
        constexpr int SH = decltype(shells)::value;
        hipLaunchKernelGGL((pyro_energy_kernel<Real, SH>), grid, dim3(64), 0, s, st, J, partial);
        return hipGetLastError();
    });
}

template <typename Real>
hipError_t launch_pyro_fields(const DeviceState<Real>& st, const PyroCouplings<Real>& J, int replica, Real* out,
                              hipStream_t s) {
    const dim3 grid(unsigned((st.N + 255) / 256));
    return dispatch_pyro(J, [&](auto shells) -> hipError_t {
        constexpr int SH = decltype(shells)::value;
        hipLaunchKernelGGL((pyro_fields_kernel<Real, SH>), grid, dim3(256), 0, s, st, J, replica, out);
        return hipGetLastError();
    });
}

template <typename Real>
hipError_t launch_pyro_resident(const DeviceState<Real>& st, const PyroCouplings<Real>& J, int kind,
                                int n_or, int n_sweeps, unsigned long long pass_index,
                                hipStream_t s, double* epart, bool fused) {
    constexpr int BS = 1024, CPT = 2;
    if (st.n_cells > BS * CPT) return hipErrorInvalidValue;
    const dim3 grid(unsigned(st.R));
    auto go = [&](auto en, auto fu) -> hipError_t {
        constexpr bool EN = decltype(en)::value;
        constexpr bool FU = decltype(fu)::value;
        return dispatch_pyro(J, [&](auto shells) -> hipError_t {
            constexpr int SH = decltype(shells)::value;
            switch (kind) {
                case 0: hipLaunchKernelGGL((pyro_resident_kernel<Real, SH, 0, BS, CPT, EN, FU>), grid, dim3(BS), 0, s, st, J, n_or, n_sweeps, pass_index, epart); break;
                case 1: hipLaunchKernelGGL((pyro_resident_kernel<Real, SH, 1, BS, CPT, EN, FU>), grid, dim3(BS), 0, s, st, J, n_or, n_sweeps, pass_index, epart); break;
                case 2: hipLaunchKernelGGL((pyro_resident_kernel<Real, SH, 2, BS, CPT, EN, FU>), grid, dim3(BS), 0, s, st, J, n_or, n_sweeps, pass_index, epart); break;
                default: return hipErrorInvalidValue;
            }
            return hipGetLastError();
        });
    };
    auto by_fused = [&](auto en) -> hipError_t {
        return fused && n_or >= 1 ? go(en, std::true_type{}) : go(en, std::false_type{});
    };
    return epart ? by_fused(std::true_type{}) : by_fused(std::false_type{});
}

template <typename Real>
hipError_t launch_pyro_xcd(const DeviceState<Real>& st, const PyroCouplings<Real>& J, int kind, int n_or,
                           int n_sweeps, unsigned long long pass_index, hipStream_t s, double* epart, bool fused,
                           unsigned int* bar, int* err, unsigned long long timeout_ticks) {
    if (kind < 0 || kind > 2) return hipErrorInvalidValue;
    const int G = pyro_xcd_workgroups(st.R, st.n_cells, J.shells);
    const bool fu = fused && n_or >= 1;
    auto go = [&](auto per_rep, auto en, auto fuc) -> hipError_t {
        constexpr bool PR = decltype(per_rep)::value;
        constexpr bool EN = decltype(en)::value;
        constexpr bool FU = decltype(fuc)::value;
        return dispatch_pyro(J, [&](auto shells) -> hipError_t {
            constexpr int SH = decltype(shells)::value;
            auto one = [&](auto kk) -> hipError_t {
                auto* fn = pyro_xcd_kernel<Real, SH, decltype(kk)::value, PR, EN, FU>;
                // every workgroup of a replica must be resident at once (they
                // wait for one another): refuse a grid beyond what fits
                int dev = 0, cus = 0, per_cu = 0;
                hipError_t e = hipGetDevice(&dev);
                if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
                if (e == hipSuccess)
                    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(fn), 64, 0);
                if (e != hipSuccess) return e;
                if (cus % 8 != 0 || size_t(per_cu) * size_t(cus) < size_t(G) || st.N >= (1 << 24))
                    return hipErrorNotSupported;
                hipLaunchKernelGGL(xcd_zero_kernel, dim3(1), dim3(256), 0, s, bar, (st.R + 8) * kXcdPad);
                hipLaunchKernelGGL(fn, dim3(unsigned(G)), dim3(64), 0, s, st, J, n_or, n_sweeps, pass_index, epart,
                                   bar, err, timeout_ticks, 0u);
                return hipGetLastError();
            };
            switch (kind) {
                case 0: return one(std::integral_constant<int, 0>{});
                case 1: return one(std::integral_constant<int, 1>{});
                default: return one(std::integral_constant<int, 2>{});
            }
        });
    };
    auto by_fused = [&](auto per_rep, auto en) -> hipError_t {
        return fu ? go(per_rep, en, std::true_type{}) : go(per_rep, en, std::false_type{});
    };
    auto by_en = [&](auto per_rep) -> hipError_t {
        return epart ? by_fused(per_rep, std::true_type{}) : by_fused(per_rep, std::false_type{});
    };
    return J.per_replica ? by_en(std::true_type{}) : by_en(std::false_type{});
}

namespace {
// Calls f with the batch kernel instance of the given template arguments.
template <typename Real, typename F>
hipError_t with_batch_kernel(int shells, bool per_rep, int kind, bool energy, bool fused, F&& f) {
    if (kind < 0 || kind > 2) return hipErrorInvalidValue;
    auto go = [&](auto sh, auto pr, auto en, auto fu) -> hipError_t {
        constexpr int SH = decltype(sh)::value;
        constexpr bool PR = decltype(pr)::value, EN = decltype(en)::value, FU = decltype(fu)::value;
        switch (kind) {
            case 0: return f(pyro_batch_kernel<Real, SH, 0, PR, EN, FU>);
            case 1: return f(pyro_batch_kernel<Real, SH, 1, PR, EN, FU>);
            default: return f(pyro_batch_kernel<Real, SH, 2, PR, EN, FU>);
        }
    };
    auto by_fu = [&](auto sh, auto pr, auto en) -> hipError_t {
        return fused ? go(sh, pr, en, std::true_type{}) : go(sh, pr, en, std::false_type{});
    };
    auto by_en = [&](auto sh, auto pr) -> hipError_t {
        return energy ? by_fu(sh, pr, std::true_type{}) : by_fu(sh, pr, std::false_type{});
    };
    auto by_pr = [&](auto sh) -> hipError_t {
        return per_rep ? by_en(sh, std::true_type{}) : by_en(sh, std::false_type{});
    };
    switch (shells) {
        case 1: return by_pr(std::integral_constant<int, 1>{});
        case 2: return by_pr(std::integral_constant<int, 2>{});
        case 3: return by_pr(std::integral_constant<int, 3>{});
        default: return hipErrorInvalidValue;
    }
}
}  // namespace

template <typename Real>
int pyro_batch_capacity(int shells, bool per_rep, int kind, bool energy, bool fused) {
    int cap = 0;
    const hipError_t e = with_batch_kernel<Real>(shells, per_rep, kind, energy, fused, [&](auto fn) -> hipError_t {
        int dev = 0, cus = 0, per_cu = 0;
        hipError_t r = hipGetDevice(&dev);
        if (r == hipSuccess) r = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (r == hipSuccess)
            r = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(fn), 64, 0);
        if (r == hipSuccess && cus % 8 == 0) cap = per_cu * cus;
        return r;
    });
    return e == hipSuccess ? cap : 0;
}

template <typename Real>
hipError_t launch_pyro_batch(const BatchMember<Real>* members, const int4* work, int per_xcd, int n_items, int shells,
                             bool per_rep, int kind, int n_or, int n_sweeps, bool energy, bool fused,
                             unsigned int* bar, hipStream_t s) {
    const bool fu = fused && n_or >= 1;
    // every workgroup of a replica must be resident at once (they wait for
    // one another): refuse a grid beyond what fits
    const int cap = pyro_batch_capacity<Real>(shells, per_rep, kind, energy, fu);
    if (per_xcd <= 0 || cap < 8 * per_xcd) return hipErrorNotSupported;
    return with_batch_kernel<Real>(shells, per_rep, kind, energy, fu, [&](auto fn) -> hipError_t {
        hipLaunchKernelGGL(xcd_zero_kernel, dim3(1), dim3(256), 0, s, bar, (n_items + 8) * kXcdPad);
        hipLaunchKernelGGL(fn, dim3(unsigned(8 * per_xcd)), dim3(64), 0, s, members, work, per_xcd, n_items, n_or,
                           n_sweeps, bar, 0u);
        return hipGetLastError();
    });
}

template <typename Real>
hipError_t launch_pyro_tiled(const DeviceState<Real>& st, const PyroCouplings<Real>& J, int tile_kind,
                             int phase, int kind, int n_or, unsigned long long pass_index, hipStream_t s,
                             bool fused) {
    auto go = [&](auto t0, auto t1, auto t2) -> hipError_t {
        constexpr int T0 = decltype(t0)::value, T1 = decltype(t1)::value, T2 = decltype(t2)::value;
        constexpr int BS = pyro_tile_threads(T0 * T1 * T2);
        const size_t lds = pyro_tile_lds(T0, T1, T2, int(sizeof(Real)), kind == 1 || kind == 2);
        const unsigned tiles = unsigned((st.L0 / (2 * T0)) * (st.L1 / (2 * T1)) * (st.L2 / (2 * T2)));
        const dim3 grid(tiles * unsigned(st.R));  // (replica, tile), XCD-chunked in the kernel
        return dispatch_pyro(J, [&](auto shells) -> hipError_t {
            constexpr int SH = decltype(shells)::value;
            auto launch = [&](auto kk, auto fu) -> hipError_t {
                auto* fn = pyro_tiled_kernel<Real, SH, decltype(kk)::value, T0, T1, T2, BS, decltype(fu)::value>;
                (void)hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
                hipLaunchKernelGGL(fn, grid, dim3(BS), lds, s, st, J, phase, n_or, pass_index);
                return hipGetLastError();
            };
            auto by_kind = [&](auto fu) -> hipError_t {
                switch (kind) {
                    case 0: return launch(std::integral_constant<int, 0>{}, fu);
                    case 1: return launch(std::integral_constant<int, 1>{}, fu);
                    case 2: return launch(std::integral_constant<int, 2>{}, fu);
                    default: return hipErrorInvalidValue;
                }
            };
            return fused ? by_kind(std::true_type{}) : by_kind(std::false_type{});
        });
    };
    using I4 = std::integral_constant<int, 4>;
    using I6 = std::integral_constant<int, 6>;
    using I8 = std::integral_constant<int, 8>;
    switch (tile_kind) {
        case 1: return go(I4{}, I4{}, I4{});
        case 2: return go(I6{}, I6{}, I4{});
        case 3: return go(I8{}, I4{}, I4{});
        case 4: return go(I4{}, I4{}, I8{});
        case 5: return go(I8{}, I8{}, I4{});
        default: return hipErrorInvalidValue;
    }
}

template <typename Real>
hipError_t launch_pyro_energy_tiled(const DeviceState<Real>& st, const PyroCouplings<Real>& J, int tile_kind,
                                    double* partial, int* nblk, hipStream_t s) {
    auto go = [&](auto t0, auto t1, auto t2) -> hipError_t {
        constexpr int T0 = decltype(t0)::value, T1 = decltype(t1)::value, T2 = decltype(t2)::value;
        constexpr int BS = 512;  // 8 waves share the tile's (sublattice, chunk) items
        if (st.L0 % T0 || st.L1 % T1 || st.L2 % T2) return hipErrorInvalidValue;
        const size_t lds = pyro_tile_lds(T0, T1, T2, int(sizeof(Real)), false);
        const unsigned tiles = unsigned((st.L0 / T0) * (st.L1 / T1) * (st.L2 / T2));
        *nblk = int(tiles);
        const dim3 grid(tiles * unsigned(st.R));  // (replica, tile), XCD-chunked in the kernel
        return dispatch_pyro(J, [&](auto shells) -> hipError_t {
            constexpr int SH = decltype(shells)::value;
            auto* fn = pyro_energy_tiled_kernel<Real, SH, T0, T1, T2, BS>;
            const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(fn, grid, dim3(BS), lds, s, st, J, partial);
            return hipGetLastError();
        });
    };
    using I4 = std::integral_constant<int, 4>;
    using I6 = std::integral_constant<int, 6>;
    using I8 = std::integral_constant<int, 8>;
    switch (tile_kind) {
        case 1: return go(I4{}, I4{}, I4{});
        case 2: return go(I6{}, I6{}, I4{});
        case 3: return go(I8{}, I4{}, I4{});
        case 4: return go(I4{}, I4{}, I8{});
        case 5: return go(I8{}, I8{}, I4{});
        default: return hipErrorInvalidValue;
    }
}

}  // namespace pyro
