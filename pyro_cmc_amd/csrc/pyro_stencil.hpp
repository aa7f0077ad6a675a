// Compile-time neighbour stencil of the pyrochlore cubic cell.
//
// Sublattice sl = 4 f + p sits at R_fcc[f] + link[p] inside a cubic cell of
// edge 8 (pyrochlore_geometry.hpp:12-32).  A neighbour of sl at
// displacement d lands on sublattice t at cell offset (d0, d1, d2) in
// {-1, 0, 1}^3.  Shells by squared length: |d|^2 = 8 (nn1), 24 (nn2) and
// 32 (nn3, both the a and b kinds; they carry the same J3).  Derived here by
// search, so the table is the geometry rather than a transcription, and the
// kernels index it with constants after unrolling.
#pragma once

namespace pyro {
namespace pyrostencil {

struct Nbr {
    int sl;      // target sublattice
    int d[3];    // target cell offset
};

struct Stencil {
    Nbr n[16][30];  // per sublattice: 6 nn1, then 12 nn2, then 12 nn3
    int link[16][3];
};

constexpr int kFcc[4][3] = {{0, 0, 0}, {0, 4, 4}, {4, 0, 4}, {4, 4, 0}};
constexpr int kLink[4][3] = {{1, 1, 1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, 1}};

constexpr int fmod8(int a) { return ((a % 8) + 8) % 8; }
constexpr int fdiv8(int a) { return (a - fmod8(a)) / 8; }

constexpr Stencil make_stencil() {
    Stencil s{};
    for (int sl = 0; sl < 16; ++sl) {
        const int f = sl / 4, p = sl % 4;
        for (int a = 0; a < 3; ++a) s.link[sl][a] = kLink[p][a];
        int r[3] = {kFcc[f][0] + kLink[p][0], kFcc[f][1] + kLink[p][1], kFcc[f][2] + kLink[p][2]};
        int count = 0;
        const int lens[3] = {8, 24, 32};
        for (int shell = 0; shell < 3; ++shell)
            for (int dx = -4; dx <= 4; ++dx)
                for (int dy = -4; dy <= 4; ++dy)
                    for (int dz = -4; dz <= 4; ++dz) {
                        if (dx * dx + dy * dy + dz * dz != lens[shell]) continue;
                        const int q[3] = {r[0] + dx, r[1] + dy, r[2] + dz};
                        for (int t = 0; t < 16; ++t) {
                            const int tf = t / 4, tp = t % 4;
                            const int b[3] = {kFcc[tf][0] + kLink[tp][0], kFcc[tf][1] + kLink[tp][1],
                                              kFcc[tf][2] + kLink[tp][2]};
                            if (fmod8(q[0] - b[0]) || fmod8(q[1] - b[1]) || fmod8(q[2] - b[2])) continue;
                            Nbr& n = s.n[sl][count++];
                            n.sl = t;
                            for (int a = 0; a < 3; ++a) n.d[a] = fdiv8(q[a] - b[a]);
                        }
                    }
    }
    return s;
}

constexpr Stencil kStencil = make_stencil();

// Number of slots used by a coupling structure: J1 only, J1+J2, J1+J2+J3.
constexpr int slots_for(int shells) { return shells == 1 ? 6 : shells == 2 ? 18 : 30; }
// Colours: sites of one pyro_sl p share no J1/J2 bond (4 colours, colour p =
// {sl : sl % 4 == p}); J3 bonds join equal p, so each sublattice is its own
// colour (16 colours, colour = sl).
constexpr int colours_for(int shells) { return shells == 3 ? 16 : 4; }

// One-sided halo of a tile.  Inside its cell a site of sublattice sl sits
// at R_fcc[f] + link[p], coordinate -1, 1 (R_fcc = 0) or 3, 5 (R_fcc = 4)
// along each axis.  Bonds reach at most 4 along an axis, so sites of the
// first kind are only reached across the tile's upper face and sites of
// the second kind only across its lower face: a tile of T cells needs T + 1
// cells of each sublattice per axis, starting one cell below the tile
// (halo_lo = 1) or at the tile (halo_lo = 0).
constexpr int halo_lo(int sl, int axis) { return kFcc[sl / 4][axis] == 4 ? 1 : 0; }
constexpr bool one_sided_halo_covers() {
    for (int sl = 0; sl < 16; ++sl)
        for (int k = 0; k < 30; ++k)
            for (int a = 0; a < 3; ++a) {
                const int v = kStencil.n[sl][k].d[a] + halo_lo(kStencil.n[sl][k].sl, a);
                if (v < 0 || v > 1) return false;
            }
    return true;
}
static_assert(one_sided_halo_covers(), "every neighbour must lie in its sublattice's one-sided halo box");

}  // namespace pyrostencil
}  // namespace pyro
