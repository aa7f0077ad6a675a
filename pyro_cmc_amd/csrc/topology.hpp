// Host-side lattice topology for the sweep kernels: site positions, the
// int32 neighbour table, per-slot coupling matrices and the graph colouring.
//
// Follows the bond-shell construction of MC_runner::setup_lattice
// (MC.cpp:15-52) and define_coupling (MC.cpp:56-66): for each coupling in
// definition order, each site's neighbours at origin + v (v from the
// displacement list of the site's kind) are split into a "source" shell
// (kind(origin) <= kind(other): matrix J) and a "target" shell (matrix J^T).
// Here every neighbour becomes one "slot" carrying its own matrix, so the
// field is h_i = sum_slots J_slot S_nbr(slot), the same sum regrouped.
// Couplings whose matrix is exactly zero contribute exact zeros and are
// dropped (cli_bits.hpp:198-200 defines them anyway).
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace pyro {

using ivec3 = std::array<int32_t, 3>;

struct CouplingDef {
    std::string name;
    std::vector<std::vector<ivec3>> rel;  // rel[kind] = displacement list
    double J[9];                          // row-major, applied by the source
};

struct LatticeGeom {
    int n_sl = 0;
    std::vector<ivec3> sl_pos;   // [n_sl]
    std::vector<int32_t> sl_kind;  // [n_sl]
    ivec3 cell{};                // orthorhombic cell edges
    ivec3 L{};                   // supercell extent in cells

    int64_t n_cells() const { return int64_t(L[0]) * L[1] * L[2]; }
    int64_t n_sites() const { return n_cells() * n_sl; }
    int64_t cell_index(int n0, int n1, int n2) const {
        return (int64_t(n0) * L[1] + n1) * L[2] + n2;
    }
    ivec3 position(int64_t site) const;
    // Site at integer position p (periodic); -1 if no site sits there.
    int64_t site_at(const ivec3& p) const;
    void validate() const;
};

struct Topology {
    int n_slot = 0;                   // slots per site (max over sublattices)
    std::vector<int32_t> n_slot_sl;   // real slots per sublattice
    std::vector<int32_t> nbr;         // [N][n_slot], padded with the site itself
    std::vector<double> Jslot;        // [n_sl][n_slot][9], zero on padding
    bool isotropic = true;            // every slot matrix is c*I
    std::vector<double> cslot;        // [n_sl][n_slot] scalar c when isotropic
    // Colouring: classes are (sublattice, n mod period); a colour is a set of
    // classes no two of which share a bond.
    ivec3 period{1, 1, 1};
    int n_colours = 0;
    std::vector<int32_t> class_colour;  // [n_sl * P], P = period product
    std::vector<std::vector<int32_t>> colour_classes;  // classes per colour
};

// Builds the topology; throws std::invalid_argument on inconsistent input
// (missing displacement lists, displacement to a non-site, no valid colouring).
Topology build_topology(const LatticeGeom& g, const std::vector<CouplingDef>& cs);

// Pyrochlore cubic cell (pyrochlore_geometry.hpp:12-32): 16 sites, sl = 4f+p,
// position R_fcc[f] + link[p], kind p, cell edge 8.
LatticeGeom pyrochlore_cell(int L0, int L1, int L2);

// Neighbour displacement shells of the pyrochlore lattice by squared length
// and, for |d|^2 = 32, by whether the bond midpoint carries a spin.
// which: 1 = nn1 (|d|^2=8), 2 = nn2 (24), 3 = nn3a (32, midpoint is a site),
// 4 = nn3b (32, midpoint empty).  Returned per kind p = 0..3.
std::vector<std::vector<ivec3>> pyrochlore_shell(int which);

// The nine couplings of build_J1J2J3_h (cli_bits.hpp:168-200).
std::vector<CouplingDef> pyrochlore_J1J2J3(double J1, double J2, double J3,
                                           double Delta);

// Directed bond alpha -> beta of the Luttinger-Tisza matrix: displacement d
// (integer coordinates) and 3x3 coupling J (row-major).
struct LtBond {
    double d[3];
    double J[9];
};

// Bonds of the pyrochlore J1-J2-J3 XXZ model for the Luttinger-Tisza matrix
// (luttinger_tisza.cpp:31-81): for every sublattice alpha of the cubic cell
// and every displacement v of the nn1, nn2, nn3a, nn3b shells of its kind,
// the bond alpha -> beta (beta = the sublattice at r_alpha + v modulo the
// cell) with J = J1 [I + (Delta - 1) z_alpha z_beta^T] for nn1 (z_p the
// local [111] axis of kind p) and J = J2 I, J3 I, J3 I for the others;
// shells with zero coupling are left out.  Bonds are grouped by block
// alpha * 16 + beta: block b holds bonds[begin[b] .. begin[b + 1]).
struct LtBondTable {
    std::vector<LtBond> bonds;
    std::vector<int32_t> begin;  // [16 * 16 + 1]
};
LtBondTable pyrochlore_lt_bonds(double J1, double J2, double J3, double Delta);

}  // namespace pyro
