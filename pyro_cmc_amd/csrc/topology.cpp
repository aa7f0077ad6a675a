#include "topology.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <set>

namespace pyro {

namespace {

inline int floor_mod(int64_t a, int64_t m) {
    int64_t r = a % m;
    return int(r < 0 ? r + m : r);
}
inline int64_t floor_div(int64_t a, int64_t m) {
    int64_t q = a / m;
    if ((a % m != 0) && ((a < 0) != (m < 0))) --q;
    return q;
}

bool matrix_is_zero(const double* J) {
    for (int i = 0; i < 9; ++i)
        if (J[i] != 0.0) return false;
    return true;
}

}  // namespace

ivec3 LatticeGeom::position(int64_t site) const {
    const int64_t nc = n_cells();
    const int sl = int(site / nc);
    int64_t c = site % nc;
    const int n2 = int(c % L[2]);
    c /= L[2];
    const int n1 = int(c % L[1]);
    const int n0 = int(c / L[1]);
    return {sl_pos[sl][0] + cell[0] * n0, sl_pos[sl][1] + cell[1] * n1,
            sl_pos[sl][2] + cell[2] * n2};
}

int64_t LatticeGeom::site_at(const ivec3& p) const {
    for (int s = 0; s < n_sl; ++s) {
        int64_t n[3];
        bool ok = true;
        for (int a = 0; a < 3; ++a) {
            const int64_t d = int64_t(p[a]) - sl_pos[s][a];
            if (floor_mod(d, cell[a]) != 0) { ok = false; break; }
            n[a] = floor_mod(floor_div(d, cell[a]), L[a]);
        }
        if (ok) return int64_t(s) * n_cells() + cell_index(int(n[0]), int(n[1]), int(n[2]));
    }
    return -1;
}

void LatticeGeom::validate() const {
    if (n_sl <= 0 || int(sl_pos.size()) != n_sl || int(sl_kind.size()) != n_sl)
        throw std::invalid_argument("lattice: n_sl does not match sl_pos/sl_kind");
    for (int a = 0; a < 3; ++a) {
        if (cell[a] <= 0) throw std::invalid_argument("lattice: cell edges must be positive");
        if (L[a] <= 0) throw std::invalid_argument("lattice: supercell extent must be positive");
    }
    if (n_sites() > (int64_t(1) << 30))
        throw std::invalid_argument("lattice: more than 2^30 sites");
    // distinct sublattices must sit at distinct positions modulo the cell
    std::set<std::array<int, 3>> seen;
    for (int s = 0; s < n_sl; ++s) {
        std::array<int, 3> r{floor_mod(sl_pos[s][0], cell[0]), floor_mod(sl_pos[s][1], cell[1]),
                             floor_mod(sl_pos[s][2], cell[2])};
        if (!seen.insert(r).second)
            throw std::invalid_argument("lattice: two sublattices share a position");
        if (sl_kind[s] < 0) throw std::invalid_argument("lattice: negative sublattice kind");
    }
}

Topology build_topology(const LatticeGeom& g, const std::vector<CouplingDef>& cs) {
    g.validate();
    const int64_t N = g.n_sites();
    const int64_t nc = g.n_cells();
    Topology t;

    // Per-sublattice slot lists (displacement + matrix), identical for every
    // site of a sublattice because kinds and displacements are translation
    // invariant.  Order: coupling order, then source shell, then target shell.
    struct Slot { ivec3 d; std::array<double, 9> J; };
    std::vector<std::vector<Slot>> slots(g.n_sl);
    for (int s = 0; s < g.n_sl; ++s) {
        const int kind = g.sl_kind[s];
        const int64_t origin = int64_t(s) * nc;  // cell 0 of sublattice s
        const ivec3 p0 = g.position(origin);
        for (const auto& c : cs) {
            if (matrix_is_zero(c.J)) continue;
            if (kind >= int(c.rel.size()))
                throw std::invalid_argument("coupling '" + c.name +
                                            "' has no displacement list for sublattice kind " +
                                            std::to_string(kind));
            std::array<double, 9> J, Jt;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    J[3 * i + j] = c.J[3 * i + j];
                    Jt[3 * i + j] = c.J[3 * j + i];
                }
            std::vector<Slot> src, tgt;
            for (const auto& v : c.rel[kind]) {
                const ivec3 q{p0[0] + v[0], p0[1] + v[1], p0[2] + v[2]};
                const int64_t other = g.site_at(q);
                if (other < 0)
                    throw std::invalid_argument("coupling '" + c.name +
                                                "': displacement does not end on a site");
                const int okind = g.sl_kind[other / nc];
                if (kind <= okind) src.push_back({v, J});
                else tgt.push_back({v, Jt});
            }
            for (auto& x : src) slots[s].push_back(x);
            for (auto& x : tgt) slots[s].push_back(x);
        }
    }
    t.n_slot = 0;
    t.n_slot_sl.resize(g.n_sl);
    for (int s = 0; s < g.n_sl; ++s) {
        t.n_slot_sl[s] = int(slots[s].size());
        t.n_slot = std::max(t.n_slot, t.n_slot_sl[s]);
    }

    // Neighbour table [N][n_slot]; padded slots point at the site itself.
    // By translation invariance slot k of sublattice s always lands on the
    // same sublattice at the same cell offset as it does from cell 0.
    t.nbr.assign(size_t(N) * std::max(t.n_slot, 1), 0);
    for (int s = 0; s < g.n_sl; ++s) {
        const ivec3 p0 = g.position(int64_t(s) * nc);
        std::vector<int> tsl(t.n_slot_sl[s]);
        std::vector<ivec3> off(t.n_slot_sl[s]);
        for (int k = 0; k < t.n_slot_sl[s]; ++k) {
            const ivec3& d = slots[s][k].d;
            const int64_t j0 = g.site_at({p0[0] + d[0], p0[1] + d[1], p0[2] + d[2]});
            tsl[k] = int(j0 / nc);
            int64_t c = j0 % nc;
            off[k][2] = int(c % g.L[2]);
            c /= g.L[2];
            off[k][1] = int(c % g.L[1]);
            off[k][0] = int(c / g.L[1]);
        }
        for (int n0 = 0; n0 < g.L[0]; ++n0)
            for (int n1 = 0; n1 < g.L[1]; ++n1)
                for (int n2 = 0; n2 < g.L[2]; ++n2) {
                    const int64_t i = int64_t(s) * nc + g.cell_index(n0, n1, n2);
                    for (int k = 0; k < t.n_slot; ++k) {
                        int64_t j = i;
                        if (k < t.n_slot_sl[s])
                            j = int64_t(tsl[k]) * nc +
                                g.cell_index((n0 + off[k][0]) % g.L[0], (n1 + off[k][1]) % g.L[1],
                                             (n2 + off[k][2]) % g.L[2]);
                        t.nbr[size_t(i) * t.n_slot + k] = int32_t(j);
                    }
                }
    }

    // Per-slot matrices; isotropy check (J == c*I exactly).
    t.Jslot.assign(size_t(g.n_sl) * t.n_slot * 9, 0.0);
    t.cslot.assign(size_t(g.n_sl) * t.n_slot, 0.0);
    t.isotropic = true;
    for (int s = 0; s < g.n_sl; ++s)
        for (int k = 0; k < t.n_slot_sl[s]; ++k) {
            const auto& J = slots[s][k].J;
            for (int e = 0; e < 9; ++e) t.Jslot[(size_t(s) * t.n_slot + k) * 9 + e] = J[e];
            const bool iso = J[1] == 0 && J[2] == 0 && J[3] == 0 && J[5] == 0 && J[6] == 0 &&
                             J[7] == 0 && J[0] == J[4] && J[4] == J[8];
            if (!iso) t.isotropic = false;
            t.cslot[size_t(s) * t.n_slot + k] = J[0];
        }

    // Colouring.  Try residue periods until every class is an independent set.
    // Candidate periods per axis in {1..4}, smallest class count first.
    std::vector<ivec3> candidates;
    for (int a = 1; a <= 4; ++a)
        for (int b = 1; b <= 4; ++b)
            for (int c = 1; c <= 4; ++c) candidates.push_back({a, b, c});
    std::stable_sort(candidates.begin(), candidates.end(), [](const ivec3& x, const ivec3& y) {
        return x[0] * x[1] * x[2] < y[0] * y[1] * y[2];
    });
    bool found = false;
    for (const auto& per : candidates) {
        if (g.L[0] % per[0] || g.L[1] % per[1] || g.L[2] % per[2]) continue;
        const int P = per[0] * per[1] * per[2];
        const int n_class = g.n_sl * P;
        auto class_of = [&](int64_t i) {
            const int s = int(i / nc);
            int64_t c = i % nc;
            const int n2 = int(c % g.L[2]);
            c /= g.L[2];
            const int n1 = int(c % g.L[1]);
            const int n0 = int(c / g.L[1]);
            return s * P + ((n0 % per[0]) * per[1] + (n1 % per[1])) * per[2] + (n2 % per[2]);
        };
        // Bonds are translation invariant, so the sites of the cells with
        // n_a < period_a represent every class.
        std::vector<std::set<int>> adj(n_class);
        bool ok = true;
        std::vector<int64_t> reps;
        for (int s = 0; s < g.n_sl; ++s)
            for (int a = 0; a < per[0]; ++a)
                for (int b = 0; b < per[1]; ++b)
                    for (int c = 0; c < per[2]; ++c) reps.push_back(int64_t(s) * nc + g.cell_index(a, b, c));
        for (size_t ri = 0; ri < reps.size() && ok; ++ri) {
            const int64_t i = reps[ri];
            const int ci = class_of(i);
            const int s = int(i / nc);
            for (int k = 0; k < t.n_slot_sl[s]; ++k) {
                const int64_t j = t.nbr[size_t(i) * t.n_slot + k];
                const int cj = class_of(j);
                if (cj == ci) { ok = false; break; }
                adj[ci].insert(cj);
            }
        }
        if (!ok) continue;
        // Greedy colouring, largest-degree first (ties by class index).
        std::vector<int> order(n_class);
        for (int c = 0; c < n_class; ++c) order[c] = c;
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return adj[a].size() > adj[b].size(); });
        std::vector<int> col(n_class, -1);
        int n_col = 0;
        for (int c : order) {
            std::set<int> used;
            for (int o : adj[c])
                if (col[o] >= 0) used.insert(col[o]);
            int k = 0;
            while (used.count(k)) ++k;
            col[c] = k;
            n_col = std::max(n_col, k + 1);
        }
        t.period = per;
        t.n_colours = n_col;
        t.class_colour.assign(col.begin(), col.end());
        t.colour_classes.assign(n_col, {});
        for (int c = 0; c < n_class; ++c) t.colour_classes[col[c]].push_back(c);
        found = true;
        break;
    }
    if (!found)
        throw std::invalid_argument("no valid colouring: a sublattice is bonded to itself "
                                    "and the supercell extents admit no residue split");
    return t;
}

// --------------------------------------------------------------------------
// Pyrochlore geometry (pyrochlore_geometry.hpp:12-32, 110-149).
// --------------------------------------------------------------------------

namespace {
const ivec3 kFcc[4] = {{0, 0, 0}, {0, 4, 4}, {4, 0, 4}, {4, 4, 0}};
const ivec3 kLink[4] = {{1, 1, 1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, 1}};
}  // namespace

LatticeGeom pyrochlore_cell(int L0, int L1, int L2) {
    LatticeGeom g;
    g.n_sl = 16;
    for (int f = 0; f < 4; ++f)
        for (int p = 0; p < 4; ++p) {
            g.sl_pos.push_back({kFcc[f][0] + kLink[p][0], kFcc[f][1] + kLink[p][1],
                                kFcc[f][2] + kLink[p][2]});
            g.sl_kind.push_back(p);
        }
    g.cell = {8, 8, 8};
    g.L = {L0, L1, L2};
    return g;
}

std::vector<std::vector<ivec3>> pyrochlore_shell(int which) {
    // Search displacements within the bounding box |d_a| <= 4 on a 4x4x4
    // supercell (large enough that no two displacements alias) and keep those
    // of the requested squared length that end on a site.
    const LatticeGeom g = pyrochlore_cell(4, 4, 4);
    const int want = which == 1 ? 8 : which == 2 ? 24 : 32;
    std::vector<std::vector<ivec3>> out(4);
    for (int p = 0; p < 4; ++p) {
        const ivec3 r = kLink[p];  // site of kind p in FCC cell 0
        for (int dx = -4; dx <= 4; ++dx)
            for (int dy = -4; dy <= 4; ++dy)
                for (int dz = -4; dz <= 4; ++dz) {
                    if (dx * dx + dy * dy + dz * dz != want) continue;
                    if (g.site_at({r[0] + dx, r[1] + dy, r[2] + dz}) < 0) continue;
                    if (which >= 3) {
                        if ((dx | dy | dz) & 1) continue;
                        const bool mid = g.site_at({r[0] + dx / 2, r[1] + dy / 2, r[2] + dz / 2}) >= 0;
                        if ((which == 3) != mid) continue;
                    }
                    out[p].push_back({dx, dy, dz});
                }
    }
    return out;
}

std::vector<CouplingDef> pyrochlore_J1J2J3(double J1, double J2, double J3, double Delta) {
    std::vector<CouplingDef> cs;
    const auto nn1 = pyrochlore_shell(1);
    const LatticeGeom g = pyrochlore_cell(4, 4, 4);
    const double inv_s3 = 1.0 / std::sqrt(3.0);
    const int pairs[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    for (const auto& pr : pairs) {
        const int mu = pr[0], nu = pr[1];
        CouplingDef c;
        c.name = "J1_" + std::to_string(mu) + std::to_string(nu);
        // nn1 bonds of kind mu towards kind nu and back (nn1_pair_MN tables)
        c.rel.assign(4, {});
        for (int p = 0; p < 4; ++p) {
            if (p != mu && p != nu) continue;
            const int want = p == mu ? nu : mu;
            for (const auto& d : nn1[p]) {
                const int64_t j = g.site_at({kLink[p][0] + d[0], kLink[p][1] + d[1], kLink[p][2] + d[2]});
                if (g.sl_kind[j / g.n_cells()] == want) c.rel[p].push_back(d);
            }
        }
        // J1 * [I + (Delta - 1) z_mu z_nu^T], z_p = link_p / sqrt(3)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                const double zz = (inv_s3 * kLink[mu][a]) * (inv_s3 * kLink[nu][b]);
                c.J[3 * a + b] = J1 * (Delta - 1) * zz + J1 * (a == b ? 1.0 : 0.0);
            }
        cs.push_back(std::move(c));
    }
    const char* names[3] = {"J2", "J3a", "J3b"};
    const double Js[3] = {J2, J3, J3};
    for (int k = 0; k < 3; ++k) {
        CouplingDef c;
        c.name = names[k];
        c.rel = pyrochlore_shell(k + 2);
        for (int e = 0; e < 9; ++e) c.J[e] = (e % 4 == 0) ? Js[k] : 0.0;
        cs.push_back(std::move(c));
    }
    return cs;
}

LtBondTable pyrochlore_lt_bonds(double J1, double J2, double J3, double Delta) {
    const LatticeGeom g = pyrochlore_cell(1, 1, 1);
    const int n_sl = g.n_sl;
    const double inv_s3 = 1.0 / std::sqrt(3.0);
    const double Js[4] = {J1, J2, J3, J3};
    std::vector<std::vector<LtBond>> by_block(size_t(n_sl) * n_sl);
    for (int shell = 1; shell <= 4; ++shell) {
        const double j = Js[shell - 1];
        if (j == 0.0) continue;
        const auto rel = pyrochlore_shell(shell);
        for (int a = 0; a < n_sl; ++a) {
            const int mu = g.sl_kind[a];
            for (const auto& v : rel[mu]) {
                const ivec3& r = g.sl_pos[a];
                const int64_t b = g.site_at({r[0] + v[0], r[1] + v[1], r[2] + v[2]});
                if (b < 0) throw std::invalid_argument("pyrochlore shell displacement leaves the lattice");
                const int nu = g.sl_kind[b];
                LtBond bd;
                for (int e = 0; e < 3; ++e) bd.d[e] = double(v[e]);
                for (int x = 0; x < 3; ++x)
                    for (int y = 0; y < 3; ++y) {
                        double m = x == y ? 1.0 : 0.0;
                        if (shell == 1 && Delta != 1.0)
                            m += (Delta - 1.0) * (inv_s3 * kLink[mu][x]) * (inv_s3 * kLink[nu][y]);
                        bd.J[3 * x + y] = j * m;
                    }
                by_block[size_t(a) * n_sl + size_t(b)].push_back(bd);
            }
        }
    }
    LtBondTable t;
    t.begin.push_back(0);
    for (const auto& blk : by_block) {
        t.bonds.insert(t.bonds.end(), blk.begin(), blk.end());
        t.begin.push_back(int32_t(t.bonds.size()));
    }
    return t;
}

}  // namespace pyro
