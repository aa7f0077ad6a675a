// oracle.cpp — CPU FP64 restatement of pyro_cmc's sweep-and-measure path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
// never as the thing measured or shipped.  The product (libpyro_cmc.so)
// does not link, load or call it.
//
// It restates the reference's algorithms with the reference's data
// structures (AoS spins, per-site bond shells holding neighbour pointers and
// a pointer to the shell's 3x3 matrix), one replica per instance:
//   lattice + periodic lookup   lattice_lib build_supercell/get_object_at
//                               (as used at MC.cpp:32,38), cell order
//                               (n0*L1 + n1)*L2 + n2 (a convention choice)
//   define_coupling             MC.cpp:56-66
//   setup_lattice (bond shells) MC.cpp:15-52
//   local_field                 MC.cpp:87-96
//   rotate_about_vector         MC.cpp:102-109
//   local_Metropolis            MC.cpp:111-130
//   overrelax_all               MC.cpp:145-153
//   lifted_Metropolis_rot       MC.cpp:194-244
//   lifted_Metropolis           MC.cpp:249-281
//   sweeps                      MC.cpp:155-163, 283-303
//   total_energy_per_unit_cell  MC.cpp:305-312
//   replica exchange decision   temper.cpp:437-473
//   SSF sample (direct DFT)     ssf_manager.hpp:42-57, 260-293
//   Xoroshiro128++ / SplitMix64 XoshiroCpp.hpp:681-702, 1127-1142 (Blackman &
//                               Vigna's published generators)
//   Philox4x32-10               Salmon et al., SC'11 (the product's RNG)
//
// Two sweep modes:
//   "reference" — sequential storage order, Xoroshiro128++ driving libstdc++
//                 normal/uniform/exponential distributions, exactly as the
//                 reference draws them;
//   "counter"   — the same update formulas visiting sites colour by colour
//                 with per-site Philox4x32-10 uniforms (the product's
//                 stream), so a GPU run can be compared draw for draw.
// Both modes guard the measure-zero h_loc = 0 case (the reference divides
// by |h| there): OR keeps the spin, lifted flips its direction.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <deque>
#include <map>
#include <random>
#include <set>
#include <string>
#include <tuple>
#include <vector>

namespace {

// ---------------------------------------------------------------- RNGs
struct SplitMix64 {
    uint64_t x;
    uint64_t operator()() {
        uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
};

struct Xoroshiro128pp {
    using result_type = uint64_t;
    uint64_t s[2];
    explicit Xoroshiro128pp(uint64_t seed = 0) {
        SplitMix64 sm{seed};
        s[0] = sm();
        s[1] = sm();
    }
    static constexpr uint64_t min() { return 0; }
    static constexpr uint64_t max() { return ~uint64_t(0); }
    static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    uint64_t operator()() {
        const uint64_t s0 = s[0];
        uint64_t s1 = s[1];
        const uint64_t r = rotl(s0 + s1, 17) + s0;
        s1 ^= s0;
        s[0] = rotl(s0, 49) ^ s1 ^ (s1 << 21);
        s[1] = rotl(s1, 28);
        return r;
    }
};

struct W4 { uint32_t v[4]; };

W4 philox10(W4 c, uint32_t k0, uint32_t k1) {
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = uint64_t(0xD2511F53u) * c.v[0];
        const uint64_t p1 = uint64_t(0xCD9E8D57u) * c.v[2];
        W4 n;
        n.v[0] = uint32_t(p1 >> 32) ^ c.v[1] ^ k0;
        n.v[1] = uint32_t(p1);
        n.v[2] = uint32_t(p0 >> 32) ^ c.v[3] ^ k1;
        n.v[3] = uint32_t(p0);
        c = n;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// five 24-bit uniforms (m + 1/2) 2^-24 from one block
void uniforms5(const W4& w, double u[5]) {
    const uint32_t m[5] = {w.v[0] >> 8, w.v[1] >> 8, w.v[2] >> 8, w.v[3] >> 8,
                           (w.v[0] & 0xFFu) | ((w.v[1] & 0xFFu) << 8) | ((w.v[2] & 0xFFu) << 16)};
    for (int i = 0; i < 5; ++i) u[i] = (double(m[i]) + 0.5) * 5.9604644775390625e-08;
}

const uint32_t kTagInit = 0x5A17C0DEu;
const uint32_t kTagSwap = 0x7E3B5EEDu;

// ---------------------------------------------------------------- vectors
using V3 = std::array<double, 3>;
inline double dot(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline V3 cross(const V3& a, const V3& b) {
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
inline V3 add(const V3& a, const V3& b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
inline V3 sub(const V3& a, const V3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline V3 scale(double s, const V3& a) { return {s * a[0], s * a[1], s * a[2]}; }
inline V3 matvec(const double* M, const V3& v) {
    return {M[0] * v[0] + M[1] * v[1] + M[2] * v[2], M[3] * v[0] + M[4] * v[1] + M[5] * v[2],
            M[6] * v[0] + M[7] * v[1] + M[8] * v[2]};
}

// ---------------------------------------------------------------- model
struct Spin;
struct BondShell {
    const double* J;
    std::vector<Spin*> bonds;
};
struct Spin {
    std::array<int, 3> ipos;
    int pyro_sl = 0;
    V3 S{0, 0, 0};
    std::vector<BondShell> bond_shells;
    int8_t lifted_dir = 1;
};
struct Coupling {
    std::string name;
    std::vector<std::vector<std::array<int, 3>>> rel;
    double J[9], Jt[9];
};

// rotate_about_vector (MC.cpp:102-109)
void rotate_about_vector(V3& v, const V3& axis, double theta) {
    const double c = std::cos(theta), s = std::sin(theta);
    const double norm2 = dot(axis, axis);
    v = add(add(scale(c, v), scale(((1 - c) / norm2) * dot(axis, v), axis)),
            scale(s / std::sqrt(norm2), cross(axis, v)));
}

struct Draws {
    // reference mode
    Xoroshiro128pp* rng = nullptr;
    std::normal_distribution<double>* normal = nullptr;
    std::exponential_distribution<double>* expd = nullptr;
    std::uniform_real_distribution<double>* rand01 = nullptr;
    // counter mode
    const double* u = nullptr;
    // Conditioning record of the last update, for FP32 error bounds:
    //   [0] the smallest perturbation |dh| of the local field that could
    //       flip one of its decision tests, to first order (the test's
    //       distance from its threshold over its sensitivity to h; +inf
    //       when it takes no decision);
    //   [1] |h_loc|;
    //   [2] sensitivity of the new spin to the local field, |dS'| / |dh| to
    //       first order (0 when the update does not follow h);
    //   [3] amplification of the relative rounding of its own arithmetic
    //       (1 for well-conditioned formulas).
    double info[4] = {HUGE_VAL, 0.0, 0.0, 1.0};
};

struct Oracle {
    int n_sl = 0;
    std::vector<std::array<int, 3>> sl_pos;
    std::vector<int> sl_kind;
    int cell[3] = {0, 0, 0}, L[3] = {0, 0, 0};
    int64_t nc = 0, N = 0;
    std::vector<Spin> spins;
    std::map<std::array<int64_t, 3>, int64_t> where;
    std::deque<Coupling> couplings;  // stable addresses for BondShell::J
    V3 B{0, 0, 0};
    double T_ref = 1.0;
    Xoroshiro128pp rng;
    std::normal_distribution<double> normal{0.0, 1.0};
    std::exponential_distribution<double> expd{1.0};
    std::uniform_real_distribution<double> rand01{0.0, 1.0};

    std::array<int64_t, 3> reduce(const std::array<int, 3>& p) const {
        std::array<int64_t, 3> r;
        for (int a = 0; a < 3; ++a) {
            const int64_t per = int64_t(cell[a]) * L[a];
            r[a] = ((int64_t(p[a]) % per) + per) % per;
        }
        return r;
    }
    Spin* at(const std::array<int, 3>& p) {
        auto it = where.find(reduce(p));
        return it == where.end() ? nullptr : &spins[it->second];
    }

    // MC_runner::local_field (MC.cpp:87-96)
    V3 local_field(const Spin* s) const {
        V3 h{0, 0, 0};
        for (const auto& shell : s->bond_shells) {
            V3 tmp{0, 0, 0};
            for (const Spin* o : shell.bonds) tmp = add(tmp, o->S);
            h = add(h, matvec(shell.J, tmp));
        }
        return h;
    }

    // local_Metropolis (MC.cpp:111-130)
    int metropolis(double T, Spin* s, Draws& d) {
        const V3 h = sub(local_field(s), B);
        const double curE = dot(s->S, h);
        V3 n;
        double U;
        if (d.u) {
            const double r1 = std::sqrt(-2.0 * std::log(d.u[0]));
            const double t1 = 2 * M_PI * d.u[1];
            const double r2 = std::sqrt(-2.0 * std::log(d.u[2]));
            const double t2 = 2 * M_PI * d.u[3];
            n = {r1 * std::cos(t1), r1 * std::sin(t1), r2 * std::cos(t2)};
            U = d.u[4];
        } else {
            n[0] = (*d.normal)(*d.rng);
            n[1] = (*d.normal)(*d.rng);
            n[2] = (*d.normal)(*d.rng);
            U = -1.0;  // drawn only if needed, as the reference does
        }
        V3 ns = add(scale(std::sqrt(T / T_ref), n), s->S);
        const double nrm = std::sqrt(dot(ns, ns));
        ns = {ns[0] / nrm, ns[1] / nrm, ns[2] / nrm};
        const double dE = dot(ns, h) - curE;
        const double dS = std::sqrt(dot(sub(ns, s->S), sub(ns, s->S)));  // d dE / dh
        d.info[0] = std::abs(dE) / dS;
        d.info[1] = std::sqrt(dot(h, h));
        if (dE < 0) {
            s->S = ns;
            return 1;
        }
        if (U < 0) U = (*d.rand01)(*d.rng);
        {
            const double p = std::exp(-dE / T);
            d.info[0] = std::min(d.info[0], std::abs(U - p) / (p * dS / T));
        }
        if (U < std::exp(-dE / T)) {
            s->S = ns;
            return 1;
        }
        return 0;
    }

    // overrelax step of overrelax_all (MC.cpp:145-153)
    void overrelax(Spin* s, Draws& d) {
        const V3 h = sub(local_field(s), B);
        const double U = d.u ? d.u[0] : (*d.rand01)(*d.rng);
        if (!(dot(h, h) > 0)) return;
        d.info[1] = std::sqrt(dot(h, h));
        d.info[2] = 1.0 / d.info[1];  // the axis direction moves by |dh| / |h|
        rotate_about_vector(s->S, h, 2 * M_PI * U);
    }

    // local_lifted_Metropolis (MC.cpp:249-281)
    int lifted(double T, Spin* s, Draws& d) {
        const V3 h = sub(local_field(s), B);
        const double curE = dot(s->S, h);
        const double hn = std::sqrt(dot(h, h));
        const double E1 = d.u ? -std::log(d.u[0]) : (*d.expd)(*d.rng);
        if (!(hn > 0)) {
            s->lifted_dir = int8_t(-s->lifted_dir);
            return 0;
        }
        const V3 hh = {h[0] / hn, h[1] / hn, h[2] / hn};
        const double u = dot(hh, s->S);
        const double du = s->lifted_dir * std::sqrt(T / T_ref) * E1;
        const double nu = u + du;
        // u moves by at most |dh| / |h|
        d.info[0] = std::min(std::abs(1 - nu), std::abs(nu + 1)) * hn;
        d.info[1] = hn;
        // S' = u' h^ + sqrt(1 - u'^2) v^ with v^ = (S - u h^) / |S - u h^|:
        // u' moves by |dh| / |h|, which moves S' by that over sqrt(1 - u'^2),
        // and v^ by |dh| / (|h| sin theta); the ratio (1 - u'^2) / (1 - u^2)
        // carries the rounding of 1 - u^2 amplified by 1 / sin^2 theta
        // (recorded only if the move is accepted: a rejected spin stays put)
        const double s2 = std::max(std::abs(1 - u * u), 1e-300);
        const double c2 = std::max(std::abs(1 - nu * nu), 1e-300);
        const double sens_h = (1.0 / std::sqrt(c2) + std::sqrt(c2 / s2)) / hn;
        if (nu > 1 || nu < -1 || std::abs(1 - u * u) < 1e-16) {
            s->lifted_dir = int8_t(-s->lifted_dir);
            return 0;
        }
        V3 ns = add(scale(nu, hh), scale(std::sqrt((1 - nu * nu) / (1 - u * u)), sub(s->S, scale(u, hh))));
        const double nrm = std::sqrt(dot(ns, ns));
        ns = {ns[0] / nrm, ns[1] / nrm, ns[2] / nrm};
        const double dE = dot(ns, h) - curE;
        const double U = d.u ? d.u[1] : (*d.rand01)(*d.rng);
        {
            // dE = (S' - S) . h moves by |S' - S| |dh| + |h| |dS'|
            const double p = std::exp(-dE / T);
            const double sens = std::sqrt(dot(sub(ns, s->S), sub(ns, s->S))) + hn * sens_h;
            d.info[0] = std::min(d.info[0], std::abs(U - p) / (p * sens / T));
        }
        if (U < std::exp(-dE / T)) {
            d.info[2] = sens_h;
            d.info[3] = 1.0 / s2 + 1.0 / std::sqrt(c2);
            s->S = ns;
            return 1;
        }
        s->lifted_dir = int8_t(-s->lifted_dir);
        return 0;
    }

    // local_lifted_Metropolis_rot (MC.cpp:194-244)
    int lifted_rot(double T, Spin* s, Draws& d) {
        const V3 h = sub(local_field(s), B);
        const double curE = dot(s->S, h);
        const double delta = std::sqrt(T / T_ref);
        const V3 axis = cross(s->S, h);
        const double a2 = dot(axis, axis);
        d.info[1] = std::sqrt(dot(h, h));
        if (a2 <= 1e-20) {
            s->lifted_dir = int8_t(-s->lifted_dir);
            return 0;
        }
        const double sens_h = 1.0 / std::sqrt(a2);  // the rotation axis S x h moves by |dh| / |S x h|
        V3 ns = s->S;
        rotate_about_vector(ns, axis, s->lifted_dir * delta);
        const double nrm = std::sqrt(dot(ns, ns));
        ns = {ns[0] / nrm, ns[1] / nrm, ns[2] / nrm};
        const V3 axis_new = cross(ns, h);
        const double b2 = dot(axis_new, axis_new);
        // a . b with a = S x h, b = S' x h moves by (|a| + |b|) |dh|
        d.info[0] = std::abs(dot(axis, axis_new)) / (std::sqrt(a2) + std::sqrt(b2));
        if (dot(axis, axis_new) < 0) {
            s->lifted_dir = int8_t(-s->lifted_dir);
            return 0;
        }
        const double dE = dot(ns, h) - curE;
        const double ratio = std::sqrt(b2 / a2) * std::exp(-dE / T);
        const double U = d.u ? d.u[0] : (*d.rand01)(*d.rng);
        {
            // log ratio moves by (1 / |a| + 1 / |b| + |S' - S| / T) |dh|
            const double sens = 1.0 / std::sqrt(a2) + 1.0 / std::sqrt(std::max(b2, 1e-300)) +
                                std::sqrt(dot(sub(ns, s->S), sub(ns, s->S))) / T;
            d.info[0] = std::min(d.info[0], std::abs(U - ratio) / (ratio * sens));
        }
        if (U < ratio) {
            d.info[2] = sens_h;  // an accepted move follows the axis
            s->S = ns;
            return 1;
        }
        s->lifted_dir = int8_t(-s->lifted_dir);
        return 0;
    }

    int update(int kind, double T, Spin* s, Draws& d) {
        d.info[0] = HUGE_VAL;
        d.info[1] = 0.0;
        d.info[2] = 0.0;
        d.info[3] = 1.0;
        switch (kind) {
            case 0: return metropolis(T, s, d);
            case 1: return lifted(T, s, d);
            case 2: return lifted_rot(T, s, d);
            default: overrelax(s, d); return 0;
        }
    }

    Draws ref_draws() {
        Draws d;
        d.rng = &rng;
        d.normal = &normal;
        d.expd = &expd;
        d.rand01 = &rand01;
        return d;
    }

    // total_energy_per_unit_cell (MC.cpp:305-312)
    double energy_per_cell() const {
        double E = 0;
        for (const auto& s : spins) {
            E += 0.5 * dot(s.S, local_field(&s));
            E -= dot(s.S, B);
        }
        return E / double(nc);
    }
};

}  // namespace

extern "C" {

void* orc_create(int n_sl, const int32_t* sl_pos, const int32_t* sl_kind, const int32_t* cell,
                 const int32_t* L, uint64_t seed) {
    auto* o = new Oracle();
    o->n_sl = n_sl;
    for (int s = 0; s < n_sl; ++s) {
        o->sl_pos.push_back({sl_pos[3 * s], sl_pos[3 * s + 1], sl_pos[3 * s + 2]});
        o->sl_kind.push_back(sl_kind[s]);
    }
    for (int a = 0; a < 3; ++a) {
        o->cell[a] = cell[a];
        o->L[a] = L[a];
    }
    o->nc = int64_t(L[0]) * L[1] * L[2];
    o->N = o->nc * n_sl;
    o->spins.resize(o->N);
    for (int s = 0; s < n_sl; ++s)
        for (int n0 = 0; n0 < L[0]; ++n0)
            for (int n1 = 0; n1 < L[1]; ++n1)
                for (int n2 = 0; n2 < L[2]; ++n2) {
                    const int64_t i = s * o->nc + (int64_t(n0) * L[1] + n1) * L[2] + n2;
                    Spin& sp = o->spins[i];
                    sp.ipos = {sl_pos[3 * s] + cell[0] * n0, sl_pos[3 * s + 1] + cell[1] * n1,
                               sl_pos[3 * s + 2] + cell[2] * n2};
                    sp.pyro_sl = sl_kind[s];
                    o->where[o->reduce(sp.ipos)] = i;
                }
    o->rng = Xoroshiro128pp(seed);
    return o;
}

void orc_destroy(void* h) { delete static_cast<Oracle*>(h); }

int64_t orc_num_sites(void* h) { return static_cast<Oracle*>(h)->N; }

// define_coupling (MC.cpp:56-66); returns -1 on a duplicate name.
int orc_define_coupling(void* h, const char* name, int n_kinds, const int32_t* counts,
                        const int32_t* rel, const double* J) {
    auto* o = static_cast<Oracle*>(h);
    for (const auto& c : o->couplings)
        if (c.name == name) return -1;
    Coupling c;
    c.name = name;
    size_t off = 0;
    for (int k = 0; k < n_kinds; ++k) {
        std::vector<std::array<int, 3>> v;
        for (int i = 0; i < counts[k]; ++i, ++off)
            v.push_back({rel[3 * off], rel[3 * off + 1], rel[3 * off + 2]});
        c.rel.push_back(v);
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            c.J[3 * i + j] = J[3 * i + j];
            c.Jt[3 * i + j] = J[3 * j + i];
        }
    o->couplings.push_back(c);
    return 0;
}

// Replace the matrix of the coupling called `name` (the bond shells point at
// it, so no new setup is needed); -1 if there is no such coupling.
int orc_set_coupling(void* h, const char* name, const double* J) {
    auto* o = static_cast<Oracle*>(h);
    for (auto& c : o->couplings)
        if (c.name == name) {
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    c.J[3 * i + j] = J[3 * i + j];
                    c.Jt[3 * i + j] = J[3 * j + i];
                }
            return 0;
        }
    return -1;
}

// setup_lattice (MC.cpp:15-52).  Returns -1 if a displacement misses a site
// or a displacement list is missing.
int orc_setup(void* h) {
    auto* o = static_cast<Oracle*>(h);
    for (const auto& c : o->couplings) {
        for (int sl = 0; sl < o->n_sl; ++sl) {
            const int kind = o->sl_kind[sl];
            if (kind >= int(c.rel.size())) return -1;
            for (int64_t cell = 0; cell < o->nc; ++cell) {
                Spin* origin = &o->spins[sl * o->nc + cell];
                std::vector<Spin*> src, tgt;
                for (const auto& v : c.rel[kind]) {
                    Spin* other = o->at({origin->ipos[0] + v[0], origin->ipos[1] + v[1], origin->ipos[2] + v[2]});
                    if (!other) return -1;
                    if (origin->pyro_sl <= other->pyro_sl) src.push_back(other);
                    else tgt.push_back(other);
                }
                if (!src.empty()) origin->bond_shells.push_back({c.J, src});
                if (!tgt.empty()) origin->bond_shells.push_back({c.Jt, tgt});
            }
        }
    }
    return 0;
}

void orc_set_field(void* h, const double* B) {
    auto* o = static_cast<Oracle*>(h);
    o->B = {B[0], B[1], B[2]};
}
void orc_set_T_ref(void* h, double T_ref) { static_cast<Oracle*>(h)->T_ref = T_ref; }

void orc_set_spins(void* h, const double* S) {
    auto* o = static_cast<Oracle*>(h);
    for (int64_t i = 0; i < o->N; ++i) o->spins[i].S = {S[3 * i], S[3 * i + 1], S[3 * i + 2]};
}
void orc_get_spins(void* h, double* S) {
    auto* o = static_cast<Oracle*>(h);
    for (int64_t i = 0; i < o->N; ++i)
        for (int a = 0; a < 3; ++a) S[3 * i + a] = o->spins[i].S[a];
}
void orc_set_dir(void* h, const int8_t* d) {
    auto* o = static_cast<Oracle*>(h);
    for (int64_t i = 0; i < o->N; ++i) o->spins[i].lifted_dir = d[i];
}
void orc_get_dir(void* h, int8_t* d) {
    auto* o = static_cast<Oracle*>(h);
    for (int64_t i = 0; i < o->N; ++i) d[i] = o->spins[i].lifted_dir;
}
void orc_positions(void* h, int32_t* pos) {
    auto* o = static_cast<Oracle*>(h);
    for (int64_t i = 0; i < o->N; ++i)
        for (int a = 0; a < 3; ++a) pos[3 * i + a] = o->spins[i].ipos[a];
}

// Bonded neighbours of every site as (site, neighbour) index pairs, one per
// bond-shell entry (so a site bonded twice to the same neighbour appears
// twice).  Returns the number of pairs; writes at most cap pairs.
int64_t orc_bonds(void* h, int64_t* pairs, int64_t cap) {
    auto* o = static_cast<Oracle*>(h);
    int64_t n = 0;
    for (int64_t i = 0; i < o->N; ++i)
        for (const auto& sh : o->spins[i].bond_shells)
            for (const Spin* b : sh.bonds) {
                if (n < cap) {
                    pairs[2 * n] = i;
                    pairs[2 * n + 1] = b - o->spins.data();
                }
                ++n;
            }
    return n;
}

void orc_local_fields(void* h, double* out) {
    auto* o = static_cast<Oracle*>(h);
    for (int64_t i = 0; i < o->N; ++i) {
        const V3 f = o->local_field(&o->spins[i]);
        for (int a = 0; a < 3; ++a) out[3 * i + a] = f[a];
    }
}

double orc_energy_per_cell(void* h) { return static_cast<Oracle*>(h)->energy_per_cell(); }

// Reference-mode sweep (MC.cpp:155-163, 283-303): kind 0/1/2 over all sites
// in storage order, then n_or OR passes.  Returns the accepted count.
uint64_t orc_sweep_ref(void* h, int kind, double T, int n_or) {
    auto* o = static_cast<Oracle*>(h);
    Draws d = o->ref_draws();
    uint64_t acc = 0;
    for (auto& s : o->spins) acc += o->update(kind, T, &s, d);
    for (int p = 0; p < n_or; ++p)
        for (auto& s : o->spins) o->update(3, T, &s, d);
    return acc;
}

// Counter-mode pass: one pass of `kind` (0..3) visiting colours 0..n_col-1
// in order (sites of a colour in storage order) with Philox uniforms keyed
// by (seed ^ tag, seed >> 32) and counter (site, replica, pass).
// info (may be NULL) receives per site the conditioning record of its
// update, info[4 i .. 4 i + 3] (Draws::info).
uint64_t orc_pass_ctr_m(void* h, int kind, double T, const int32_t* colour_of_site, int n_col,
                        uint64_t seed, uint32_t replica, uint64_t pass, double* info) {
    auto* o = static_cast<Oracle*>(h);
    uint64_t acc = 0;
    const uint32_t k0 = uint32_t(seed), k1 = uint32_t(seed >> 32);
    for (int c = 0; c < n_col; ++c)
        for (int64_t i = 0; i < o->N; ++i) {
            if (colour_of_site[i] != c) continue;
            const W4 w = philox10(W4{{uint32_t(i), replica, uint32_t(pass), uint32_t(pass >> 32)}}, k0, k1);
            double u[5];
            uniforms5(w, u);
            Draws d;
            d.u = u;
            acc += o->update(kind, T, &o->spins[i], d);
            if (info) std::copy(d.info, d.info + 4, info + 4 * i);
        }
    return acc;
}

uint64_t orc_pass_ctr(void* h, int kind, double T, const int32_t* colour_of_site, int n_col,
                      uint64_t seed, uint32_t replica, uint64_t pass) {
    return orc_pass_ctr_m(h, kind, T, colour_of_site, n_col, seed, replica, pass, nullptr);
}

// Counter-mode updates in an explicit order: update k visits site[k] with
// kind first_kind if pass_offset[k] == 0 and overrelaxation otherwise, pass
// index pass0 + pass_offset[k].  Returns the accepted count of the first kind.
// info (may be NULL) receives the conditioning record of update k at
// info[4 k .. 4 k + 3].  shared != 0: the sweep's uniforms of a site are
// drawn in order from its blocks (the product's fused passes): uniform j is
// component j % 5 of block pass0 + j / 5, the first kind taking 5 / 2 / 1
// (Metropolis / lifted / lifted rotation) and OR pass o >= 1 uniform
// (first kind's count) + o - 1; else every pass draws its own block.
uint64_t orc_sweep_order_m(void* h, int first_kind, double T, const int32_t* site,
                           const int32_t* pass_offset, int64_t n, uint64_t seed, uint32_t replica,
                           uint64_t pass0, double* info, int shared) {
    auto* o = static_cast<Oracle*>(h);
    uint64_t acc = 0;
    const uint32_t k0 = uint32_t(seed), k1 = uint32_t(seed >> 32);
    const int n_first = first_kind == 0 ? 5 : (first_kind == 1 ? 2 : 1);
    for (int64_t k = 0; k < n; ++k) {
        const int64_t i = site[k];
        uint64_t pass = pass0 + uint64_t(pass_offset[k]);
        int comp = 0;
        if (shared && pass_offset[k] > 0) {
            const int j = n_first + pass_offset[k] - 1;
            pass = pass0 + uint64_t(j / 5);
            comp = j % 5;
        }
        const W4 w = philox10(W4{{uint32_t(i), replica, uint32_t(pass), uint32_t(pass >> 32)}}, k0, k1);
        double u[5];
        uniforms5(w, u);
        u[0] = u[comp];
        Draws d;
        d.u = u;
        const int kind = pass_offset[k] == 0 ? first_kind : 3;
        const int a = o->update(kind, T, &o->spins[i], d);
        if (info) std::copy(d.info, d.info + 4, info + 4 * k);
        if (pass_offset[k] == 0) acc += a;
    }
    return acc;
}

uint64_t orc_sweep_order(void* h, int first_kind, double T, const int32_t* site,
                         const int32_t* pass_offset, int64_t n, uint64_t seed, uint32_t replica,
                         uint64_t pass0) {
    return orc_sweep_order_m(h, first_kind, T, site, pass_offset, n, seed, replica, pass0, nullptr, 0);
}

// Uniform-on-sphere initial spins of replica r (z = 2u0 - 1, phi = 2 pi u1).
void orc_randomize_ctr(void* h, uint64_t seed, uint32_t replica) {
    auto* o = static_cast<Oracle*>(h);
    const uint32_t k0 = uint32_t(seed) ^ kTagInit, k1 = uint32_t(seed >> 32);
    for (int64_t i = 0; i < o->N; ++i) {
        const W4 w = philox10(W4{{uint32_t(i), replica, 0u, 0u}}, k0, k1);
        double u[5];
        uniforms5(w, u);
        const double z = 2.0 * u[0] - 1.0, phi = 6.283185307179586 * u[1];
        const double s = std::sqrt(1.0 - z * z);
        o->spins[i].S = {s * std::cos(phi), s * std::sin(phi), z};
        o->spins[i].lifted_dir = 1;
    }
}

// Raw Philox block (for known-answer tests).
void orc_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    const W4 w = philox10(W4{{ctr[0], ctr[1], ctr[2], ctr[3]}}, key[0], key[1]);
    for (int i = 0; i < 4; ++i) out[i] = w.v[i];
}

// First n outputs of Xoroshiro128++ seeded through SplitMix64.
void orc_xoroshiro(uint64_t seed, int n, uint64_t* out) {
    Xoroshiro128pp g(seed);
    for (int i = 0; i < n; ++i) out[i] = g();
}

// Replica exchange round (temper.cpp:437-473) on slot-indexed arrays:
// pairs (s, s+1) with s = round parity; total energies e[r] * n_cells.
void orc_exchange(int R, const double* T_slot, const double* e, double n_cells, int32_t* rep_of,
                  uint64_t round, uint64_t swap_seed, uint64_t* att, uint64_t* accp) {
    const uint32_t k0 = uint32_t(swap_seed) ^ kTagSwap, k1 = uint32_t(swap_seed >> 32);
    for (int s = int(round & 1); s + 1 < R; s += 2) {
        const int ra = rep_of[s], rb = rep_of[s + 1];
        const double arg = (1.0 / T_slot[s] - 1.0 / T_slot[s + 1]) * (e[ra] * n_cells - e[rb] * n_cells);
        const W4 w = philox10(W4{{uint32_t(s), 0xFFFFFFFFu, uint32_t(round), uint32_t(round >> 32)}}, k0, k1);
        const uint64_t m = (uint64_t(w.v[0]) << 21) ^ uint64_t(w.v[1] >> 11);
        const double U = (double(m) + 0.5) * 1.1102230246251565e-16;
        att[s] += 1;
        if (arg >= 0.0 || U < std::exp(arg)) {
            accp[s] += 1;
            rep_of[s] = rb;
            rep_of[s + 1] = ra;
        }
    }
}

// Raw per-sublattice DFT  A^a_mu(K) = sum_n e^{-2 pi i K.n / L} S^a_mu(n),
// ft[3][n_sl][n_k][2], k order (K0*L1 + K1)*L2 + K2.  Direct O(N n_k) sum.
void orc_spin_ft(void* h, double* ft) {
    auto* o = static_cast<Oracle*>(h);
    const int64_t nc = o->nc;
    for (int a = 0; a < 3; ++a)
        for (int mu = 0; mu < o->n_sl; ++mu)
            for (int64_t K = 0; K < nc; ++K) {
                const int64_t K2 = K % o->L[2], K1 = (K / o->L[2]) % o->L[1], K0 = K / (int64_t(o->L[1]) * o->L[2]);
                double re = 0, im = 0;
                for (int64_t c = 0; c < nc; ++c) {
                    const int64_t n2 = c % o->L[2], n1 = (c / o->L[2]) % o->L[1], n0 = c / (int64_t(o->L[1]) * o->L[2]);
                    const double ph = -2 * M_PI *
                                      (double(K0 * n0) / o->L[0] + double(K1 * n1) / o->L[1] + double(K2 * n2) / o->L[2]);
                    const double x = o->spins[mu * nc + c].S[a];
                    re += x * std::cos(ph);
                    im += x * std::sin(ph);
                }
                ft[2 * ((a * o->n_sl + mu) * nc + K)] = re;
                ft[2 * ((a * o->n_sl + mu) * nc + K) + 1] = im;
            }
}

// One SSF sample (ssf_manager.hpp:260-293): for each correlator (alpha,beta)
// S(K) = sum_{mu,nu} w_{mu nu}(K) conj(A^alpha_mu(K)) A^beta_nu(K) with
// w_{mu nu} = exp(+i q(K).(r_mu - r_nu)), q_a = 2 pi K'_a / (cell_a L_a),
// K' the centred representative of K.  out[n_corr][n_k][2].
// If munu is non-null it receives C_{mu nu}(K) for each correlator,
// [n_corr][n_k][n_sl][n_sl][2] (the temper layout, temper.cpp:98-159).
void orc_ssf_sample(void* h, int n_corr, const int32_t* alpha, const int32_t* beta, double* out,
                    double* munu) {
    auto* o = static_cast<Oracle*>(h);
    const int64_t nc = o->nc;
    const int ns = o->n_sl;
    std::vector<double> ft(size_t(3) * ns * nc * 2);
    orc_spin_ft(h, ft.data());
    for (int c = 0; c < n_corr; ++c)
        for (int64_t K = 0; K < nc; ++K) {
            const int64_t Ks[3] = {K / (int64_t(o->L[1]) * o->L[2]), (K / o->L[2]) % o->L[1], K % o->L[2]};
            double q[3];
            for (int a = 0; a < 3; ++a) {
                const int64_t Kc = (2 * Ks[a] < o->L[a]) ? Ks[a] : Ks[a] - o->L[a];
                q[a] = 2 * M_PI * double(Kc) / double(o->cell[a] * o->L[a]);
            }
            double sre = 0, sim = 0;
            for (int mu = 0; mu < ns; ++mu)
                for (int nu = 0; nu < ns; ++nu) {
                    const double* A = &ft[2 * ((alpha[c] * ns + mu) * nc + K)];
                    const double* Bv = &ft[2 * ((beta[c] * ns + nu) * nc + K)];
                    // conj(A) * B
                    const double cr = A[0] * Bv[0] + A[1] * Bv[1];
                    const double ci = A[0] * Bv[1] - A[1] * Bv[0];
                    if (munu) {
                        double* m = &munu[2 * ((((size_t(c) * nc + K) * ns) + mu) * ns + nu)];
                        m[0] = cr;
                        m[1] = ci;
                    }
                    double ph = 0;
                    for (int a = 0; a < 3; ++a) ph += q[a] * double(o->sl_pos[mu][a] - o->sl_pos[nu][a]);
                    const double wr = std::cos(ph), wi = std::sin(ph);
                    sre += wr * cr - wi * ci;
                    sim += wr * ci + wi * cr;
                }
            out[2 * (c * nc + K)] = sre;
            out[2 * (c * nc + K) + 1] = sim;
        }
}

}  // extern "C"
