// bench_speed: throughput of the two lifted-Metropolis sweeps (displacement
// and rotation), each followed by one overrelaxation pass, on the J1-J2
// Heisenberg model of an L x L square lattice.  Mirrors the reference
// harness (test/bench_speed.cpp:97-207): the same flags and defaults, the
// same initial spins (normalised Gaussian draws from std::mt19937_64(seed)),
// warm-up sweeps before the first repetition, variants alternated per
// repetition, best of the repetitions reported as ns per sweep and per spin.
// Here a variant's timed sweeps run on the device as one call (one
// synchronisation), so the figure is device throughput, not host latency.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <random>

#include "cli.hpp"
#include "pyro_host.hpp"

namespace {

using clk = std::chrono::steady_clock;

// One spin per cell of a simple square lattice; J1 on the four nearest and
// J2 on the four diagonal neighbours (test/bench_speed.cpp:36-50).
CMC::UnitCellSpecifier square_cell() {
    CMC::UnitCellSpecifier spec(imat33_t::from_cols({1, 0, 0}, {0, 1, 0}, {0, 0, 1}));
    CMC::HeisenbergSpin s;
    s.ipos = {0, 0, 0};
    s.pyro_sl = 0;
    spec.add(s);
    return spec;
}
const std::vector<std::vector<ipos_t>> kJ1 = {{{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}}};
const std::vector<std::vector<ipos_t>> kJ2 = {{{1, 1, 0}, {1, -1, 0}, {-1, 1, 0}, {-1, -1, 0}}};

void randomise(CMC::Lattice& lat, std::mt19937_64& rng) {
    std::normal_distribution<double> g(0.0, 1.0);
    for (auto& s : lat.get_objects<CMC::HeisenbergSpin>()) {
        const double x = g(rng), y = g(rng), z = g(rng);
        const double n = std::sqrt(x * x + y * y + z * z);
        s.S = vector3::vec3d(x / n, y / n, z / n);
        s.lifted_dir = 1;
    }
}

struct Timing {
    double ns_per_sweep = 1e300, sweeps_per_sec = 0, accept = 0;
};

Timing time_variant(CMC::MC_runner& mc, int kind, double T, size_t n_spins, size_t n_warm, size_t n_sweep) {
    if (n_warm) mc.sweeps(kind, T, 1, n_warm);
    const auto t0 = clk::now();
    const size_t accepted = mc.sweeps(kind, T, 1, n_sweep);
    const auto t1 = clk::now();
    Timing t;
    t.ns_per_sweep = std::chrono::duration<double, std::nano>(t1 - t0).count() / double(n_sweep);
    t.sweeps_per_sec = 1e9 / t.ns_per_sweep;
    t.accept = double(accepted) / (double(n_sweep) * double(n_spins));
    return t;
}

}  // namespace

int main(int argc, char* argv[]) {
    ArgParser prog("bench_speed");
    prog.add({"--J1"}, "Nearest-neighbour Heisenberg coupling").def = {"-1"};
    prog.add({"--J2"}, "Next-nearest-neighbour Heisenberg coupling").def = {"0.3"};
    prog.add({"--T"}, "Temperature").def = {"0.5"};
    prog.add({"--T_ref"}, "Reference temperature of the proposal width").def = {"1"};
    prog.add({"-L", "--L"}, "Linear size of the L x L square supercell").def = {"32"};
    prog.add({"--n_warm"}, "Untimed warm-up sweeps per variant").def = {"2000"};
    prog.add({"--n_sweep"}, "Timed sweeps per variant per repetition").def = {"20000"};
    prog.add({"--reps"}, "Repetitions (variants alternated); best time reported").def = {"3"};
    prog.add({"--seed"}, "Seed").def = {"2718281828"};
    prog.add({"--fp32"}, "FP32 spins and fields (FP64 accumulators)").flag = true;
    try {
        prog.parse(argc, argv);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 2;
    }
    try {
        const double J1 = prog.dbl("--J1"), J2 = prog.dbl("--J2"), T = prog.dbl("--T");
        const double T_ref = prog.dbl("--T_ref");
        const int L = int(prog.integer("--L"));
        const size_t n_warm = size_t(prog.integer("--n_warm")), n_sweep = size_t(prog.integer("--n_sweep"));
        const size_t reps = size_t(prog.integer("--reps")), seed = size_t(prog.integer("--seed"));
        if (L < 1 || n_sweep < 1 || reps < 1) throw std::invalid_argument("L, n_sweep and reps must be positive");

        std::printf("=== lifted Metropolis speed benchmark ===\n");
        std::printf("Model : %dx%d square-lattice J1-J2 Heisenberg (%zu spins)\n", L, L, size_t(L) * size_t(L));
        std::printf("        J1=%g  J2=%g  T=%g  T_ref=%g\n", J1, J2, T, T_ref);
        std::printf("Stats : n_warm=%zu  n_sweep=%zu  reps=%zu  seed=%zu\n\n", n_warm, n_sweep, reps, seed);

        const imat33_t supercell = imat33_t::from_cols({L, 0, 0}, {0, L, 0}, {0, 0, 1});
        const size_t n_spins = size_t(L) * size_t(L);
        const vector3::mat33d I = vector3::mat33d::from_cols({1, 0, 0}, {0, 1, 0}, {0, 0, 1});
        // one lattice and runner per variant, from the same initial spins
        CMC::Lattice lat_d = CMC::build_supercell(square_cell(), supercell);
        CMC::Lattice lat_r = CMC::build_supercell(square_cell(), supercell);
        std::mt19937_64 init(seed);
        randomise(lat_d, init);
        init.seed(seed);
        randomise(lat_r, init);
        auto make = [&](CMC::Lattice& lat) {
            CMC::MC_runner mc(lat, seed + 1);
            mc.settings.precision = prog.flag("--fp32") ? PYRO_FP32 : PYRO_FP64;
            mc.settings.T_ref = T_ref;
            mc.define_coupling("J1", kJ1, J1 * I);
            mc.define_coupling("J2", kJ2, J2 * I);
            mc.setup_lattice();
            return mc;
        };
        CMC::MC_runner mc_d = make(lat_d);
        CMC::MC_runner mc_r = make(lat_r);

        Timing best_d, best_r;
        for (size_t r = 0; r < reps; ++r) {
            Timing d, ro;
            if (r % 2 == 0) {
                d = time_variant(mc_d, PYRO_LIFTED, T, n_spins, r == 0 ? n_warm : 0, n_sweep);
                ro = time_variant(mc_r, PYRO_LIFTED_ROT, T, n_spins, r == 0 ? n_warm : 0, n_sweep);
            } else {
                ro = time_variant(mc_r, PYRO_LIFTED_ROT, T, n_spins, 0, n_sweep);
                d = time_variant(mc_d, PYRO_LIFTED, T, n_spins, 0, n_sweep);
            }
            if (d.ns_per_sweep < best_d.ns_per_sweep) best_d = d;
            if (ro.ns_per_sweep < best_r.ns_per_sweep) best_r = ro;
            std::printf("rep %zu:  displacement %8.1f ns/sweep   rotation %8.1f ns/sweep\n", r, d.ns_per_sweep,
                        ro.ns_per_sweep);
        }
        std::printf("\n%-14s %14s %14s\n", "quantity", "displacement", "rotation");
        std::printf("%-14s %14.1f %14.1f\n", "ns / sweep", best_d.ns_per_sweep, best_r.ns_per_sweep);
        std::printf("%-14s %14.3f %14.3f\n", "ns / spin", best_d.ns_per_sweep / double(n_spins),
                    best_r.ns_per_sweep / double(n_spins));
        std::printf("%-14s %14.3g %14.3g\n", "sweeps / sec", best_d.sweeps_per_sec, best_r.sweeps_per_sec);
        std::printf("%-14s %13.1f%% %13.1f%%\n", "accept", 100 * best_d.accept, 100 * best_r.accept);
        const double ratio = best_r.ns_per_sweep / best_d.ns_per_sweep;
        std::printf("\nFaster: %s  (by %.2fx, best-of-%zu)\n", ratio > 1.0 ? "displacement" : "rotation",
                    ratio > 1.0 ? ratio : 1.0 / ratio, reps);
    } catch (const std::exception& e) {
        std::cerr << "bench_speed: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
