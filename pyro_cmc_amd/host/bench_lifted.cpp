// bench_lifted: lifted and standard Metropolis sample the same distribution.
// Mirrors the reference harness (test/bench_lifted.cpp:124-272): the J1-J2
// Heisenberg model on an L x L square lattice, the same flags and defaults,
// SplitMix64 fan-out of the base seed into four streams (initial spins and
// sampler seed per sampler), n_burn sweeps then n_sample sweeps with one
// energy measurement each, batch-means errors, and PASS when both |<E>| and
// |<E^2>| of the two samplers differ by less than tol_frac * min(|J1|,|J2|)
// (exit status 0, else 1).  Each sweep is one pass of the sampler plus one
// overrelaxation pass, as sweep_local_Metropolis / sweep_lifted_Metropolis
// default to.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <iostream>
#include <random>

#include "cli.hpp"
#include "pyro_host.hpp"

namespace {

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

// SplitMix64 (Steele, Lea, Flood 2014): next output of the state x.
uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Standard error of the mean from n_batch contiguous batch means.
double batch_stderr(const std::vector<double>& x, size_t n_batch) {
    if (x.size() < n_batch || n_batch < 2) return 0.0;
    const size_t bs = x.size() / n_batch;
    std::vector<double> m(n_batch, 0.0);
    for (size_t b = 0; b < n_batch; ++b) {
        for (size_t i = 0; i < bs; ++i) m[b] += x[b * bs + i];
        m[b] /= double(bs);
    }
    double mean = 0.0;
    for (double v : m) mean += v;
    mean /= double(n_batch);
    double var = 0.0;
    for (double v : m) var += (v - mean) * (v - mean);
    return std::sqrt(var / double(n_batch - 1) / double(n_batch));
}

struct Estimates {
    double E = 0, E2 = 0, E_err = 0, E2_err = 0, accept = 0;
};

Estimates run_sampler(CMC::MC_runner& mc, int kind, double T, size_t n_spins, size_t n_burn, size_t n_sample) {
    if (n_burn) mc.sweeps(kind, T, 1, n_burn);
    std::vector<double> e(n_sample), e2(n_sample);
    size_t accepted = 0;
    for (size_t i = 0; i < n_sample; ++i) {
        accepted += mc.sweeps(kind, T, 1, 1);
        e[i] = mc.total_energy_per_unit_cell();  // one site per cell: per spin
        e2[i] = e[i] * e[i];
    }
    Estimates est;
    for (size_t i = 0; i < n_sample; ++i) {
        est.E += e[i];
        est.E2 += e2[i];
    }
    est.E /= double(n_sample);
    est.E2 /= double(n_sample);
    const size_t n_batch = std::min<size_t>(50, std::max<size_t>(2, n_sample / 20));
    est.E_err = batch_stderr(e, n_batch);
    est.E2_err = batch_stderr(e2, n_batch);
    est.accept = double(accepted) / (double(n_sample) * double(n_spins));
    return est;
}

}  // namespace

int main(int argc, char* argv[]) {
    ArgParser prog("bench_lifted");
    prog.add({"--J1"}, "Nearest-neighbour Heisenberg coupling (default -1, ferromagnetic)").def = {"-1"};
    prog.add({"--J2"}, "Next-nearest-neighbour Heisenberg coupling (default +0.3)").def = {"0.3"};
    prog.add({"--T"}, "Temperature at which to compare the two samplers").def = {"0.5"};
    prog.add({"--T_ref"}, "Reference temperature for the proposal distribution").def = {"1"};
    prog.add({"-L", "--L"}, "Linear size of the L x L square supercell").def = {"8"};
    prog.add({"--n_burn"}, "Thermalisation sweeps before sampling").def = {"5000"};
    prog.add({"--n_sample"}, "Measurement sweeps (one energy measurement per sweep)").def = {"50000"};
    prog.add({"--tol_frac"}, "Tolerance as a fraction of min(|J1|,|J2|)").def = {"0.1"};
    prog.add({"--seed"}, "Base RNG seed").def = {"2718281828"};
    prog.add({"--fp32"}, "FP32 spins and fields (FP64 accumulators)").flag = true;
    try {
        prog.parse(argc, argv);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 2;
    }
    try {
        const double J1 = prog.dbl("--J1"), J2 = prog.dbl("--J2"), T = prog.dbl("--T");
        const double T_ref = prog.dbl("--T_ref"), tol_frac = prog.dbl("--tol_frac");
        const int L = int(prog.integer("--L"));
        const size_t n_burn = size_t(prog.integer("--n_burn")), n_sample = size_t(prog.integer("--n_sample"));
        const size_t seed = size_t(prog.integer("--seed"));
        if (L < 1 || n_sample < 1) throw std::invalid_argument("L and n_sample must be positive");
        const double jmin = std::min(std::abs(J1), std::abs(J2));
        const double tol = tol_frac * jmin;

        std::printf("=== lifted vs. standard Metropolis consistency benchmark ===\n");
        std::printf("Model : %dx%d square-lattice J1-J2 Heisenberg\n", L, L);
        std::printf("        J1=%g  J2=%g  T=%g  T_ref=%g\n", J1, J2, T, T_ref);
        std::printf("Stats : n_burn=%zu  n_sample=%zu  seed=%zu\n", n_burn, n_sample, seed);
        std::printf("Tol   : %g  (= %g * min(|J1|,|J2|)=%g)\n\n", tol, tol_frac, jmin);

        uint64_t sm = seed;
        const uint64_t seed_std_init = splitmix64(sm), seed_std_mc = splitmix64(sm);
        const uint64_t seed_lft_init = splitmix64(sm), seed_lft_mc = splitmix64(sm);
        const imat33_t supercell = imat33_t::from_cols({L, 0, 0}, {0, L, 0}, {0, 0, 1});
        const size_t n_spins = size_t(L) * size_t(L);
        const vector3::mat33d I = vector3::mat33d::from_cols({1, 0, 0}, {0, 1, 0}, {0, 0, 1});
        auto run = [&](uint64_t init_seed, uint64_t mc_seed, int kind, const char* what) {
            CMC::Lattice lat = CMC::build_supercell(square_cell(), supercell);
            std::mt19937_64 init(init_seed);
            randomise(lat, init);
            CMC::MC_runner mc(lat, mc_seed);
            mc.settings.precision = prog.flag("--fp32") ? PYRO_FP32 : PYRO_FP64;
            mc.settings.T_ref = T_ref;
            mc.define_coupling("J1", kJ1, J1 * I);
            mc.define_coupling("J2", kJ2, J2 * I);
            mc.setup_lattice();
            std::printf("Running %s Metropolis...\n", what);
            return run_sampler(mc, kind, T, n_spins, n_burn, n_sample);
        };
        const Estimates s = run(seed_std_init, seed_std_mc, PYRO_METROPOLIS, "standard");
        const Estimates l = run(seed_lft_init, seed_lft_mc, PYRO_LIFTED, "lifted");

        const double dE = std::abs(s.E - l.E), dE2 = std::abs(s.E2 - l.E2);
        std::printf("\n%-10s %14s %14s %14s\n", "quantity", "standard", "lifted", "|difference|");
        std::printf("%-10s %14.6f %14.6f %14.6f\n", "<E>", s.E, l.E, dE);
        std::printf("%-10s %14.6f %14.6f %14.6f\n", "<E^2>", s.E2, l.E2, dE2);
        std::printf("%-10s %14.6f %14.6f\n", "stderr<E>", s.E_err, l.E_err);
        std::printf("%-10s %14.6f %14.6f\n", "stderr<E2>", s.E2_err, l.E2_err);
        std::printf("%-10s %13.1f%% %13.1f%%\n", "accept", 100 * s.accept, 100 * l.accept);
        const bool pass_E = dE < tol, pass_E2 = dE2 < tol;
        std::printf("\n<E>   : |diff|=%.6f  tol=%.6f  -> %s\n", dE, tol, pass_E ? "PASS" : "FAIL");
        std::printf("<E^2> : |diff|=%.6f  tol=%.6f  -> %s\n", dE2, tol, pass_E2 ? "PASS" : "FAIL");
        std::printf("\nRESULT: %s\n", pass_E && pass_E2 ? "PASS" : "FAIL");
        return pass_E && pass_E2 ? 0 : 1;
    } catch (const std::exception& e) {
        std::cerr << "bench_lifted: " << e.what() << std::endl;
        return 3;
    }
}
