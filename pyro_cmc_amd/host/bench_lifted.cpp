// bench_lifted: do lifted and standard Metropolis sample the same Boltzmann
// distribution?  The question and the interface are the reference harness's
// (test/bench_lifted.cpp): the J1-J2 Heisenberg model on an L x L square
// lattice, the same flags and defaults, the same report and the same verdict
// (PASS when both |<E>| and |<E^2>| per spin of the two samplers differ by
// less than tol_frac * min(|J1|, |J2|); exit status 0, else 1).
//
// The sampling runs as a GPU ensemble.  Each sampler is one device context
// holding --chains independent chains of the lattice (independent counter-
// RNG streams and uniform random starting spins); after n_burn sweeps every
// chain takes n_sample measurement sweeps of one sampler pass plus one
// overrelaxation pass, through a constant-temperature device schedule that
// records every chain's energy per spin after every sweep on the device (one
// replayed graph, no host round trip per sweep).  <E> and <E^2> are ensemble
// means over chains and sweeps, and their standard errors come from the
// spread of the chains' own means, which needs no autocorrelation model.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <iostream>
#include <numeric>
#include <vector>

#include "cli.hpp"
#include "pyro_host.hpp"

namespace {

struct Sampler {
    const char* name;
    int kind;
    uint64_t tag;  // distinguishes the two samplers' streams under one --seed
};

struct Result {
    double E = 0, E2 = 0, E_err = 0, E2_err = 0, accept = 0;
};

void ok(int status, const char* what) {
    if (status != PYRO_OK) throw std::runtime_error(std::string(what) + ": " + pyro_last_error());
}

// Mean over chains of per-chain means, and its standard error from the
// chain-to-chain spread.
void ensemble(const std::vector<double>& chain_means, double& mean, double& err) {
    const double n = double(chain_means.size());
    mean = std::accumulate(chain_means.begin(), chain_means.end(), 0.0) / n;
    double ss = 0;
    for (double m : chain_means) ss += (m - mean) * (m - mean);
    err = std::sqrt(ss / (n - 1.0) / n);
}

Result run(const Sampler& smp, int L, int chains, double J1, double J2, double T, double T_ref, size_t n_burn,
           size_t n_sample, uint64_t seed, int precision) {
    // one site per cell of a 1 x 1 x 1 cubic cell; an L x L x 1 supercell is
    // the square lattice, J1 on the four axial and J2 on the four diagonal
    // neighbours in the plane
    const int32_t pos[3] = {0, 0, 0}, kind[1] = {0}, cell[3] = {1, 1, 1}, Ls[3] = {L, L, 1};
    pyro_ctx* ctx = nullptr;
    const uint64_t key = seed_key(seed) ^ smp.tag;
    ok(pyro_create(1, pos, kind, cell, Ls, chains, precision, key, &ctx), "create");
    struct Closer {
        pyro_ctx* c;
        ~Closer() { pyro_destroy(c); }
    } closer{ctx};
    const int32_t counts[1] = {4};
    const int32_t axial[12] = {1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1, 0};
    const int32_t diagonal[12] = {1, 1, 0, 1, -1, 0, -1, 1, 0, -1, -1, 0};
    const double I1[9] = {J1, 0, 0, 0, J1, 0, 0, 0, J1}, I2[9] = {J2, 0, 0, 0, J2, 0, 0, 0, J2};
    ok(pyro_define_coupling(ctx, "J1", 1, counts, axial, I1), "define J1");
    ok(pyro_define_coupling(ctx, "J2", 1, counts, diagonal, I2), "define J2");
    ok(pyro_setup_lattice(ctx), "setup");
    ok(pyro_set_T_ref(ctx, T_ref), "T_ref");
    ok(pyro_randomize_spins(ctx, key), "randomize");
    const std::vector<double> Ts(size_t(chains), T);
    ok(pyro_set_temperatures(ctx, Ts.data()), "temperatures");
    std::printf("Running %s Metropolis (%d chains)...\n", smp.name, chains);
    if (n_burn) ok(pyro_sweep(ctx, smp.kind, 1, int(n_burn), nullptr), "burn-in");

    // measurement: n_sample steps of a constant schedule, energies on the device
    const std::vector<double> sched(n_sample, T);
    ok(pyro_anneal_schedule(ctx, sched.data(), int(n_sample)), "schedule");
    ok(pyro_anneal_steps(ctx, smp.kind, 1, 1, 0, int(n_sample)), "measurement sweeps");
    std::vector<double> hist(size_t(chains) * n_sample);  // [chain][sweep], per cell = per spin
    ok(pyro_anneal_history(ctx, hist.data()), "history");
    // acceptance of the sampler pass over a further stretch of sweeps
    std::vector<uint64_t> acc(static_cast<size_t>(chains));
    const int n_acc = int(std::min<size_t>(n_sample, 1000));
    ok(pyro_sweep(ctx, smp.kind, 1, n_acc, acc.data()), "acceptance sweeps");

    std::vector<double> mE(static_cast<size_t>(chains)), mE2(static_cast<size_t>(chains));
    for (int c = 0; c < chains; ++c) {
        double s = 0, s2 = 0;
        for (size_t k = 0; k < n_sample; ++k) {
            const double e = hist[size_t(c) * n_sample + k];
            s += e;
            s2 += e * e;
        }
        mE[size_t(c)] = s / double(n_sample);
        mE2[size_t(c)] = s2 / double(n_sample);
    }
    Result res;
    ensemble(mE, res.E, res.E_err);
    ensemble(mE2, res.E2, res.E2_err);
    const double n_acc_total = double(std::accumulate(acc.begin(), acc.end(), uint64_t(0)));
    res.accept = n_acc_total / (double(chains) * double(n_acc) * double(L) * double(L));
    return res;
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
    prog.add({"--chains"}, "Independent chains per sampler, run side by side on the GPU").def = {"32"};
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
        const int L = int(prog.integer("--L")), chains = int(prog.integer("--chains"));
        const size_t n_burn = size_t(prog.integer("--n_burn")), n_sample = size_t(prog.integer("--n_sample"));
        const uint64_t seed = uint64_t(prog.integer("--seed"));
        if (L < 1 || n_sample < 1 || chains < 2)
            throw std::invalid_argument("L and n_sample must be positive and --chains at least 2");
        const double jmin = std::min(std::abs(J1), std::abs(J2)), tol = tol_frac * jmin;
        const int precision = prog.flag("--fp32") ? PYRO_FP32 : PYRO_FP64;

        std::printf("=== lifted vs. standard Metropolis consistency benchmark ===\n");
        std::printf("Model : %dx%d square-lattice J1-J2 Heisenberg\n", L, L);
        std::printf("        J1=%g  J2=%g  T=%g  T_ref=%g\n", J1, J2, T, T_ref);
        std::printf("Stats : n_burn=%zu  n_sample=%zu  seed=%llu  chains=%d\n", n_burn, n_sample,
                    (unsigned long long)seed, chains);
        std::printf("Tol   : %g  (= %g * min(|J1|,|J2|)=%g)\n\n", tol, tol_frac, jmin);

        const Sampler standard{"standard", PYRO_METROPOLIS, 0x5354444D45545230ull};
        const Sampler lifted{"lifted", PYRO_LIFTED, 0x4C49465445444D43ull};
        const Result s = run(standard, L, chains, J1, J2, T, T_ref, n_burn, n_sample, seed, precision);
        const Result l = run(lifted, L, chains, J1, J2, T, T_ref, n_burn, n_sample, seed, precision);

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
