// bench_speed: throughput of the two lifted-Metropolis sweeps (displacement
// and rotation), each sweep one lifted pass plus one overrelaxation pass, on
// the J1-J2 Heisenberg model of an L x L square lattice.  The question, the
// flags, their defaults and the report are those of the reference harness
// (test/bench_speed.cpp); the measurement is the device's.
//
// Both variants run in one device context with --replicas copies of the
// lattice (1 by default: the reference's single lattice).  Every timed
// stretch starts from the same state: the context's spins and lifted
// directions are reset to the initial uniform random configuration and
// n_warm untimed sweeps of that variant are run, then n_sweep sweeps are
// bracketed by HIP events on the context stream (pyro_mark), so the figure
// is device time per sweep with no host clock or launch-queue gap in it.
// Each repetition times both variants; the best repetition of each is
// reported.  Acceptance is counted on a separate stretch after the timing,
// so counting does not enter the timed sweeps.
#include <cstdint>
#include <cstdio>
#include <iostream>
#include <string>
#include <vector>

#include "cli.hpp"
#include "pyro_host.hpp"

namespace {

void ok(int status, const char* what) {
    if (status != PYRO_OK) throw std::runtime_error(std::string(what) + ": " + pyro_last_error());
}

struct Variant {
    const char* name;
    int kind;
    double best_ms = 1e300;  // best device time of n_sweep sweeps
    double accept = 0;
};

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
    prog.add({"--reps"}, "Repetitions; best time reported").def = {"3"};
    prog.add({"--seed"}, "Seed").def = {"2718281828"};
    prog.add({"--replicas"}, "Copies of the lattice swept together (throughput at scale)").def = {"1"};
    prog.add({"--fp32"}, "FP32 spins and fields (FP64 accumulators)").flag = true;
    try {
        prog.parse(argc, argv);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 2;
    }
    pyro_ctx* ctx = nullptr;
    try {
        const double J1 = prog.dbl("--J1"), J2 = prog.dbl("--J2"), T = prog.dbl("--T");
        const double T_ref = prog.dbl("--T_ref");
        const int L = int(prog.integer("--L")), R = int(prog.integer("--replicas"));
        const size_t n_warm = size_t(prog.integer("--n_warm")), n_sweep = size_t(prog.integer("--n_sweep"));
        const size_t reps = size_t(prog.integer("--reps"));
        const uint64_t seed = uint64_t(prog.integer("--seed"));
        if (L < 1 || n_sweep < 1 || reps < 1 || R < 1)
            throw std::invalid_argument("L, n_sweep, reps and replicas must be positive");
        const size_t n_spins = size_t(L) * size_t(L);

        std::printf("=== lifted Metropolis speed benchmark ===\n");
        std::printf("Model : %dx%d square-lattice J1-J2 Heisenberg (%zu spins)\n", L, L, n_spins);
        std::printf("        J1=%g  J2=%g  T=%g  T_ref=%g\n", J1, J2, T, T_ref);
        std::printf("Stats : n_warm=%zu  n_sweep=%zu  reps=%zu  seed=%llu  replicas=%d\n", n_warm, n_sweep, reps,
                    (unsigned long long)seed, R);
        std::printf("Timing: HIP events on the device stream, from the same starting state each time\n\n");

        // the square lattice: one site per 1 x 1 x 1 cell, an L x L x 1 supercell
        const int32_t pos[3] = {0, 0, 0}, kind[1] = {0}, cell[3] = {1, 1, 1}, Ls[3] = {L, L, 1};
        const uint64_t key = seed_key(seed);
        ok(pyro_create(1, pos, kind, cell, Ls, R, prog.flag("--fp32") ? PYRO_FP32 : PYRO_FP64, key, &ctx), "create");
        const int32_t counts[1] = {4};
        const int32_t axial[12] = {1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1, 0};
        const int32_t diagonal[12] = {1, 1, 0, 1, -1, 0, -1, 1, 0, -1, -1, 0};
        const double I1[9] = {J1, 0, 0, 0, J1, 0, 0, 0, J1}, I2[9] = {J2, 0, 0, 0, J2, 0, 0, 0, J2};
        ok(pyro_define_coupling(ctx, "J1", 1, counts, axial, I1), "define J1");
        ok(pyro_define_coupling(ctx, "J2", 1, counts, diagonal, I2), "define J2");
        ok(pyro_setup_lattice(ctx), "setup");
        ok(pyro_set_T_ref(ctx, T_ref), "T_ref");
        const std::vector<double> Ts(size_t(R), T);
        ok(pyro_set_temperatures(ctx, Ts.data()), "temperatures");

        // the common starting state, kept on the host
        ok(pyro_randomize_spins(ctx, key), "randomize");
        std::vector<std::vector<double>> S0(size_t(R), std::vector<double>(3 * n_spins));
        for (int r = 0; r < R; ++r) ok(pyro_get_spins(ctx, r, S0[size_t(r)].data()), "get_spins");
        const std::vector<int8_t> up(n_spins, int8_t(1));
        auto reset = [&] {
            for (int r = 0; r < R; ++r) {
                ok(pyro_set_spins(ctx, r, S0[size_t(r)].data()), "set_spins");
                ok(pyro_set_lifted_dir(ctx, r, up.data()), "set_lifted_dir");
            }
        };

        Variant var[2] = {{"displacement", PYRO_LIFTED}, {"rotation", PYRO_LIFTED_ROT}};
        for (size_t rep = 0; rep < reps; ++rep) {
            double ms[2];
            for (int v = 0; v < 2; ++v) {
                reset();
                if (n_warm) ok(pyro_sweep(ctx, var[v].kind, 1, int(n_warm), nullptr), "warm-up");
                ok(pyro_mark(ctx, 0), "mark");
                ok(pyro_sweep(ctx, var[v].kind, 1, int(n_sweep), nullptr), "timed sweeps");
                ok(pyro_mark(ctx, 1), "mark");
                ok(pyro_elapsed_ms(ctx, &ms[v]), "elapsed");
                var[v].best_ms = std::min(var[v].best_ms, ms[v]);
            }
            std::printf("rep %zu:  displacement %8.1f ns/sweep   rotation %8.1f ns/sweep\n", rep,
                        1e6 * ms[0] / double(n_sweep), 1e6 * ms[1] / double(n_sweep));
        }
        for (auto& v : var) {
            std::vector<uint64_t> acc(static_cast<size_t>(R));
            const int n_acc = int(std::min<size_t>(n_sweep, 2000));
            ok(pyro_sweep(ctx, v.kind, 1, n_acc, acc.data()), "acceptance sweeps");
            uint64_t a = 0;
            for (uint64_t x : acc) a += x;
            v.accept = double(a) / (double(R) * double(n_acc) * double(n_spins));
        }

        const double ns_d = 1e6 * var[0].best_ms / double(n_sweep), ns_r = 1e6 * var[1].best_ms / double(n_sweep);
        std::printf("\n%-14s %14s %14s\n", "quantity", "displacement", "rotation");
        std::printf("%-14s %14.1f %14.1f\n", "ns / sweep", ns_d, ns_r);
        std::printf("%-14s %14.3f %14.3f\n", "ns / spin", ns_d / double(n_spins * size_t(R)),
                    ns_r / double(n_spins * size_t(R)));
        std::printf("%-14s %14.3g %14.3g\n", "sweeps / sec", 1e9 / ns_d, 1e9 / ns_r);
        std::printf("%-14s %13.1f%% %13.1f%%\n", "accept", 100 * var[0].accept, 100 * var[1].accept);
        const double ratio = ns_r / ns_d;
        std::printf("\nFaster: %s  (by %.2fx, best-of-%zu)\n", ratio > 1.0 ? "displacement" : "rotation",
                    ratio > 1.0 ? ratio : 1.0 / ratio, reps);
    } catch (const std::exception& e) {
        std::cerr << "bench_speed: " << e.what() << std::endl;
        if (ctx) pyro_destroy(ctx);
        return 1;
    }
    pyro_destroy(ctx);
    return 0;
}
