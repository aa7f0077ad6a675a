// anneal: simulated annealing of the pyrochlore J1-J2-J3 Heisenberg model
// with the static structure factor sampled at the T_sample temperatures.
// Mirrors the reference driver (anneal.cpp:33-248): same flags, run name,
// temperature schedule, streaming HDF5 layout (/ssf, /energy, /geometry) and
// state dumps, with the sweeps on the GPU.
//
// Differences, all deliberate:
//   * spins start uniform on the sphere from the seed (the reference leaves
//     them at their default value, MC.hpp:55);
//   * --seed indexes a SplitMix64-derived 64-bit key, not the reference's
//     table of stored random numbers (random_seeds.hpp);
//   * --save_state writes .spins.h5 and --save_state_ft writes .spindft.h5
//     (the reference swaps the two writers, anneal.cpp:235-245);
//   * --fp32 selects FP32 spins with FP64 accumulators; --strategy forces a
//     sweep launch strategy.
#include <cstdio>
#include <iostream>

#include "model_args.hpp"

int main(int argc, char* argv[]) {
    ArgParser prog("anneal");
    prog.add({"--output_dir", "-o"}, "Path to output").required = true;
    prog.add({"--seed", "-s"}, "Seed index of the RNG").required = true;
    prog.add({"--save_state"}, "Write the final spins (.spins.h5)").flag = true;
    prog.add({"--save_state_ft"}, "Write the final spin DFT (.spindft.h5)").flag = true;
    prog.add({"--init_spiral"}, "Start from the spiral given by --Q").flag = true;
    prog.add({"--T_hot"}, "Temperature to anneal from (default 10)");
    prog.add({"--T_cold"}, "Lowest temperature (default min T_sample)");
    prog.add({"--T_ref"}, "Reference temperature of the proposal width").def = {"1.0"};
    auto& ts = prog.add({"--T_sample"}, "Temperatures at which the SSF is sampled");
    ts.nargs = -1;
    ts.required = true;
    prog.add({"--n_steps"}, "Number of annealing steps").def = {"100"};
    prog.add({"--n_sweep", "-w"}, "Sweeps per temperature step").def = {"16"};
    prog.add({"--n_burn_in"}, "Sweeps at T_hot before annealing").def = {"16"};
    prog.add({"--n_sample"}, "SSF samples per T_sample").def = {"64"};
    prog.add({"--prefix"}, "Run name prefix").def = {"run"};
    prog.add({"--lifted"}, "Use lifted Metropolis instead of standard Metropolis").flag = true;
    prog.add({"--fp32"}, "FP32 spins and fields (FP64 accumulators)").flag = true;
    prog.add({"--strategy"}, "Sweep launch strategy 0..5 (0 = automatic)").def = {"0"};
    declare_physical(prog);
    try {
        prog.parse(argc, argv);
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }

    try {
        const std::filesystem::path outdir(prog.str("--output_dir"));
        if (!std::filesystem::exists(outdir)) throw std::runtime_error("Cannot open outdir");
        const size_t seed = size_t(prog.integer("--seed"));
        const double T_hot = prog.is_used("--T_hot") ? prog.dbl("--T_hot") : 10.0;
        const std::vector<double> T_sample = prog.dbls("--T_sample");
        const double T_cold =
            prog.is_used("--T_cold") ? prog.dbl("--T_cold") : *std::min_element(T_sample.begin(), T_sample.end());
        const size_t n_steps = size_t(prog.integer("--n_steps"));
        const size_t n_sweep = size_t(prog.integer("--n_sweep"));
        const size_t n_burn_in = size_t(prog.integer("--n_burn_in"));
        const size_t n_sample = size_t(prog.integer("--n_sample"));
        const int L = int(prog.integer("L"));
        const std::vector<double> B = prog.dbls("-B");
        const double J3 = resolve_J3(prog);

        CMC::Lattice lat = build_pyro_lat(L);
        CMC::MC_runner mc(lat, seed);
        mc.settings.precision = prog.flag("--fp32") ? PYRO_FP32 : PYRO_FP64;
        mc.settings.strategy = int(prog.integer("--strategy"));
        build_J1J2J3_h(mc, prog.dbl("--J1"), prog.dbl("--J2"), J3, prog.dbl("--Delta"),
                       vector3::vec3d(B[0], B[1], B[2]), prog.dbl("--T_ref"));
        mc.randomize_spins(seed);

        if (prog.flag("--init_spiral")) {
            if (!prog.is_used("--Q")) throw std::runtime_error("--init_spiral requires --Q");
            const double Qr = round_Qz_to_supercell(prog.dbl("--Q"), L);
            const int axis = int(prog.integer("--spiral_axis"));
            std::printf("Pre-initialising to spiral order (Q=%.10g, axis=%d)...\n", Qr, axis);
            init_spiral_state(lat, Qr, axis);
        }

        std::stringstream name;
        name << prog.str("--prefix") << DELIM << name_LJ123(prog, J3) << "B=" << B[0] << "," << B[1] << "," << B[2]
             << DELIM << "seed=" << seed << DELIM << "Tc=" << T_cold << DELIM << "sw=" << n_sweep << DELIM;

        auto [T_grid, S_idx] = generate_T_profile(T_hot, T_cold, T_sample, n_steps);
        energy_manager e_manager;
        const auto file_path = outdir / (name.str() + ".out.h5");
        hid_t file_id = H5Fcreate(file_path.string().c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
        if (file_id < 0) throw std::runtime_error("Failed to create HDF5 file: " + file_path.string());
        ssf_manager ssfm(lat, {"xx", "yy", "zz", "xy", "xz", "yz"}, file_id, "/ssf", T_sample, true);

        const bool lifted = prog.flag("--lifted");
        const int kind = lifted ? PYRO_LIFTED : PYRO_METROPOLIS;
        const size_t N = size_t(16) * L * L * L;
        std::printf("Done. Begin anneal%s...\n", lifted ? " (lifted Metropolis)" : "");
        std::printf("Burning in (%zu sweeps)...\n", n_burn_in);
        mc.sweeps(kind, T_hot, 1, n_burn_in);

        for (size_t i = 0; i < T_grid.size(); ++i) {
            const double T = T_grid[i];
            e_manager.new_T(T);
            const size_t accepted = mc.sweeps(kind, T, 1, n_sweep);
            e_manager.sample(mc.total_energy_per_unit_cell());
            if (S_idx.count(i)) {
                std::printf("Sampling at T=%lf (%zu sweeps)...\n", T, n_sample);
                ssfm.new_T(T);
                for (size_t k = 0; k < n_sample; ++k) {
                    const size_t acc = mc.sweeps(kind, T, 1, n_sweep);
                    std::printf("[ssf] Iter %4zu T=%.3e Acceptance rate: %.2f%%\n", k, T,
                                acc * 100.0 / double(N) / double(n_sweep));
                    ssfm.sample();
                }
                ssfm.flush();
            }
            std::printf("[anneal] Iter %4zu T=%.3e E=%3e Acceptance rate: %.2f%%\n", i, T, e_manager.curr_E(),
                        accepted * 100.0 / double(N) / double(n_sweep));
        }

        ssfm.write_group(file_id, "/ssf");
        e_manager.write_group(file_id, "/energy");
        CMC::write_geometry_group(file_id, lat);
        H5Fclose(file_id);
        std::cout << "Saved to \n" << file_path << std::endl;

        if (prog.flag("--save_state_ft")) {
            const auto f = outdir / (name.str() + ".spindft.h5");
            std::printf("Saving Fourier transformed spin state to %s\n", f.string().c_str());
            CMC::save_ft_spin_state(lat, f);
        }
        if (prog.flag("--save_state")) {
            const auto f = outdir / (name.str() + ".spins.h5");
            std::printf("Saving spin state to %s\n", f.string().c_str());
            CMC::save_spin_state(lat, f);
        }
    } catch (const std::exception& e) {
        std::cerr << "anneal: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
