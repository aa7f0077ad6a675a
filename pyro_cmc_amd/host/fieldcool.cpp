// fieldcool: anneal in a field, release it, re-burn and sample the static
// structure factor at T_cold.  Mirrors the reference driver
// (fieldcool.cpp:31-214): same flags and run name; burn-in at T_hot; n_steps
// log steps T_hot * f^i (f = (T_cold/T_hot)^(1/n_steps)) of n_sweep
// Metropolis + OR sweeps with one energy sample each; then the field set
// to zero, n_burn_in sweeps at T_cold, and n_sample x (n_sweep sweeps + one
// SSF sample of xx, yy, zz); output /ssf, /energy, /geometry and optionally
// the spins.  Spins start uniform on the sphere from the seed (the
// reference leaves them at their default value).
#include <cmath>
#include <cstdio>
#include <iostream>

#include "model_args.hpp"

int main(int argc, char* argv[]) {
    ArgParser prog("fieldcool");
    prog.add({"--output_dir", "-o"}, "Path to output").required = true;
    prog.add({"--seed", "-s"}, "Seed index of the RNG").required = true;
    prog.add({"--save_state"}, "Write the final spins (.spins.h5)").flag = true;
    prog.add({"--init_spiral"}, "Pre-initialise spins to the spiral given by --Q").flag = true;
    prog.add({"--T_hot"}, "Temperature to begin annealing from (default 10)");
    prog.add({"--T_ref"}, "Reference temperature of the proposal width").def = {"1.0"};
    prog.add({"--T_cold"}, "Final temperature").required = true;
    prog.add({"--n_steps"}, "Number of annealing steps").def = {"100"};
    prog.add({"--n_sweep"}, "Sweeps per temperature step").def = {"16"};
    prog.add({"--n_burn_in"}, "Sweeps at T_hot before annealing, and after releasing the field").def = {"16"};
    prog.add({"--n_sample"}, "SSF samples at T_cold").def = {"64"};
    prog.add({"--prefix"}, "Run name prefix").def = {"run"};
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
        const double T_cold = prog.dbl("--T_cold");
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
             << DELIM << "seed=" << seed << DELIM << "Tc=" << T_cold << DELIM;

        const size_t N = size_t(16) * L * L * L;
        std::printf("Burning in (%zu sweeps)...\n", n_burn_in);
        mc.sweeps(PYRO_METROPOLIS, T_hot, 1, n_burn_in);
        std::printf("Done. Begin anneal...\n");
        const double factor = std::pow(T_cold / T_hot, 1. / double(n_steps));
        double T = T_hot;
        energy_manager e_manager;
        for (size_t i = 0; i < n_steps; ++i) {
            e_manager.new_T(T);
            const size_t accepted = mc.sweeps(PYRO_METROPOLIS, T, 1, n_sweep);
            e_manager.sample(mc.total_energy_per_unit_cell());
            std::printf("Iter %4zu T=%.3e E=%3e Acceptance rate: %.2f%%\n", i, T, e_manager.curr_E(),
                        accepted * 100.0 / double(N) / double(n_sweep));
            T *= factor;
        }

        ssf_manager ssfm(lat, {"xx", "yy", "zz"}, 1);
        ssfm.new_T(T);
        std::printf("Releasing field and re-burning (%zu sweeps)...\n", n_burn_in);
        mc.set_global_field(vector3::vec3d(0, 0, 0));
        mc.sweeps(PYRO_METROPOLIS, T, 1, n_burn_in);
        std::printf("Sampling at T=%lf (%zu sweeps)...\n", T, n_sample);
        for (size_t i = 0; i < n_sample; ++i) {
            const size_t accepted = mc.sweeps(PYRO_METROPOLIS, T, 1, n_sweep);
            std::printf("[cold] Iter %4zu T=%.3e Acceptance rate: %.2f%%\n", i, T,
                        accepted * 100.0 / double(N) / double(n_sweep));
            ssfm.sample();
        }

        const auto file_path = outdir / (name.str() + ".out.h5");
        hid_t file_id = H5Fcreate(file_path.string().c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
        if (file_id < 0) throw std::runtime_error("Failed to create HDF5 file: " + file_path.string());
        ssfm.write_group(file_id, "/ssf");
        e_manager.write_group(file_id, "/energy");
        CMC::write_geometry_group(file_id, lat);
        H5Fclose(file_id);
        std::cout << "Saved to \n" << file_path << std::endl;
        if (prog.flag("--save_state")) {
            const auto f = outdir / (name.str() + ".spins.h5");
            std::printf("Saving spin state to %s\n", f.string().c_str());
            CMC::save_spin_state(lat, f);
        }
    } catch (const std::exception& e) {
        std::cerr << "fieldcool: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
