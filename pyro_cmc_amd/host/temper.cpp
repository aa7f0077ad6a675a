// temper: parallel tempering of the pyrochlore J1-J2-J3 Heisenberg model with
// every rung held on one GPU.  Mirrors the reference driver
// (temper.cpp:242-621): same flags, temperature ladder, round structure
// (n_sweep sweeps, SSF sample of the T_sample rungs, one energy sample per
// round, checkerboard exchange of neighbouring rungs decided by
// arg = (1/T_s - 1/T_s+1)(E_s - E_s+1)), per-pair swap summary and HDF5
// output (/energy, /ssf, /geometry).
//
// The reference runs one MPI rank per rung and swaps spin arrays between
// ranks; here the rungs are replicas of one device context, a swap exchanges
// temperature labels on the device, and a whole run of rounds is one HIP
// graph launch.  Differences, all deliberate:
//   * --n_replicas sets the rung count (the reference takes the MPI world
//     size; its --n_steps help text describes the same quantity and is the
//     default here);
//   * every rung does n_sweep sweeps per round (the reference gives the
//     unsampled rungs one extra sweep while the sampled ones run their FFT,
//     temper.cpp:424-429, an artefact of the per-rank load balance);
//   * /ssf/static_corr and static_corr_2 use the contracted layout
//     [n_corr, n_T, n_k, 2] of anneal; --munu adds the sublattice-resolved
//     sums [n_corr, n_T, n_k, n_sl, n_sl, 2] of temper.cpp:121-159 as
//     /ssf/static_corr_munu;
//   * one counter-based swap stream replaces the per-rank Xoroshiro streams.
#include <cstdio>
#include <iostream>
#include <numeric>

#include "model_args.hpp"

int main(int argc, char* argv[]) {
    ArgParser prog("temper");
    prog.add({"--output_dir", "-o"}, "Path to output").required = true;
    prog.add({"--seed", "-s"}, "Seed index of the RNG").required = true;
    prog.add({"--save_state"}, "Save the spin state of the coldest replica").flag = true;
    prog.add({"--init_spiral"}, "Pre-initialise spins to a spiral (requires --Q)").flag = true;
    prog.add({"--T_hot"}, "Temperature of the hottest replica (default 10)");
    prog.add({"--T_cold"}, "Temperature of the coldest replica (default min T_sample)");
    prog.add({"--T_ref"}, "Reference temperature of the proposal width").def = {"1.0"};
    auto& ts = prog.add({"--T_sample"}, "Temperatures at which the SSF is accumulated");
    ts.nargs = -1;
    ts.required = true;
    prog.add({"--n_steps"}, "Total number of replicas (including T_sample rungs)").def = {"100"};
    prog.add({"--n_replicas", "-R"}, "Number of rungs (overrides --n_steps)");
    prog.add({"--n_sweep", "-w"}, "Local-update sweeps per replica between exchange attempts").def = {"16"};
    prog.add({"--n_burn_in"}, "Sweeps at fixed T before exchange attempts begin").def = {"16"};
    prog.add({"--n_sample"}, "Number of exchange rounds while collecting statistics").def = {"64"};
    prog.add({"--n_overrelax"}, "Overrelaxation passes per sweep").def = {"1"};
    prog.add({"--prefix"}, "Run name prefix").def = {"run"};
    prog.add({"--lifted"}, "Use lifted Metropolis instead of standard Metropolis").flag = true;
    prog.add({"--munu"}, "Also write the sublattice-resolved correlators").flag = true;
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
        if (!std::filesystem::exists(outdir)) {
            std::fprintf(stderr, "ERROR: output_dir does not exist: %s\n", outdir.string().c_str());
            return 1;
        }
        const size_t seed = size_t(prog.integer("--seed"));
        const double T_hot = prog.is_used("--T_hot") ? prog.dbl("--T_hot") : 10.0;
        const std::vector<double> T_sample = prog.dbls("--T_sample");
        const double T_cold =
            prog.is_used("--T_cold") ? prog.dbl("--T_cold") : *std::min_element(T_sample.begin(), T_sample.end());
        const int n_sweep = int(prog.integer("--n_sweep"));
        const int n_burn_in = int(prog.integer("--n_burn_in"));
        const int n_sample = int(prog.integer("--n_sample"));
        const int n_or = int(prog.integer("--n_overrelax"));
        const int R = int(prog.is_used("--n_replicas") ? prog.integer("--n_replicas") : prog.integer("--n_steps"));
        const int L = int(prog.integer("L"));
        const std::vector<double> B = prog.dbls("-B");
        if (R <= int(T_sample.size())) {
            std::fprintf(stderr, "ERROR: n_replicas (%d) must exceed T_sample count (%zu)\n", R, T_sample.size());
            return 1;
        }
        if (R < 2) {
            std::fprintf(stderr, "ERROR: need at least 2 replicas\n");
            return 1;
        }
        const double J3 = resolve_J3(prog);

        auto [T_grid, S_idx] = generate_T_profile(T_hot, T_cold, T_sample, size_t(R) - T_sample.size());

        CMC::Lattice lat = build_pyro_lat(L);
        CMC::MC_runner mc(lat, seed);
        mc.settings.precision = prog.flag("--fp32") ? PYRO_FP32 : PYRO_FP64;
        mc.settings.strategy = int(prog.integer("--strategy"));
        mc.settings.n_replicas = R;
        build_J1J2J3_h(mc, prog.dbl("--J1"), prog.dbl("--J2"), J3, prog.dbl("--Delta"),
                       vector3::vec3d(B[0], B[1], B[2]), prog.dbl("--T_ref"));
        mc.randomize_spins(seed);
        pyro_ctx* ctx = mc.context();

        if (prog.flag("--init_spiral")) {
            if (!prog.is_used("--Q")) {
                std::fprintf(stderr, "ERROR: --init_spiral requires --Q\n");
                return 1;
            }
            init_spiral_state(lat, round_Qz_to_supercell(prog.dbl("--Q"), L), int(prog.integer("--spiral_axis")));
            const auto& spins = lat.get_objects<CMC::HeisenbergSpin>();
            std::vector<double> S(3 * spins.size());
            for (size_t i = 0; i < spins.size(); ++i)
                for (int a = 0; a < 3; ++a) S[3 * i + a] = spins[i].S[a];
            for (int r = 0; r < R; ++r) CMC::check(pyro_set_spins(ctx, r, S.data()), "set_spins");
            lat.host_newer = false;
        }

        std::stringstream name;
        name << prog.str("--prefix") << DELIM << name_LJ123(prog, J3) << "B=" << B[0] << "," << B[1] << "," << B[2]
             << DELIM << "seed=" << seed << DELIM << "Tc=" << T_cold << DELIM << "sw=" << n_sweep << DELIM;

        // Ladder: slot 0 hottest, replica r starts in slot r.
        CMC::check(pyro_pt_init(ctx, T_grid.data(), seed_key(seed + size_t(R))), "pt_init");

        // SSF accumulators: one per T_sample rung, indexed by ascending T.
        std::vector<size_t> sampled(S_idx.begin(), S_idx.end());
        std::sort(sampled.begin(), sampled.end(), [&](size_t a, size_t b) { return T_grid[a] < T_grid[b]; });
        std::vector<int32_t> ssf_map(size_t(R), -1);
        for (size_t j = 0; j < sampled.size(); ++j) ssf_map[sampled[j]] = int32_t(j);
        std::vector<double> T_slots;
        for (size_t idx : sampled) T_slots.push_back(T_grid[idx]);
        const bool munu = prog.flag("--munu");
        ssf_manager ssfm(lat, {"xx", "yy", "zz"}, T_slots, true, munu);
        CMC::check(pyro_pt_ssf_map(ctx, ssf_map.data()), "pt_ssf_map");

        const int kind = prog.flag("--lifted") ? PYRO_LIFTED : PYRO_METROPOLIS;
        std::printf("[temper] Burning in %d replicas (%d sweeps each)...\n", R, n_burn_in);
        if (n_burn_in > 0) CMC::check(pyro_sweep(ctx, kind, n_or, n_burn_in, nullptr), "sweep");

        std::printf("[temper] Begin sampling (%d rungs, %zu sampled, %d rounds)...\n", R, sampled.size(), n_sample);
        CMC::check(pyro_pt_rounds(ctx, kind, n_or, n_sweep, 1, n_sample), "pt_rounds");
        CMC::check(pyro_synchronize(ctx), "synchronize");
        lat.device_newer = true;

        std::vector<uint64_t> att(static_cast<size_t>(R - 1)), acc(static_cast<size_t>(R - 1)), nE(static_cast<size_t>(R));
        std::vector<double> E(static_cast<size_t>(R)), E2(static_cast<size_t>(R));
        CMC::check(pyro_pt_stats(ctx, att.data(), acc.data(), E.data(), E2.data(), nE.data()), "pt_stats");

        std::printf("\n[temper] per-pair swap acceptance:\n");
        for (int a = 0; a + 1 < R; ++a)
            std::printf("  pair %3d<->%3d  T=%.4e <-> %.4e  accept=%5.1f%%  (%llu/%llu)\n", a, a + 1, T_grid[a],
                        T_grid[a + 1], att[a] ? 100.0 * double(acc[a]) / double(att[a]) : 0.0,
                        (unsigned long long)acc[a], (unsigned long long)att[a]);

        const auto file_path = outdir / (name.str() + ".out.h5");
        hid_t file_id = H5Fcreate(file_path.string().c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
        if (file_id < 0) throw std::runtime_error("Failed to create HDF5 file: " + file_path.string());

        // /energy, ascending T (slot R-1 is the coldest).
        energy_manager e_manager;
        for (int s = R - 1; s >= 0; --s) e_manager.append(T_grid[s], E[s], E2[s], size_t(nE[s]));
        e_manager.write_group(file_id, "/energy");

        // /ssf: contracted sums per sampled rung (ascending T), optional munu.
        ssfm.write_group(file_id, "/ssf");
        if (munu) ssfm.write_munu(file_id, "/ssf");
        CMC::write_geometry_group(file_id, lat);
        H5Fclose(file_id);
        std::printf("[temper] Saved to %s\n", file_path.string().c_str());

        if (prog.flag("--save_state")) {
            std::vector<int32_t> rep_of(static_cast<size_t>(R));
            CMC::check(pyro_pt_permutation(ctx, rep_of.data()), "pt_permutation");
            lat.view_replica(rep_of[size_t(R - 1)]);
            const auto f = outdir / (name.str() + ".spins.h5");
            std::printf("[temper] Saving spin state of coldest replica (T=%.3e) to %s\n", T_grid[size_t(R - 1)],
                        f.string().c_str());
            CMC::save_spin_state(lat, f);
        }
    } catch (const std::exception& e) {
        std::cerr << "temper: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
