// spiral_byhand: known-answer generator.  Sets the spiral
// S_i = (0, cos q.r_i, sin q.r_i) for the integer wave-vector index Q, takes
// one energy and one SSF sample and writes <L,Q>.out.h5 (/ssf, /energy,
// /geometry) and .spins.h5, as the reference tool does
// (test/spiral_byhand.cpp:27-137): J1 on the unsplit nearest-neighbour
// shell (default -1), J2, J3a, J3b isotropic.
#include <cstdio>
#include <iostream>

#include "cli.hpp"
#include "pyro_host.hpp"

int main(int argc, char* argv[]) {
    ArgParser prog("spiral_byhand");
    prog.add({"--output_dir", "-o"}, "Path to output").required = true;
    prog.add({"L"}, "Linear dimension of the cubic supercell").required = true;
    auto& q = prog.add({"Q"}, "Integer wave-vector index (3 values)");
    q.nargs = 3;
    q.required = true;
    prog.add({"--J1"}, "Nearest-neighbour Heisenberg coupling strength").def = {"-1"};
    prog.add({"--J2"}, "Second-nearest-neighbour Heisenberg coupling strength").def = {"0"};
    prog.add({"--J3"}, "Third-nearest-neighbour Heisenberg coupling strength").def = {"0"};
    prog.add({"--fp32"}, "FP32 spins and fields (FP64 accumulators)").flag = true;
    try {
        prog.parse(argc, argv);
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    try {
        const std::filesystem::path outdir(prog.str("--output_dir"));
        if (!std::filesystem::exists(outdir)) throw std::runtime_error("Cannot open outdir");
        const int L = int(prog.integer("L"));
        const auto Qv = prog.raw("Q");
        const ipos_t Q{std::stoll(Qv[0]), std::stoll(Qv[1]), std::stoll(Qv[2])};

        CMC::Lattice lat = build_pyro_lat(L);
        CMC::MC_runner mc(lat, 0);
        mc.settings.precision = prog.flag("--fp32") ? PYRO_FP32 : PYRO_FP64;
        const vector3::mat33d I = vector3::mat33d::from_cols({1, 0, 0}, {0, 1, 0}, {0, 0, 1});
        mc.define_coupling("J1", pyrochlore::shell(1), prog.dbl("--J1") * I);
        mc.define_coupling("J2", pyrochlore::shell(2), prog.dbl("--J2") * I);
        mc.define_coupling("J3a", pyrochlore::shell(3), prog.dbl("--J3") * I);
        mc.define_coupling("J3b", pyrochlore::shell(4), prog.dbl("--J3") * I);
        mc.setup_lattice();

        std::stringstream name;
        name << "L=" << L << DELIM << "Q=" << Q[0] << "," << Q[1] << "," << Q[2] << DELIM;

        ssf_manager ssfm(lat, {"xx", "yy", "zz"}, 1);
        energy_manager e_manager;
        const vector3::vec3d qv = lat.lattice.wavevector_from_idx3(Q);
        std::printf("q = (%.17g, %.17g, %.17g)\n", qv[0], qv[1], qv[2]);
        for (auto& s : lat.get_objects<CMC::HeisenbergSpin>()) {
            const vector3::vec3d r(double(s.ipos[0]), double(s.ipos[1]), double(s.ipos[2]));
            const double phase = vector3::dot(qv, r);
            s.S = vector3::vec3d(0, std::cos(phase), std::sin(phase));
        }
        ssfm.new_T(0);
        e_manager.new_T(0);
        const double total_E = mc.total_energy_per_unit_cell();
        ssfm.sample();
        e_manager.sample(total_E);
        std::printf("Total E = %.17g\n", total_E);

        const auto file_path = outdir / (name.str() + ".out.h5");
        hid_t file_id = H5Fcreate(file_path.string().c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
        if (file_id < 0) throw std::runtime_error("Failed to create HDF5 file: " + file_path.string());
        ssfm.write_group(file_id, "/ssf");
        e_manager.write_group(file_id, "/energy");
        CMC::write_geometry_group(file_id, lat);
        H5Fclose(file_id);
        std::cout << "Saved to \n" << file_path << std::endl;
        const auto f = outdir / (name.str() + ".spins.h5");
        std::printf("Saving spin state to %s\n", f.string().c_str());
        CMC::save_spin_state(lat, f);
    } catch (const std::exception& e) {
        std::cerr << "spiral_byhand: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
