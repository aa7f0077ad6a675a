// ltgs: Luttinger-Tisza ground-state bound of the pyrochlore J1-J2-J3 XXZ
// model.  Mirrors the reference driver (luttinger_tisza.cpp:84-307): the
// same flags, printed summary and output file <name_LJ123>ltgs.out.h5 with
// eigenvalue_map [Nk], k_star_idx int64[3], k_star_vec [3], the scalar
// E_min (= lambda_min / 2, the energy per spin), eigenvector [48][2],
// /geometry {index_cell, recip_vectors} and the root attribute k_dims.  The
// spectrum comes from pyro_lt_spectrum (one workgroup per k builds M(k),
// tridiagonalises it and finds its lowest eigenvalue on the device) instead
// of a per-k Eigen loop.
#include <cmath>
#include <cstdio>
#include <iostream>

#include "model_args.hpp"

int main(int argc, char* argv[]) {
    ArgParser prog("ltgs");
    prog.add({"--output_dir", "-o"}, "Path to output directory");
    declare_physical(prog);
    try {
        prog.parse(argc, argv);
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    try {
        std::filesystem::path outdir;
        if (prog.is_used("--output_dir")) {
            outdir = prog.str("--output_dir");
            if (!std::filesystem::exists(outdir))
                throw std::runtime_error("output_dir does not exist: " + outdir.string());
        }
        const double J1 = prog.dbl("--J1"), J2 = prog.dbl("--J2"), J3 = resolve_J3(prog);
        const double Delta = prog.dbl("--Delta");
        const int L = int(prog.integer("L"));
        const int32_t Ls[3] = {L, L, L};
        const size_t Nk = size_t(L) * L * L;
        const int M_dim = 48;

        std::vector<double> eigenvalue_map(Nk), ev(2 * M_dim);
        int32_t ks[3];
        CMC::check(pyro_lt_spectrum(Ls, J1, J2, J3, Delta, eigenvalue_map.data(), ks, ev.data()), "pyro_lt_spectrum");
        const size_t k_flat = (size_t(ks[0]) * L + ks[1]) * L + ks[2];
        const double E_min = eigenvalue_map[k_flat];
        const double E_per_spin = 0.5 * E_min;
        // q of the centred index, as the spectrum uses
        double k_star_vec[3];
        for (int a = 0; a < 3; ++a) {
            const int K = 2 * ks[a] < L ? ks[a] : ks[a] - L;
            k_star_vec[a] = 2 * M_PI * double(K) / double(8 * L);
        }

        std::printf("LT minimum: λ_min = %.6f  E/spin = %.6f\n", E_min, E_per_spin);
        std::printf("Eigvec sublattice |s_α|² :");
        for (int s = 0; s < 16; s++) {
            double w = 0;
            for (int c = 0; c < 3; ++c) w += ev[2 * (3 * s + c)] * ev[2 * (3 * s + c)] + ev[2 * (3 * s + c) + 1] * ev[2 * (3 * s + c) + 1];
            std::printf(" %.4f", w);
        }
        std::printf("\n");
        std::printf("k*  idx = [%lld, %lld, %lld]\n", (long long)ks[0], (long long)ks[1], (long long)ks[2]);
        std::printf("k*  vec = [%.4f, %.4f, %.4f]  (rad / coord-unit)\n", k_star_vec[0], k_star_vec[1], k_star_vec[2]);

        if (!prog.is_used("--output_dir")) return 0;

        const auto file_path = outdir / (name_LJ123(prog, J3) + "ltgs.out.h5");
        hid_t fid = H5Fcreate(file_path.string().c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
        if (fid < 0) throw std::runtime_error("Failed to create " + file_path.string());
        auto write = [&](hid_t where, const char* dname, hid_t type, int rank, const hsize_t* dims, const void* data) {
            hid_t sp = rank == 0 ? H5Screate(H5S_SCALAR) : H5Screate_simple(rank, dims, nullptr);
            hid_t ds = H5Dcreate2(where, dname, type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
            if (ds < 0 || H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
                throw std::runtime_error(std::string("cannot write ") + dname);
            H5Dclose(ds);
            H5Sclose(sp);
        };
        const hsize_t nk = Nk, three = 3, evd[2] = {hsize_t(M_dim), 2};
        write(fid, "eigenvalue_map", H5T_NATIVE_DOUBLE, 1, &nk, eigenvalue_map.data());
        const int64_t ksi[3] = {ks[0], ks[1], ks[2]};
        write(fid, "k_star_idx", H5T_NATIVE_INT64, 1, &three, ksi);
        write(fid, "k_star_vec", H5T_NATIVE_DOUBLE, 1, &three, k_star_vec);
        write(fid, "E_min", H5T_NATIVE_DOUBLE, 0, nullptr, &E_per_spin);
        write(fid, "eigenvector", H5T_NATIVE_DOUBLE, 2, evd, ev.data());
        CMC::Lattice lat = build_pyro_lat(L);
        CMC::write_geometry_group(fid, lat);
        {
            const int kd[3] = {L, L, L};
            hid_t sp = H5Screate_simple(1, &three, nullptr);
            hid_t at = H5Acreate2(fid, "k_dims", H5T_NATIVE_INT, sp, H5P_DEFAULT, H5P_DEFAULT);
            H5Awrite(at, H5T_NATIVE_INT, kd);
            H5Aclose(at);
            H5Sclose(sp);
        }
        H5Fclose(fid);
        std::cout << "Saved to " << file_path << "\n";
    } catch (const std::exception& e) {
        std::cerr << "ltgs: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
