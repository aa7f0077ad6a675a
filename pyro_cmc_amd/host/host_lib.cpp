// Implementation of the C++ host interface (pyro_host.hpp).
#include "pyro_host.hpp"

#include <cstring>

namespace {
void h5_check(hid_t id, const std::string& what) {
    if (id < 0) throw std::runtime_error("HDF5: failed to " + what);
}
void write_2d(hid_t loc, const char* name, hid_t type, hsize_t d0, hsize_t d1, const void* data) {
    const hsize_t dims[2] = {d0, d1};
    hid_t sp = H5Screate_simple(2, dims, nullptr);
    hid_t ds = H5Dcreate2(loc, name, type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (ds < 0) {
        H5Sclose(sp);
        throw std::runtime_error(std::string("HDF5: failed to create ") + name);
    }
    H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
    H5Dclose(ds);
    H5Sclose(sp);
}
}  // namespace

namespace CMC {

void throw_status(int code, const std::string& where) {
    const std::string msg = where + ": " + pyro_last_error();
    switch (code) {
        case PYRO_E_INVALID: throw std::invalid_argument(msg);
        case PYRO_E_LOGIC:
        case PYRO_E_STATE: throw std::logic_error(msg);
        default: throw std::runtime_error(msg);
    }
}

namespace {
int64_t fmod64(int64_t a, int64_t m) { return ((a % m) + m) % m; }
}  // namespace

Lattice build_supercell(const UnitCellSpecifier& cell, const imat33_t& supercell) {
    Lattice lat;
    for (int a = 0; a < 3; ++a) {
        for (int b = 0; b < 3; ++b)
            if (a != b && (cell.cell(a, b) != 0 || supercell(a, b) != 0))
                throw std::invalid_argument("build_supercell: only diagonal cells and supercells are supported");
        lat.lattice.a[a] = cell.cell(a, a);
        lat.lattice.L[a] = supercell(a, a);
        if (lat.lattice.a[a] <= 0 || lat.lattice.L[a] <= 0)
            throw std::invalid_argument("build_supercell: cell edges and extents must be positive");
    }
    const int64_t nc = lat.lattice.L[0] * lat.lattice.L[1] * lat.lattice.L[2];
    for (const auto& s : cell.sites) {
        lat.sl_positions.push_back(s.ipos);
        lat.sl_kind.push_back(s.pyro_sl);
    }
    lat.spins_.resize(cell.sites.size() * nc);
    for (size_t sl = 0; sl < cell.sites.size(); ++sl)
        for (int64_t n0 = 0; n0 < lat.lattice.L[0]; ++n0)
            for (int64_t n1 = 0; n1 < lat.lattice.L[1]; ++n1)
                for (int64_t n2 = 0; n2 < lat.lattice.L[2]; ++n2) {
                    HeisenbergSpin& sp = lat.spins_[sl * nc + (n0 * lat.lattice.L[1] + n1) * lat.lattice.L[2] + n2];
                    sp = cell.sites[sl];
                    sp.ipos = ipos_t{cell.sites[sl].ipos[0] + lat.lattice.a[0] * n0,
                                     cell.sites[sl].ipos[1] + lat.lattice.a[1] * n1,
                                     cell.sites[sl].ipos[2] + lat.lattice.a[2] * n2};
                }
    return lat;
}

template <>
HeisenbergSpin* Lattice::get_object_at<HeisenbergSpin>(const ipos_t& p) {
    const int64_t nc = lattice.L[0] * lattice.L[1] * lattice.L[2];
    for (size_t s = 0; s < sl_positions.size(); ++s) {
        int64_t n[3];
        bool ok = true;
        for (int a = 0; a < 3; ++a) {
            const int64_t d = p[a] - sl_positions[s][a];
            if (fmod64(d, lattice.a[a]) != 0) { ok = false; break; }
            n[a] = fmod64((d - fmod64(d, lattice.a[a])) / lattice.a[a], lattice.L[a]);
        }
        if (ok) return &get_objects<HeisenbergSpin>()[s * nc + (n[0] * lattice.L[1] + n[1]) * lattice.L[2] + n[2]];
    }
    return nullptr;
}

void Lattice::download() const {
    if (!device || !device_newer) return;
    const size_t N = spins_.size();
    std::vector<double> S(3 * N);
    std::vector<int8_t> d(N);
    check(pyro_get_spins(device->ctx, device->replica, S.data()), "get_spins");
    check(pyro_get_lifted_dir(device->ctx, device->replica, d.data()), "get_lifted_dir");
    for (size_t i = 0; i < N; ++i) {
        for (int a = 0; a < 3; ++a) spins_[i].S[a] = S[3 * i + a];
        spins_[i].lifted_dir = d[i];
    }
    device_newer = false;
}

void Lattice::upload() {
    if (!device || !host_newer) return;
    const size_t N = spins_.size();
    std::vector<double> S(3 * N);
    std::vector<int8_t> d(N);
    for (size_t i = 0; i < N; ++i) {
        for (int a = 0; a < 3; ++a) S[3 * i + a] = spins_[i].S[a];
        d[i] = spins_[i].lifted_dir >= 0 ? 1 : -1;
    }
    check(pyro_set_spins(device->ctx, device->replica, S.data()), "set_spins");
    check(pyro_set_lifted_dir(device->ctx, device->replica, d.data()), "set_lifted_dir");
    host_newer = false;
}

void Lattice::view_replica(int r) {
    if (!device) throw std::logic_error("view_replica needs a lattice set up on the device");
    if (r < 0 || r >= device->n_replicas) throw std::invalid_argument("view_replica: replica out of range");
    upload();
    device->replica = r;
    host_newer = false;
    device_newer = true;
}

void MC_runner::define_coupling(const std::string& name, const std::vector<std::vector<ipos_t>>& rel_vecs,
                                const vector3::mat33<double>& J) {
    if (index_.count(name)) throw std::logic_error("Coupling names must be unique");
    if (lat_->device) throw std::logic_error("define_coupling after setup_lattice");
    index_[name] = coupling_specs_.size();
    coupling_specs_.push_back({name, rel_vecs, J, J.tr()});
    j1j2j3_ = false;
}

void MC_runner::define_J1J2J3(double J1, double J2, double J3, double Delta) {
    const bool first = coupling_specs_.empty();
    // the nine specs of build_J1J2J3_h: six J1 sublattice pairs, J2, J3a, J3b
    const auto nn1 = pyrochlore::shell(1);
    static const ipos_t link[4] = {{1, 1, 1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, 1}};
    auto kind_of = [&](int p, const ipos_t& d) {
        // pyro_sl of the site at link[p] + d: the link vector equal mod 8 / fcc
        for (int q = 0; q < 4; ++q) {
            bool ok = true;
            for (int a = 0; a < 3; ++a) ok = ok && fmod64(link[p][a] + d[a] - link[q][a], 4) == 0;
            if (ok) return q;
        }
        return -1;
    };
    const int pairs[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    for (const auto& pr : pairs) {
        std::vector<std::vector<ipos_t>> rel(4);
        for (int p : {pr[0], pr[1]})
            for (const auto& d : nn1[p])
                if (kind_of(p, d) == (p == pr[0] ? pr[1] : pr[0])) rel[p].push_back(d);
        const auto& zm = pyrochlore::axis_z[pr[0]];
        const auto& zn = pyrochlore::axis_z[pr[1]];
        vector3::mat33d ising;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) ising(a, b) = zn[b] * zm[a];
        vector3::mat33d I = vector3::mat33d::from_cols({1, 0, 0}, {0, 1, 0}, {0, 0, 1});
        define_coupling("J1_" + std::to_string(pr[0]) + std::to_string(pr[1]), rel,
                        J1 * (Delta - 1) * ising + J1 * I);
    }
    const vector3::mat33d I = vector3::mat33d::from_cols({1, 0, 0}, {0, 1, 0}, {0, 0, 1});
    define_coupling("J2", pyrochlore::shell(2), J2 * I);
    define_coupling("J3a", pyrochlore::shell(3), J3 * I);
    define_coupling("J3b", pyrochlore::shell(4), J3 * I);
    j1j2j3_ = first;
    jparams_[0] = J1;
    jparams_[1] = J2;
    jparams_[2] = J3;
    jparams_[3] = Delta;
}

void MC_runner::set_global_field(const vector3::vec3<double>& h) {
    field_ = h;
    if (context()) check(pyro_set_field(context(), -1, field_.v), "set_field");
}

void MC_runner::setup_lattice() {
    if (lat_->device) throw std::logic_error("setup_lattice called twice");
    const int n_sl = int(lat_->sl_positions.size());
    std::vector<int32_t> pos(3 * n_sl), kind(n_sl);
    for (int s = 0; s < n_sl; ++s) {
        for (int a = 0; a < 3; ++a) pos[3 * s + a] = int32_t(lat_->sl_positions[s][a]);
        kind[s] = lat_->sl_kind[s];
    }
    const int32_t cell[3] = {int32_t(lat_->lattice.a[0]), int32_t(lat_->lattice.a[1]), int32_t(lat_->lattice.a[2])};
    const int32_t L[3] = {int32_t(lat_->lattice.L[0]), int32_t(lat_->lattice.L[1]), int32_t(lat_->lattice.L[2])};
    if (settings.n_replicas < 1) throw std::invalid_argument("MC_parameters::n_replicas must be positive");
    auto binding = std::make_shared<DeviceBinding>();
    binding->n_replicas = settings.n_replicas;
    check(pyro_create(n_sl, pos.data(), kind.data(), cell, L, settings.n_replicas, settings.precision,
                      seed_key(seed_), &binding->ctx),
          "create");
    if (j1j2j3_) {
        check(pyro_define_J1J2J3(binding->ctx, jparams_[0], jparams_[1], jparams_[2], jparams_[3]), "define_J1J2J3");
    } else {
        for (const auto& c : coupling_specs_) {
            std::vector<int32_t> counts, vecs;
            for (const auto& lst : c.relative_vectors) {
                counts.push_back(int32_t(lst.size()));
                for (const auto& v : lst)
                    for (int a = 0; a < 3; ++a) vecs.push_back(int32_t(v[a]));
            }
            if (vecs.empty()) vecs.assign(3, 0);
            check(pyro_define_coupling(binding->ctx, c.name.c_str(), int(counts.size()), counts.data(), vecs.data(),
                                       c.J.m),
                  "define_coupling");
        }
    }
    check(pyro_setup_lattice(binding->ctx), "setup_lattice");
    check(pyro_set_T_ref(binding->ctx, settings.T_ref), "set_T_ref");
    check(pyro_set_field(binding->ctx, -1, field_.v), "set_field");
    if (settings.strategy) check(pyro_set_strategy(binding->ctx, settings.strategy), "set_strategy");
    lat_->device = binding;
    lat_->host_newer = true;
    lat_->device_newer = false;
}

void MC_runner::ensure_ready() const {
    if (!lat_->device) throw std::logic_error("setup_lattice() has not been called");
    check(pyro_set_T_ref(lat_->device->ctx, settings.T_ref), "set_T_ref");
    lat_->upload();
}

void MC_runner::set_T(double T) {
    if (T != T_set_) {
        const std::vector<double> Ts(size_t(lat_->device->n_replicas), T);
        check(pyro_set_temperatures(context(), Ts.data()), "set_temperatures");
        T_set_ = T;
    }
}

size_t MC_runner::sweeps(int kind, double T, size_t n_overrelax, size_t n_sweeps) {
    ensure_ready();
    set_T(T);
    std::vector<uint64_t> acc(size_t(lat_->device->n_replicas), 0);
    check(pyro_sweep(context(), kind, int(n_overrelax), int(n_sweeps), acc.data()), "sweep");
    lat_->device_newer = true;
    return size_t(acc[size_t(lat_->device->replica)]);
}

void MC_runner::overrelax_all() {
    ensure_ready();
    check(pyro_overrelax(context(), 1), "overrelax");
    lat_->device_newer = true;
}

double MC_runner::total_energy_per_unit_cell() const {
    ensure_ready();
    std::vector<double> e(size_t(lat_->device->n_replicas), 0.0);
    check(pyro_energy_per_cell(context(), e.data()), "energy_per_cell");
    return e[size_t(lat_->device->replica)];
}

void MC_runner::rebind(Lattice& new_lat) {
    if (new_lat.get_objects<HeisenbergSpin>().size() != lat_->get_objects<HeisenbergSpin>().size())
        throw std::invalid_argument("rebind: lattices differ in size");
    auto binding = lat_->device;
    lat_->download();
    lat_->device.reset();
    lat_->device_newer = false;
    new_lat.device = binding;
    new_lat.host_newer = true;
    new_lat.device_newer = false;
    lat_ = &new_lat;
}

void MC_runner::randomize_spins(uint64_t seed) {
    if (!lat_->device) throw std::logic_error("setup_lattice() has not been called");
    check(pyro_randomize_spins(context(), seed_key(seed)), "randomize");
    lat_->host_newer = false;
    lat_->device_newer = true;
}

// ---------------------------------------------------------------------------
void write_geometry_group(hid_t file_id, Lattice& sc, const char* group_name) {
    hid_t grp = H5Gcreate2(file_id, group_name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    h5_check(grp, std::string("create group ") + group_name);
    const imat33_t A = sc.lattice.get_lattice_vectors();
    const vector3::mat33d B = sc.lattice.get_reciprocal_lattice_vectors();
    int64_t a[9];
    double b[9];
    for (int i = 0; i < 9; ++i) {
        a[i] = A(i / 3, i % 3);
        b[i] = B(i / 3, i % 3);
    }
    write_2d(grp, "index_cell", H5T_NATIVE_INT64, 3, 3, a);
    write_2d(grp, "recip_vectors", H5T_NATIVE_DOUBLE, 3, 3, b);
    H5Gclose(grp);
}

void save_spin_state(Lattice& lat, const std::filesystem::path& file_path) {
    const auto& spins = lat.get_objects<HeisenbergSpin>();
    const size_t N = spins.size();
    std::vector<int32_t> pos(3 * N);
    std::vector<double> ori(3 * N);
    for (size_t i = 0; i < N; ++i)
        for (int a = 0; a < 3; ++a) {
            pos[3 * i + a] = int32_t(spins[i].ipos[a]);
            ori[3 * i + a] = spins[i].S[a];
        }
    hid_t f = H5Fcreate(file_path.string().c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    h5_check(f, "create " + file_path.string());
    write_2d(f, "spin_pos", H5T_STD_I32LE, N, 3, pos.data());
    write_2d(f, "spin_orientation", H5T_IEEE_F64LE, N, 3, ori.data());
    H5Fclose(f);
}

void save_ft_spin_state(Lattice& lat, const std::filesystem::path& file_path) {
    if (!lat.device) throw std::logic_error("save_ft_spin_state needs a lattice set up on the device");
    lat.upload();
    const hsize_t n_sl = lat.sl_positions.size(), n_k = hsize_t(lat.lattice.num_primitive_cells());
    std::vector<double> ft(3 * n_sl * n_k * 2);
    check(pyro_spin_ft(lat.device->ctx, lat.device->replica, ft.data()), "spin_ft");
    hid_t f = H5Fcreate(file_path.string().c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    h5_check(f, "create " + file_path.string());
    const hsize_t dims[4] = {3, n_sl, n_k, 2};
    hid_t sp = H5Screate_simple(4, dims, nullptr);
    hid_t ds = H5Dcreate2(f, "spin_ft", H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, ft.data());
    H5Dclose(ds);
    H5Sclose(sp);
    H5Fclose(f);
}

}  // namespace CMC

// ---------------------------------------------------------------------------
void energy_manager::save(const std::filesystem::path& file_path) {
    hid_t f = H5Fcreate(file_path.string().c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) throw std::runtime_error("Failed to create HDF5 file: " + file_path.string());
    write_group(f);
    H5Fclose(f);
}

void energy_manager::write_group(hid_t file_id, const char* group_name) {
    hid_t grp = H5Gcreate2(file_id, group_name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (grp < 0) throw std::runtime_error("Failed to create group");
    const hsize_t dims[1] = {E.size()};
    hid_t sp = H5Screate_simple(1, dims, nullptr);
    auto w = [&](const char* name, hid_t type, const void* data) {
        hid_t ds = H5Dcreate2(grp, name, type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        if (ds < 0) {
            H5Sclose(sp);
            H5Gclose(grp);
            throw std::runtime_error(std::string("Failed to create dataset ") + name);
        }
        const herr_t st = H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
        H5Dclose(ds);
        if (st < 0) throw std::runtime_error(std::string("Failed to write dataset ") + name);
    };
    std::vector<unsigned long long> n(n_samples.begin(), n_samples.end());
    w("E", H5T_NATIVE_DOUBLE, E.data());
    w("E2", H5T_NATIVE_DOUBLE, E2.data());
    w("T_list", H5T_NATIVE_DOUBLE, T_list.data());
    w("n_samples", H5T_NATIVE_ULLONG, n.data());
    H5Sclose(sp);
    H5Gclose(grp);
}

// ---------------------------------------------------------------------------
void ssf_manager::parse(const std::vector<std::string>& names) {
    auto axis_idx = [](char c) -> int {
        if (c == 'x') return 0;
        if (c == 'y') return 1;
        if (c == 'z') return 2;
        throw std::invalid_argument(std::string("ssf_manager: unknown spin component '") + c + "'");
    };
    for (const auto& n : names) {
        if (n.size() != 2) throw std::invalid_argument("ssf_manager: correlator name must be 2 chars: " + n);
        corr_specs_.push_back({axis_idx(n[0]), axis_idx(n[1]), n});
    }
}

void ssf_manager::configure_device(int n_slots, bool munu) {
    if (!lat_->device)
        throw std::logic_error("ssf_manager needs a lattice set up on the device (MC_runner::setup_lattice)");
    std::vector<int32_t> a, b;
    for (const auto& c : corr_specs_) {
        a.push_back(c.alpha);
        b.push_back(c.beta);
    }
    CMC::check(pyro_ssf_configure(lat_->device->ctx, int(a.size()), a.data(), b.data(), n_slots,
                                  store_error_term_ ? 1 : 0, munu ? 1 : 0, munu ? 16 : 0),
               "ssf_configure");
}

ssf_manager::ssf_manager(CMC::Lattice& sc, const std::vector<std::string>& names, size_t, bool store_error_term)
    : lat_(&sc),
      n_sl_(int(sc.sl_positions.size())),
      n_kpoints_(sc.lattice.num_primitive_cells()),
      n_spins_(int(sc.sl_positions.size()) * sc.lattice.num_primitive_cells()),
      k_dims_(sc.lattice.size()),
      sl_positions_(sc.sl_positions),
      store_error_term_(store_error_term) {
    parse(names);
    configure_device(1, false);
}

ssf_manager::ssf_manager(CMC::Lattice& sc, const std::vector<std::string>& names, hid_t file_id,
                         const char* group_name, const std::vector<double>& T_sample, bool store_error_term)
    : ssf_manager(sc, names, 0, store_error_term) {
    streaming_mode_ = true;
    T_sorted_ = T_sample;
    std::sort(T_sorted_.begin(), T_sorted_.end());
    grp_ = H5Lexists(file_id, group_name, H5P_DEFAULT) > 0
               ? H5Gopen2(file_id, group_name, H5P_DEFAULT)
               : H5Gcreate2(file_id, group_name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (grp_ < 0) throw std::runtime_error(std::string("ssf_manager: failed to open/create group ") + group_name);
    const hsize_t dims[4] = {corr_specs_.size(), T_sorted_.size(), hsize_t(n_kpoints_), 2};
    hid_t fs = H5Screate_simple(4, dims, nullptr);
    hid_t dcpl = H5Pcreate(H5P_DATASET_CREATE);
    H5Pset_layout(dcpl, H5D_CONTIGUOUS);
    H5Pset_alloc_time(dcpl, H5D_ALLOC_TIME_EARLY);
    sum_ds_ = H5Dcreate2(grp_, "static_corr", H5T_NATIVE_DOUBLE, fs, H5P_DEFAULT, dcpl, H5P_DEFAULT);
    if (store_error_term_)
        sum_sq_ds_ = H5Dcreate2(grp_, "static_corr_2", H5T_NATIVE_DOUBLE, fs, H5P_DEFAULT, dcpl, H5P_DEFAULT);
    H5Pclose(dcpl);
    H5Sclose(fs);
    if (sum_ds_ < 0 || (store_error_term_ && sum_sq_ds_ < 0))
        throw std::runtime_error("ssf_manager: failed to create static_corr");
}

ssf_manager::ssf_manager(CMC::Lattice& sc, const std::vector<std::string>& names, const std::vector<double>& T_slots,
                         bool store_error_term, bool munu)
    : lat_(&sc),
      n_sl_(int(sc.sl_positions.size())),
      n_kpoints_(sc.lattice.num_primitive_cells()),
      n_spins_(int(sc.sl_positions.size()) * sc.lattice.num_primitive_cells()),
      k_dims_(sc.lattice.size()),
      sl_positions_(sc.sl_positions),
      store_error_term_(store_error_term) {
    if (T_slots.empty()) throw std::invalid_argument("ssf_manager: no temperature slots");
    parse(names);
    configure_device(int(T_slots.size()), munu);
    ladder_mode_ = true;
    munu_ = munu;
    T_list = T_slots;
    n_samples.assign(T_slots.size(), 0);
    curr_idx = 0;
}

void ssf_manager::refresh_ladder_counts() {
    for (size_t j = 0; j < T_list.size(); ++j) {
        uint64_t ns = 0;
        CMC::check(pyro_ssf_read(lat_->device->ctx, int(j), nullptr, nullptr, &ns), "ssf_read");
        n_samples[j] = size_t(ns);
    }
}

void ssf_manager::write_munu(hid_t file_id, const char* group_name) {
    if (!ladder_mode_ || !munu_) throw std::logic_error("write_munu needs a tempering-mode ssf_manager with munu");
    hid_t grp = H5Lexists(file_id, group_name, H5P_DEFAULT) > 0
                    ? H5Gopen2(file_id, group_name, H5P_DEFAULT)
                    : H5Gcreate2(file_id, group_name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    h5_check(grp, std::string("open group ") + group_name);
    const size_t n_T = T_list.size();
    std::vector<size_t> ord(n_T);
    std::iota(ord.begin(), ord.end(), 0);
    std::sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return T_list[a] < T_list[b]; });
    const hsize_t nc = corr_specs_.size(), nk = hsize_t(n_kpoints_), ns = hsize_t(n_sl_);
    const hsize_t dims[6] = {nc, n_T, nk, ns, ns, 2};
    hid_t fs = H5Screate_simple(6, dims, nullptr);
    hid_t ds = H5Dcreate2(grp, "static_corr_munu", H5T_NATIVE_DOUBLE, fs, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    H5Sclose(fs);
    if (ds < 0) {
        H5Gclose(grp);
        throw std::runtime_error("ssf_manager: failed to create static_corr_munu");
    }
    const hsize_t md = nk * ns * ns * 2;
    std::vector<double> C(md);
    hid_t ms = H5Screate_simple(1, &md, nullptr);
    herr_t rc = 0;
    for (hsize_t c = 0; c < nc; ++c)
        for (hsize_t t = 0; t < n_T; ++t) {
            CMC::check(pyro_ssf_read_munu(lat_->device->ctx, int(ord[t]), int(c), C.data()), "ssf_read_munu");
            const hsize_t start[6] = {c, t, 0, 0, 0, 0}, count[6] = {1, 1, nk, ns, ns, 2};
            hid_t f2 = H5Dget_space(ds);
            H5Sselect_hyperslab(f2, H5S_SELECT_SET, start, nullptr, count, nullptr);
            if (H5Dwrite(ds, H5T_NATIVE_DOUBLE, ms, f2, H5P_DEFAULT, C.data()) < 0) rc = -1;
            H5Sclose(f2);
        }
    H5Sclose(ms);
    H5Dclose(ds);
    H5Gclose(grp);
    if (rc < 0) throw std::runtime_error("ssf_manager: H5Dwrite of static_corr_munu failed");
}

ssf_manager::~ssf_manager() {
    if (sum_ds_ >= 0) H5Dclose(sum_ds_);
    if (sum_sq_ds_ >= 0) H5Dclose(sum_sq_ds_);
    if (grp_ >= 0) H5Gclose(grp_);
}

void ssf_manager::point_at(CMC::Lattice& sc2) {
    const bool same = sc2.device && lat_->device && sc2.device->ctx == lat_->device->ctx;
    if (!same && ladder_mode_) throw std::logic_error("point_at: a tempering-mode manager stays on its context");
    lat_ = &sc2;
    if (!same) configure_device(1, false);
}

std::vector<std::string> ssf_manager::get_corr_labels() const {
    std::vector<std::string> l;
    for (const auto& c : corr_specs_) l.push_back(c.label);
    return l;
}

void ssf_manager::on_new_temp() {
    if (ladder_mode_) throw std::logic_error("new_T: a tempering-mode manager has fixed temperature slots");
    CMC::check(pyro_ssf_reset(lat_->device->ctx, 0), "ssf_reset");
    has_data_ = true;
}

void ssf_manager::sample() {
    if (ladder_mode_) throw std::logic_error("sample: tempering-mode slots are filled by the tempering rounds");
    if (T_list.empty() || !has_data_) throw std::logic_error("ssf_manager::sample() called without new_T()");
    lat_->upload();
    const int replica = lat_->device->replica;
    int64_t info[8];
    CMC::check(pyro_info(lat_->device->ctx, info), "info");
    std::vector<int32_t> slot(size_t(info[3]), -1);
    slot[replica] = 0;
    CMC::check(pyro_ssf_sample(lat_->device->ctx, slot.data()), "ssf_sample");
    n_samples[curr_idx]++;
}

void ssf_manager::flatten_current(std::vector<double>& buf, std::vector<double>& sq) const {
    const size_t n = corr_specs_.size() * size_t(n_kpoints_) * 2;
    buf.assign(n, 0.0);
    if (store_error_term_) sq.assign(n, 0.0);
    uint64_t ns = 0;
    const int slot = ladder_mode_ ? int(curr_idx) : 0;
    CMC::check(pyro_ssf_read(lat_->device->ctx, slot, buf.data(), store_error_term_ ? sq.data() : nullptr, &ns),
               "ssf_read");
}

void ssf_manager::get_flat_buffer(size_t t_idx, std::vector<double>& corr_buf, std::vector<double>& corr_sq_buf) const {
    if (ladder_mode_) {
        if (t_idx >= T_list.size()) throw std::out_of_range("get_flat_buffer: no such temperature slot");
        const size_t n = corr_specs_.size() * size_t(n_kpoints_) * 2;
        corr_buf.assign(n, 0.0);
        if (store_error_term_) corr_sq_buf.assign(n, 0.0);
        uint64_t ns = 0;
        CMC::check(pyro_ssf_read(lat_->device->ctx, int(t_idx), corr_buf.data(),
                                 store_error_term_ ? corr_sq_buf.data() : nullptr, &ns),
                   "ssf_read");
        return;
    }
    if (t_idx != curr_idx) throw std::logic_error("get_flat_buffer: only the current temperature is held");
    flatten_current(corr_buf, corr_sq_buf);
}

void ssf_manager::flush() {
    if (!streaming_mode_) throw std::logic_error("flush() called on non-streaming ssf_manager");
    if (T_list.empty() || !has_data_) throw std::logic_error("flush() called with no accumulated data");
    const double T = T_list[curr_idx];
    auto it = std::find(T_sorted_.begin(), T_sorted_.end(), T);
    if (it == T_sorted_.end()) throw std::runtime_error("ssf_manager::flush: current temperature not found in T_sample");
    const hsize_t slot = hsize_t(it - T_sorted_.begin());
    std::vector<double> buf, sq;
    flatten_current(buf, sq);
    const hsize_t nc = corr_specs_.size(), nk = hsize_t(n_kpoints_);
    const hsize_t start[4] = {0, slot, 0, 0}, count[4] = {nc, 1, nk, 2};
    const hsize_t md = nc * nk * 2;
    hid_t ms = H5Screate_simple(1, &md, nullptr);
    herr_t rc = 0;
    for (int pass = 0; pass < (store_error_term_ ? 2 : 1); ++pass) {
        hid_t ds = pass == 0 ? sum_ds_ : sum_sq_ds_;
        hid_t fs = H5Dget_space(ds);
        H5Sselect_hyperslab(fs, H5S_SELECT_SET, start, nullptr, count, nullptr);
        const herr_t r = H5Dwrite(ds, H5T_NATIVE_DOUBLE, ms, fs, H5P_DEFAULT, pass == 0 ? buf.data() : sq.data());
        H5Sclose(fs);
        if (r < 0) rc = r;
    }
    H5Sclose(ms);
    if (rc < 0) throw std::runtime_error("ssf_manager::flush: H5Dwrite failed");
    has_data_ = false;
}

void ssf_manager::write_metadata(hid_t grp) const {
    const size_t n_T = T_list.size();
    std::vector<size_t> ord(n_T);
    std::iota(ord.begin(), ord.end(), 0);
    std::sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return T_list[a] < T_list[b]; });
    std::vector<double> sT(n_T);
    std::vector<unsigned long long> sN(n_T);
    for (size_t i = 0; i < n_T; ++i) {
        sT[i] = T_list[ord[i]];
        sN[i] = n_samples[ord[i]];
    }
    auto w1 = [&](const char* name, hid_t type, hsize_t len, const void* data) {
        hid_t sp = H5Screate_simple(1, &len, nullptr);
        hid_t ds = H5Dcreate2(grp, name, type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        if (ds < 0) {
            H5Sclose(sp);
            throw std::runtime_error(std::string("ssf_manager: failed to create ") + name);
        }
        H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
        H5Dclose(ds);
        H5Sclose(sp);
    };
    w1("T_list", H5T_NATIVE_DOUBLE, n_T, sT.data());
    w1("n_samples", H5T_NATIVE_ULLONG, n_T, sN.data());
    {
        hid_t st = H5Tcopy(H5T_C_S1);
        H5Tset_size(st, H5T_VARIABLE);
        H5Tset_cset(st, H5T_CSET_UTF8);
        const hsize_t nc = corr_specs_.size();
        hid_t sp = H5Screate_simple(1, &nc, nullptr);
        hid_t ds = H5Dcreate2(grp, "corr_lookup", st, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        std::vector<const char*> p;
        for (const auto& c : corr_specs_) p.push_back(c.label.c_str());
        H5Dwrite(ds, st, H5S_ALL, H5S_ALL, H5P_DEFAULT, p.data());
        H5Dclose(ds);
        H5Sclose(sp);
        H5Tclose(st);
    }
    {
        std::vector<int64_t> pos;
        for (const auto& s : sl_positions_)
            for (int a = 0; a < 3; ++a) pos.push_back(s[a]);
        write_2d(grp, "sl_positions", H5T_NATIVE_INT64, sl_positions_.size(), 3, pos.data());
    }
    {
        hid_t sp = H5Screate(H5S_SCALAR);
        hid_t at = H5Acreate2(grp, "n_spins", H5T_NATIVE_INT, sp, H5P_DEFAULT, H5P_DEFAULT);
        H5Awrite(at, H5T_NATIVE_INT, &n_spins_);
        H5Aclose(at);
        H5Sclose(sp);
        const hsize_t three = 3;
        const int kd[3] = {int(k_dims_[0]), int(k_dims_[1]), int(k_dims_[2])};
        sp = H5Screate_simple(1, &three, nullptr);
        at = H5Acreate2(grp, "k_dims", H5T_NATIVE_INT, sp, H5P_DEFAULT, H5P_DEFAULT);
        H5Awrite(at, H5T_NATIVE_INT, kd);
        H5Aclose(at);
        H5Sclose(sp);
    }
}

void ssf_manager::write_group(hid_t file_id, const char* group_name) {
    if (streaming_mode_) {
        if (has_data_ && !T_list.empty()) flush();
        write_metadata(grp_);
        if (sum_ds_ >= 0) { H5Dclose(sum_ds_); sum_ds_ = -1; }
        if (sum_sq_ds_ >= 0) { H5Dclose(sum_sq_ds_); sum_sq_ds_ = -1; }
        if (grp_ >= 0) { H5Gclose(grp_); grp_ = -1; }
        return;
    }
    hid_t grp = H5Lexists(file_id, group_name, H5P_DEFAULT) > 0
                    ? H5Gopen2(file_id, group_name, H5P_DEFAULT)
                    : H5Gcreate2(file_id, group_name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (grp < 0) throw std::runtime_error(std::string("ssf_manager: failed to open/create group ") + group_name);
    const hsize_t nc = corr_specs_.size(), nk = hsize_t(n_kpoints_);
    const hsize_t dims[4] = {nc, T_list.size(), nk, 2};
    hid_t fs = H5Screate_simple(4, dims, nullptr);
    // (T index, file slot) pairs to write: the current temperature only, as
    // in the reference, or every slot (ascending T) in tempering mode.
    std::vector<std::pair<size_t, hsize_t>> rows;
    if (ladder_mode_) {
        refresh_ladder_counts();
        std::vector<size_t> ord(T_list.size());
        std::iota(ord.begin(), ord.end(), 0);
        std::sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return T_list[a] < T_list[b]; });
        for (size_t t = 0; t < ord.size(); ++t) rows.push_back({ord[t], hsize_t(t)});
    } else if (has_data_) {
        rows.push_back({curr_idx, hsize_t(curr_idx)});
    }
    const hsize_t md = nc * nk * 2;
    hid_t ms = H5Screate_simple(1, &md, nullptr);
    hid_t dsets[2] = {-1, -1};
    for (int pass = 0; pass < (store_error_term_ ? 2 : 1); ++pass) {
        dsets[pass] = H5Dcreate2(grp, pass == 0 ? "static_corr" : "static_corr_2", H5T_NATIVE_DOUBLE, fs,
                                 H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        if (dsets[pass] < 0) throw std::runtime_error("ssf_manager: failed to create static_corr or static_corr_2");
    }
    std::vector<double> buf, sq;
    for (const auto& [t_idx, file_slot] : rows) {
        if (ladder_mode_) {
            get_flat_buffer(t_idx, buf, sq);
        } else {
            flatten_current(buf, sq);
        }
        const hsize_t start[4] = {0, file_slot, 0, 0}, count[4] = {nc, 1, nk, 2};
        for (int pass = 0; pass < (store_error_term_ ? 2 : 1); ++pass) {
            hid_t f2 = H5Dget_space(dsets[pass]);
            H5Sselect_hyperslab(f2, H5S_SELECT_SET, start, nullptr, count, nullptr);
            H5Dwrite(dsets[pass], H5T_NATIVE_DOUBLE, ms, f2, H5P_DEFAULT, pass == 0 ? buf.data() : sq.data());
            H5Sclose(f2);
        }
    }
    for (hid_t d : dsets)
        if (d >= 0) H5Dclose(d);
    H5Sclose(ms);
    H5Sclose(fs);
    write_metadata(grp);
    H5Gclose(grp);
}

// ---------------------------------------------------------------------------
namespace pyrochlore {

std::vector<std::vector<ipos_t>> shell(int which) {
    int32_t counts[4];
    CMC::check(pyro_pyrochlore_shell(which, counts, nullptr, 0), "pyrochlore_shell");
    const int total = counts[0] + counts[1] + counts[2] + counts[3];
    std::vector<int32_t> v(3 * size_t(total));
    CMC::check(pyro_pyrochlore_shell(which, counts, v.data(), total), "pyrochlore_shell");
    std::vector<std::vector<ipos_t>> out(4);
    int o = 0;
    for (int p = 0; p < 4; ++p)
        for (int i = 0; i < counts[p]; ++i, ++o) out[p].push_back(ipos_t{v[3 * o], v[3 * o + 1], v[3 * o + 2]});
    return out;
}

namespace {
const double kS3 = 1.732050807568877293527446341505872366942805253810380628055;
}
const vector3::vec3d axis_z[4] = {(1. / kS3) * vector3::vec3d(1, 1, 1), (1. / kS3) * vector3::vec3d(1, -1, -1),
                                  (1. / kS3) * vector3::vec3d(-1, 1, -1), (1. / kS3) * vector3::vec3d(-1, -1, 1)};

}  // namespace pyrochlore

CMC::UnitCellSpecifier PyroCubicCell() {
    CMC::UnitCellSpecifier spec(imat33_t::from_cols({8, 0, 0}, {0, 8, 0}, {0, 0, 8}));
    const ipos_t fcc[4] = {{0, 0, 0}, {0, 4, 4}, {4, 0, 4}, {4, 4, 0}};
    const ipos_t link[4] = {{1, 1, 1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, 1}};
    for (int f = 0; f < 4; ++f)
        for (int p = 0; p < 4; ++p) {
            CMC::HeisenbergSpin s;
            s.ipos = link[p] + fcc[f];
            s.pyro_sl = p;
            spec.add(s);
        }
    return spec;
}

CMC::Lattice build_pyro_lat(int L) {
    return CMC::build_supercell(PyroCubicCell(), imat33_t::from_cols({L, 0, 0}, {0, L, 0}, {0, 0, L}));
}

void build_J1J2J3_h(CMC::MC_runner& mc, double J1, double J2, double J3, double Delta, const vector3::vec3d& B,
                    double T_ref) {
    mc.define_J1J2J3(J1, J2, J3, Delta);
    mc.set_global_field(B);
    mc.settings.T_ref = T_ref;
    mc.setup_lattice();
}

double J3_from_Qz(double J2, double Qz) {
    const double q = Qz * 2 * M_PI / 8;  // 8: the cubic cell edge
    return (-4 * J2 + 1. / std::cos(2 * q) - 2. * J2 / std::cos(2 * q)) / 8.;
}

double round_Qz_to_supercell(double Qz, int L) {
    const double step = 1. / L;
    return std::round(Qz / step) * step;
}

void init_spiral_state(CMC::Lattice& lat, double Q_rounded, int axis) {
    const int i_sin = (axis + 1) % 3, i_cos = (axis + 2) % 3;
    for (auto& s : lat.get_objects<CMC::HeisenbergSpin>()) {
        const double ph = Q_rounded * 2 * M_PI * double(s.ipos[axis]) / 8.0;
        vector3::vec3d S(0., 0., 0.);
        S[i_cos] = std::cos(ph);
        S[i_sin] = std::sin(ph);
        s.S = S;
    }
}

std::pair<std::vector<double>, std::set<size_t>> generate_T_profile(double T_hot, double T_cold,
                                                                     std::vector<double> T_sample,
                                                                     size_t total_steps) {
    // log grid T_hot * f^i, i < total_steps, f = (T_cold/T_hot)^(1/total_steps)
    // (it stops one factor short of T_cold), merged with T_sample, all
    // descending; the set holds the indices of the T_sample entries.
    std::vector<double> sweep, out;
    const double f = std::pow(T_cold / T_hot, 1. / double(total_steps));
    double t = T_hot / f;
    for (size_t i = 0; i < total_steps; ++i) {
        t *= f;
        sweep.push_back(t);
    }
    auto desc = [](double a, double b) { return a > b; };
    std::sort(sweep.begin(), sweep.end(), desc);
    std::sort(T_sample.begin(), T_sample.end(), desc);
    std::set<size_t> idx;
    size_t i = 0, j = 0;
    while (i < sweep.size() && j < T_sample.size()) {
        if (sweep[i] > T_sample[j]) {
            out.push_back(sweep[i++]);
        } else {
            idx.insert(out.size());
            out.push_back(T_sample[j++]);
        }
    }
    for (; j < T_sample.size(); ++j) {
        idx.insert(out.size());
        out.push_back(T_sample[j]);
    }
    for (; i < sweep.size(); ++i) out.push_back(sweep[i]);
    return {out, idx};
}

uint64_t seed_key(uint64_t seed_index) {
    uint64_t z = seed_index + 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

void randomize_spins(CMC::MC_runner& mc, uint64_t seed) { mc.randomize_spins(seed); }
