// C++ host interface of the sweep-and-measure path, over the C ABI of
// libpyro_cmc.so (include/pyro_cmc.h).  It keeps the reference's class and
// function names and semantics so that drivers written against pyro_cmc
// read the same:
//   CMC::Lattice / build_supercell / PyroCubicCell   (lattice_lib Supercell,
//                                                   pyrochlore_geometry.hpp:12-32)
//   CMC::MC_runner                                  (MC.hpp:87-140)
//   abstract_manager, energy_manager                (abstract_manager.hpp,
//                                                   energy_manager.hpp)
//   ssf_manager                                     (ssf_manager.hpp:84-309)
//   write_geometry_group, save_spin_state,
//   save_ft_spin_state                              (MC.cpp:320-454)
//   generate_T_profile, name_LJ123, J3_from_Qz, ... (cli_bits.hpp)
//
// Spins live on the GPU once a runner is set up on a lattice.  The
// Lattice's host copy is kept coherent: reading objects through
// get_objects() downloads the device state first, and writing through it
// marks the host copy as the newer one, which the next device operation
// uploads.  Errors from the C ABI become the reference's exception types
// (std::invalid_argument, std::logic_error, std::runtime_error).
#pragma once

#include <hdf5.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <filesystem>
#include <map>
#include <memory>
#include <numeric>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "pyro_cmc.h"

namespace vector3 {

template <typename T>
struct vec3 {
    T v[3] = {T(0), T(0), T(0)};
    vec3() = default;
    vec3(T a, T b, T c) : v{a, b, c} {}
    T& operator[](int i) { return v[i]; }
    const T& operator[](int i) const { return v[i]; }
    vec3& operator+=(const vec3& o) { for (int i = 0; i < 3; ++i) v[i] += o.v[i]; return *this; }
    vec3& operator-=(const vec3& o) { for (int i = 0; i < 3; ++i) v[i] -= o.v[i]; return *this; }
    vec3& operator/=(T s) { for (int i = 0; i < 3; ++i) v[i] /= s; return *this; }
    friend vec3 operator+(vec3 a, const vec3& b) { return a += b; }
    friend vec3 operator-(vec3 a, const vec3& b) { return a -= b; }
    friend vec3 operator*(T s, vec3 a) { for (int i = 0; i < 3; ++i) a.v[i] *= s; return a; }
    friend bool operator==(const vec3& a, const vec3& b) { return a.v[0] == b.v[0] && a.v[1] == b.v[1] && a.v[2] == b.v[2]; }
};
template <typename T>
T dot(const vec3<T>& a, const vec3<T>& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// 3x3 matrix, element (i, j) = row i, column j.
template <typename T>
struct mat33 {
    T m[9] = {};
    T& operator()(int i, int j) { return m[3 * i + j]; }
    const T& operator()(int i, int j) const { return m[3 * i + j]; }
    static mat33 from_cols(const std::array<T, 3>& a, const std::array<T, 3>& b, const std::array<T, 3>& c) {
        mat33 r;
        for (int i = 0; i < 3; ++i) {
            r(i, 0) = a[i];
            r(i, 1) = b[i];
            r(i, 2) = c[i];
        }
        return r;
    }
    mat33 tr() const {
        mat33 r;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) r(i, j) = (*this)(j, i);
        return r;
    }
    friend mat33 operator*(T s, mat33 a) { for (auto& x : a.m) x *= s; return a; }
    friend mat33 operator+(mat33 a, const mat33& b) { for (int i = 0; i < 9; ++i) a.m[i] += b.m[i]; return a; }
};

using vec3d = vec3<double>;
using mat33d = mat33<double>;

}  // namespace vector3

using ipos_t = vector3::vec3<int64_t>;
using imat33_t = vector3::mat33<int64_t>;

namespace CMC {

using dmat33_t = vector3::mat33<double>;

struct HeisenbergSpin {
    ipos_t ipos;
    int pyro_sl = 0;
    vector3::vec3d S;
    int8_t lifted_dir = 1;
};

// Sites of one orthorhombic unit cell (cell edges on the diagonal).
struct UnitCellSpecifier {
    imat33_t cell;
    std::vector<HeisenbergSpin> sites;
    explicit UnitCellSpecifier(const imat33_t& c) : cell(c) {}
    void add(HeisenbergSpin s) { sites.push_back(std::move(s)); }
};

class MC_runner;

// Device state behind a lattice: one C-ABI context and the replica index.
struct DeviceBinding {
    pyro_ctx* ctx = nullptr;
    int replica = 0;     // replica the Lattice view shows
    int n_replicas = 1;  // replicas held by the context
    ~DeviceBinding() {
        if (ctx) pyro_destroy(ctx);
    }
};

void throw_status(int code, const std::string& where);
inline void check(int code, const char* where) {
    if (code != PYRO_OK) throw_status(code, where);
}

// Supercell of Heisenberg spins: L0 x L1 x L2 copies of a unit cell, site
// index sl * n_cells + (n0 * L1 + n1) * L2 + n2.
class Lattice {
public:
    struct Bravais {
        int64_t L[3] = {1, 1, 1};
        int64_t a[3] = {1, 1, 1};  // cell edges
        int num_primitive_cells() const { return int(L[0] * L[1] * L[2]); }
        ipos_t size() const { return ipos_t{L[0], L[1], L[2]}; }
        imat33_t get_lattice_vectors() const {
            return imat33_t::from_cols({a[0], 0, 0}, {0, a[1], 0}, {0, 0, a[2]});
        }
        // q = 2 pi Q_i / (a_i L_i) for the integer wave-vector index Q, taken
        // literally (no folding into the first Brillouin zone).
        vector3::vec3d wavevector_from_idx3(const ipos_t& Q) const {
            return vector3::vec3d(2 * M_PI * double(Q[0]) / double(a[0] * L[0]),
                                  2 * M_PI * double(Q[1]) / double(a[1] * L[1]),
                                  2 * M_PI * double(Q[2]) / double(a[2] * L[2]));
        }
        vector3::mat33d get_reciprocal_lattice_vectors() const {
            return vector3::mat33d::from_cols({2 * M_PI / double(a[0]), 0, 0}, {0, 2 * M_PI / double(a[1]), 0},
                                              {0, 0, 2 * M_PI / double(a[2])});
        }
    } lattice;
    std::vector<ipos_t> sl_positions;
    std::vector<int> sl_kind;

    template <typename T>
    std::vector<T>& get_objects();
    template <typename T>
    const std::vector<T>& get_objects() const;
    template <typename T>
    T* get_object_at(const ipos_t& p);

    // device coherence
    std::shared_ptr<DeviceBinding> device;
    mutable bool device_newer = false;  // device holds changes not yet downloaded
    bool host_newer = true;             // host holds changes not yet uploaded
    void download() const;
    void upload();
    // Point the host view at replica r of the device context (host changes
    // are uploaded to the previous replica first).
    void view_replica(int r);

private:
    mutable std::vector<HeisenbergSpin> spins_;
    friend Lattice build_supercell(const UnitCellSpecifier&, const imat33_t&);
};

template <>
inline std::vector<HeisenbergSpin>& Lattice::get_objects<HeisenbergSpin>() {
    download();
    host_newer = true;  // the caller may write
    return spins_;
}
template <>
inline const std::vector<HeisenbergSpin>& Lattice::get_objects<HeisenbergSpin>() const {
    download();
    return spins_;
}

Lattice build_supercell(const UnitCellSpecifier& cell, const imat33_t& supercell);

struct CouplingSpec {
    std::string name;
    std::vector<std::vector<ipos_t>> relative_vectors;
    dmat33_t J;
    dmat33_t Jt;
};

struct MC_parameters {
    double T_ref = 1.0;
    size_t verbosity = 2;
    int precision = PYRO_FP64;  // PYRO_FP32 for the mixed-precision path
    int strategy = 0;           // sweep launch strategy (0 = automatic)
    int n_replicas = 1;         // replicas in the device context (tempering rungs)
};

class MC_runner {
public:
    MC_parameters settings;

    MC_runner(Lattice& lat, size_t seed) : lat_(&lat), seed_(seed) {}

    void rebind(Lattice& new_lat);
    double total_energy_per_unit_cell() const;
    void define_coupling(const std::string& name, const std::vector<std::vector<ipos_t>>& rel_vecs,
                         const vector3::mat33<double>& J);
    // The nine couplings of build_J1J2J3_h in one call (lets the library use
    // its pyrochlore stencil kernels).
    void define_J1J2J3(double J1, double J2, double J3, double Delta);
    void set_global_field(const vector3::vec3<double>& h);
    vector3::vec3d get_global_field() const { return field_; }
    void setup_lattice();

    void overrelax_all();
    size_t sweep_local_Metropolis(double T, size_t n_overrelax = 1) { return sweeps(PYRO_METROPOLIS, T, n_overrelax, 1); }
    size_t sweep_lifted_Metropolis(double T, size_t n_overrelax = 1) { return sweeps(PYRO_LIFTED, T, n_overrelax, 1); }
    size_t sweep_lifted_Metropolis_rot(double T, size_t n_overrelax = 1) { return sweeps(PYRO_LIFTED_ROT, T, n_overrelax, 1); }
    // n_sweeps sweeps in one call (one device synchronisation), every
    // replica at T; returns the accepted moves of the viewed replica.
    size_t sweeps(int kind, double T, size_t n_overrelax, size_t n_sweeps);
    // Uniform-on-sphere spins from the seed index (device kernel).
    void randomize_spins(uint64_t seed);

    pyro_ctx* context() const { return lat_->device ? lat_->device->ctx : nullptr; }
    const std::vector<CouplingSpec>& couplings() const { return coupling_specs_; }

private:
    Lattice* lat_;
    size_t seed_;
    std::vector<CouplingSpec> coupling_specs_;
    std::map<std::string, size_t> index_;
    vector3::vec3d field_;
    bool j1j2j3_ = false;
    double jparams_[4] = {0, 0, 0, 1};
    double T_set_ = -1.0;
    void ensure_ready() const;
    void set_T(double T);
};

void write_geometry_group(hid_t file_id, Lattice& sc, const char* group_name = "geometry");
void save_spin_state(Lattice& lat, const std::filesystem::path& file_path);
void save_ft_spin_state(Lattice& lat, const std::filesystem::path& file_path);

}  // namespace CMC

// ---------------------------------------------------------------------------
// Observable managers (abstract_manager.hpp, energy_manager.hpp).
// ---------------------------------------------------------------------------
class abstract_manager {
protected:
    std::vector<double> T_list;
    std::vector<size_t> n_samples;
    size_t curr_idx = 0;
    virtual void on_new_temp() {}

public:
    virtual ~abstract_manager() = default;
    void new_T(double T) {
        curr_idx = T_list.size();
        T_list.push_back(T);
        n_samples.push_back(0);
        on_new_temp();
    }
    void set_T(double T, double tol = 1e-8) {
        for (size_t i = 0; i < T_list.size(); ++i)
            if (std::abs(T_list[i] - T) < tol) {
                curr_idx = i;
                return;
            }
        new_T(T);
    }
    double curr_T() const { return T_list[curr_idx]; }
    virtual void write_group(hid_t file_id, const char* group_name) = 0;
};

class energy_manager : public abstract_manager {
    std::vector<double> E, E2;
    void on_new_temp() override {
        E.push_back(0);
        E2.push_back(0);
    }

public:
    explicit energy_manager(size_t n_temperatures_reserve = 0) {
        E.reserve(n_temperatures_reserve);
        E2.reserve(n_temperatures_reserve);
    }
    double curr_E() const { return E[curr_idx] / double(n_samples[curr_idx]); }
    void sample(double e) {
        if (T_list.empty()) throw std::logic_error("energy_manager::sample before new_T");
        E[curr_idx] += e;
        E2[curr_idx] += e * e;
        n_samples[curr_idx]++;
    }
    // Per-slot sums gathered elsewhere (the tempering path accumulates on
    // the device); appends one slot per entry.
    void append(double T, double Esum, double E2sum, size_t n) {
        new_T(T);
        E.back() = Esum;
        E2.back() = E2sum;
        n_samples.back() = n;
    }
    void save(const std::filesystem::path& file_path);
    void write_group(hid_t file_id, const char* group_name = "/energy") override;
};

// ---------------------------------------------------------------------------
// Static structure factor (ssf_manager.hpp).  Accumulation runs on the
// device; this class keeps the reference's bookkeeping and HDF5 layout.
// ---------------------------------------------------------------------------
struct CorrSpec {
    int alpha, beta;
    std::string label;
};

class ssf_manager : public abstract_manager {
public:
    ssf_manager(CMC::Lattice& sc, const std::vector<std::string>& correlator_names,
                size_t n_temperatures_reserve = 0, bool store_error_term = true);
    ssf_manager(CMC::Lattice& sc, const std::vector<std::string>& correlator_names, hid_t file_id,
                const char* group_name, const std::vector<double>& T_sample, bool store_error_term = true);
    // Tempering constructor (no counterpart in the reference, whose temper
    // gathers one manager per MPI rank): one device accumulator slot per
    // entry of T_slots, filled by the tempering rounds through a slot map
    // (pyro_pt_ssf_map); write_group writes every slot, ascending in T.
    // munu also accumulates the sublattice-resolved sums (write_munu).
    ssf_manager(CMC::Lattice& sc, const std::vector<std::string>& correlator_names,
                const std::vector<double>& T_slots, bool store_error_term, bool munu);
    ~ssf_manager();

    void point_at(CMC::Lattice& sc2);
    int get_n_sl() const { return n_sl_; }
    int get_n_kpoints() const { return n_kpoints_; }
    ipos_t get_k_dims() const { return k_dims_; }
    int get_n_spins() const { return n_spins_; }
    size_t get_n_corr() const { return corr_specs_.size(); }
    size_t get_n_samples_at(size_t t) const { return n_samples.at(t); }
    double get_T_at(size_t t) const { return T_list.at(t); }
    std::vector<std::string> get_corr_labels() const;
    const std::vector<ipos_t>& get_sl_positions() const { return sl_positions_; }
    void get_flat_buffer(size_t t_idx, std::vector<double>& corr_buf, std::vector<double>& corr_sq_buf) const;

    void sample();
    void flush();
    void write_group(hid_t file_id, const char* group_name = "/ssf") override;
    // Sublattice-resolved sums [n_corr, n_T, n_k, n_sl, n_sl, 2] (layout of
    // temper.cpp:121-159) as <group>/static_corr_munu; tempering mode with munu.
    void write_munu(hid_t file_id, const char* group_name = "/ssf");

private:
    CMC::Lattice* lat_;
    std::vector<CorrSpec> corr_specs_;
    int n_sl_, n_kpoints_, n_spins_;
    ipos_t k_dims_;
    std::vector<ipos_t> sl_positions_;
    const bool store_error_term_;
    bool streaming_mode_ = false;
    bool ladder_mode_ = false;
    bool munu_ = false;
    bool has_data_ = false;
    hid_t grp_ = -1, sum_ds_ = -1, sum_sq_ds_ = -1;
    std::vector<double> T_sorted_;

    void parse(const std::vector<std::string>& names);
    void configure_device(int n_slots, bool munu);
    void refresh_ladder_counts();
    void on_new_temp() override;
    void flatten_current(std::vector<double>& buf, std::vector<double>& sq) const;
    void write_metadata(hid_t grp) const;
};

// ---------------------------------------------------------------------------
// Model builders and schedules (cli_bits.hpp, pyrochlore_geometry.hpp).
// ---------------------------------------------------------------------------
namespace pyrochlore {
// displacement tables by pyro_sl from the library (geometry search)
std::vector<std::vector<ipos_t>> shell(int which);  // 1 nn1, 2 nn2, 3 nn3a, 4 nn3b
extern const vector3::vec3d axis_z[4];              // local [111] axes
}  // namespace pyrochlore

CMC::UnitCellSpecifier PyroCubicCell();
CMC::Lattice build_pyro_lat(int L);
// Couplings of build_J1J2J3_h (cli_bits.hpp:144-207) on a runner.
void build_J1J2J3_h(CMC::MC_runner& mc, double J1, double J2, double J3, double Delta,
                    const vector3::vec3d& B, double T_ref);
double J3_from_Qz(double J2, double Qz);
double round_Qz_to_supercell(double Qz, int L);
void init_spiral_state(CMC::Lattice& lat, double Q_rounded, int axis);
std::pair<std::vector<double>, std::set<size_t>> generate_T_profile(double T_hot, double T_cold,
                                                                     std::vector<double> T_sample,
                                                                     size_t total_steps);
// 64-bit RNG key of a seed index (SplitMix64 of the index).
uint64_t seed_key(uint64_t seed_index);
// Uniform-on-sphere spins from the seed (device kernel), all replicas.
void randomize_spins(CMC::MC_runner& mc, uint64_t seed);

static const char* DELIM = "_";
