// acc_runs: merge seed-repeated run files into per-parameter .avg.h5 files.
// Restates the reference's merge tool (scripts/acc_runs.py:13-57, 176-438):
//   * every *.out.h5 given (or found in a given directory) is grouped by the
//     key=value tags of its name, ignoring seed=...; files without a seed
//     tag are skipped;
//   * per group: /energy E, E2, n_samples summed over seeds, with e_mean =
//     E/n, e2_mean = E2/n, var = e2_mean - e_mean^2 and n_seeds;
//     /ssf static_corr summed, static_corr_2 = sum over seeds of the
//     per-seed sums squared, var_inter (variance over seeds of the per-seed
//     mean) and var_intra (sum of the per-seed intrinsic variances, when
//     every file holds static_corr_2), with the per-seed sample count taken
//     as n_samples / K; /geometry and the /ssf metadata copied from the
//     first file, checked for agreement across files;
//   * the readers' derived quantities, written beside the sums: /energy/C =
//     var / T^2 per temperature (plot_ssf.py:181-186) and /ssf/S_mean =
//     Re static_corr / n_samples, the sample-normalised contracted S(q)
//     [n_corr][n_T][n_k] (plot_ssf.py:42-57);
//   * the output is <name with seed=<hex>_ replaced by mergeK_>.avg.h5 in the
//     directory of the first file.
// Pre-existing .avg.h5 files among the inputs are listed and deleted only
// with --delete-avg (the reference asks interactively).
#include <hdf5.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <iostream>
#include <map>
#include <regex>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "cli.hpp"

namespace fs = std::filesystem;

namespace {

// ---------------------------------------------------------------- HDF5 I/O
struct Array {
    std::vector<hsize_t> dims;
    std::vector<double> v;
    size_t size() const { return v.size(); }
};

struct H5 {
    hid_t id = -1;
    explicit H5(const std::string& path, bool write = false) {
        id = write ? H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)
                   : H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
        if (id < 0) throw std::runtime_error("cannot open " + path);
    }
    ~H5() {
        if (id >= 0) H5Fclose(id);
    }
    bool exists(const std::string& name) const {
        std::string cur;
        size_t p = 1;
        while (true) {
            const size_t q = name.find('/', p);
            cur = name.substr(0, q);
            if (H5Lexists(id, cur.c_str(), H5P_DEFAULT) <= 0) return false;
            if (q == std::string::npos) return true;
            p = q + 1;
        }
    }
    Array read(const std::string& name) const {
        hid_t ds = H5Dopen2(id, name.c_str(), H5P_DEFAULT);
        if (ds < 0) throw std::runtime_error("missing dataset " + name);
        hid_t sp = H5Dget_space(ds);
        Array a;
        const int nd = H5Sget_simple_extent_ndims(sp);
        a.dims.resize(size_t(std::max(nd, 0)));
        if (nd > 0) H5Sget_simple_extent_dims(sp, a.dims.data(), nullptr);
        size_t n = 1;
        for (hsize_t d : a.dims) n *= size_t(d);
        a.v.resize(n);
        const herr_t rc = n ? H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, a.v.data()) : 0;
        H5Sclose(sp);
        H5Dclose(ds);
        if (rc < 0) throw std::runtime_error("cannot read " + name);
        return a;
    }
    std::vector<std::string> read_strings(const std::string& name) const {
        hid_t ds = H5Dopen2(id, name.c_str(), H5P_DEFAULT);
        if (ds < 0) throw std::runtime_error("missing dataset " + name);
        hid_t sp = H5Dget_space(ds), ft = H5Dget_type(ds);
        hsize_t n = 1;
        if (H5Sget_simple_extent_ndims(sp) == 1) H5Sget_simple_extent_dims(sp, &n, nullptr);
        hid_t mt = H5Tcopy(H5T_C_S1);
        H5Tset_size(mt, H5T_VARIABLE);
        H5Tset_cset(mt, H5Tget_cset(ft));
        std::vector<char*> p(n, nullptr);
        const herr_t rc = H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, p.data());
        std::vector<std::string> out;
        for (char* c : p) out.emplace_back(c ? c : "");
        if (rc >= 0) H5Dvlen_reclaim(mt, sp, H5P_DEFAULT, p.data());
        H5Tclose(mt);
        H5Tclose(ft);
        H5Sclose(sp);
        H5Dclose(ds);
        if (rc < 0) throw std::runtime_error("cannot read " + name);
        return out;
    }
    std::vector<int> read_int_attr(const std::string& obj, const std::string& name) const {
        hid_t at = H5Aopen_by_name(id, obj.c_str(), name.c_str(), H5P_DEFAULT, H5P_DEFAULT);
        if (at < 0) throw std::runtime_error("missing attribute " + obj + "@" + name);
        hid_t sp = H5Aget_space(at);
        hsize_t n = 1;
        if (H5Sget_simple_extent_ndims(sp) == 1) H5Sget_simple_extent_dims(sp, &n, nullptr);
        std::vector<int> v(n);
        H5Aread(at, H5T_NATIVE_INT, v.data());
        H5Sclose(sp);
        H5Aclose(at);
        return v;
    }
};

hid_t make_group(hid_t f, const char* name) {
    hid_t g = H5Gcreate2(f, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (g < 0) throw std::runtime_error(std::string("cannot create group ") + name);
    return g;
}

void write(hid_t loc, const char* name, hid_t type, const std::vector<hsize_t>& dims, const void* data) {
    hid_t sp = dims.empty() ? H5Screate(H5S_SCALAR) : H5Screate_simple(int(dims.size()), dims.data(), nullptr);
    hid_t ds = H5Dcreate2(loc, name, type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (ds < 0) {
        H5Sclose(sp);
        throw std::runtime_error(std::string("cannot create dataset ") + name);
    }
    H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
    H5Dclose(ds);
    H5Sclose(sp);
}

void write_d(hid_t loc, const char* name, const Array& a) { write(loc, name, H5T_NATIVE_DOUBLE, a.dims, a.v.data()); }

void write_u64(hid_t loc, const char* name, const Array& a) {
    std::vector<unsigned long long> u(a.v.size());
    for (size_t i = 0; i < u.size(); ++i) u[i] = (unsigned long long)std::llround(a.v[i]);
    write(loc, name, H5T_NATIVE_ULLONG, a.dims, u.data());
}

void write_strings(hid_t loc, const char* name, const std::vector<std::string>& s) {
    hid_t st = H5Tcopy(H5T_C_S1);
    H5Tset_size(st, H5T_VARIABLE);
    H5Tset_cset(st, H5T_CSET_UTF8);
    std::vector<const char*> p;
    for (const auto& x : s) p.push_back(x.c_str());
    write(loc, name, st, {hsize_t(s.size())}, p.data());
    H5Tclose(st);
}

void write_int_attr(hid_t loc, const char* name, const std::vector<int>& v, bool scalar) {
    const hsize_t n = v.size();
    hid_t sp = scalar ? H5Screate(H5S_SCALAR) : H5Screate_simple(1, &n, nullptr);
    hid_t at = H5Acreate2(loc, name, H5T_NATIVE_INT, sp, H5P_DEFAULT, H5P_DEFAULT);
    H5Awrite(at, H5T_NATIVE_INT, v.data());
    H5Aclose(at);
    H5Sclose(sp);
}

// ---------------------------------------------------------------- names
std::string stem_of(const std::string& path) {
    std::string b = fs::path(path).filename().string();
    return std::regex_replace(b, std::regex(R"(\.(out|avg)\.h5$)"), "");
}

std::map<std::string, std::string> parse_tags(const std::string& path) {
    static const std::regex tag(R"((spiral_axis|[A-Za-z][A-Za-z0-9]*)=([^_]*))");
    std::map<std::string, std::string> t;
    const std::string s = stem_of(path);
    for (auto it = std::sregex_iterator(s.begin(), s.end(), tag); it != std::sregex_iterator(); ++it)
        t[(*it)[1]] = (*it)[2];
    return t;
}

// numeric tags compare by value ("0.10" == "0.1"), others as text
std::string normalise(const std::string& v) {
    char* end = nullptr;
    const double d = std::strtod(v.c_str(), &end);
    if (!v.empty() && end && *end == '\0') {
        char buf[64];
        std::snprintf(buf, sizeof buf, "#%.17g", d);
        return buf;
    }
    return v;
}

std::string group_key(const std::map<std::string, std::string>& tags) {
    std::string k;
    for (const auto& [name, v] : tags)
        if (name != "seed") k += name + "=" + normalise(v) + ";";
    return k;
}

std::string merged_name(const std::string& first, size_t K) {
    const std::string base = fs::path(first).filename().string();
    std::string out = std::regex_replace(base, std::regex(R"(seed=[0-9a-fA-F]+_)"), "merge" + std::to_string(K) + "_");
    out = std::regex_replace(out, std::regex(R"(\.out\.h5$)"), ".avg.h5");
    if (out == base) throw std::runtime_error("cannot form the merged name of " + base);
    return (fs::path(first).parent_path() / out).string();
}

bool allclose(const Array& a, const Array& b) {
    if (a.dims != b.dims) return false;
    for (size_t i = 0; i < a.size(); ++i)
        if (std::fabs(a.v[i] - b.v[i]) > 1e-8 + 1e-5 * std::fabs(b.v[i])) return false;
    return true;
}

// ---------------------------------------------------------------- merge
std::string merge_group(const std::vector<std::string>& files_in) {
    std::vector<std::string> files;
    for (const auto& f : files_in) {
        try {
            H5 h(f);
            h.read("/energy/T_list");
            h.read("/energy/E");
            h.read("/energy/E2");
            h.read("/energy/n_samples");
            files.push_back(f);
        } catch (const std::exception& e) {
            std::printf("Failed to process %s: %s\n", f.c_str(), e.what());
        }
    }
    if (files.empty()) throw std::runtime_error("no readable file in the group");
    const size_t K = files.size();

    Array T, E, E2, n;
    Array Ts, C, C2, Cin, ns, sl_pos, idx_cell, recip;
    std::vector<std::string> lookup;
    std::vector<int> n_spins, k_dims;
    bool have_intra = true;
    for (size_t f = 0; f < K; ++f) {
        H5 h(files[f]);
        const Array t = h.read("/energy/T_list"), e = h.read("/energy/E"), e2 = h.read("/energy/E2"),
                    nn = h.read("/energy/n_samples");
        if (f == 0) {
            T = t;
            E = e;
            E2 = e2;
            n = nn;
        } else {
            if (!allclose(t, T)) throw std::runtime_error("T_list mismatch in " + files[f]);
            for (size_t i = 0; i < E.size(); ++i) {
                E.v[i] += e.v[i];
                E2.v[i] += e2.v[i];
                n.v[i] += nn.v[i];
            }
        }
        const Array t_s = h.read("/ssf/T_list"), c = h.read("/ssf/static_corr"), nsamp = h.read("/ssf/n_samples"),
                    slp = h.read("/ssf/sl_positions");
        const auto lk = h.read_strings("/ssf/corr_lookup");
        const auto nsp = h.read_int_attr("/ssf", "n_spins"), kd = h.read_int_attr("/ssf", "k_dims");
        const bool has2 = h.exists("/ssf/static_corr_2");
        if (f == 0) {
            Ts = t_s;
            lookup = lk;
            sl_pos = slp;
            n_spins = nsp;
            k_dims = kd;
            C = c;
            C2 = c;
            for (auto& x : C2.v) x *= x;
            ns = nsamp;
            have_intra = has2;
            if (has2) Cin = h.read("/ssf/static_corr_2");
            idx_cell = h.read("/geometry/index_cell");
            recip = h.read("/geometry/recip_vectors");
        } else {
            if (!allclose(t_s, Ts)) throw std::runtime_error("Temperature range of file " + files[f] + " incompatible");
            if (lk != lookup) throw std::runtime_error("Stored correlators of file " + files[f] + " incompatible");
            if (!allclose(slp, sl_pos)) throw std::runtime_error("Sublattice convention incompatible in file " + files[f]);
            if (nsp != n_spins || kd != k_dims)
                throw std::runtime_error("SSF group attribute mismatch in " + files[f]);
            if (c.dims != C.dims) throw std::runtime_error("static_corr shape mismatch in " + files[f]);
            for (size_t i = 0; i < C.size(); ++i) {
                C.v[i] += c.v[i];
                C2.v[i] += c.v[i] * c.v[i];
            }
            for (size_t i = 0; i < ns.size(); ++i) ns.v[i] += nsamp.v[i];
            if (have_intra) {
                if (has2) {
                    const Array c2 = h.read("/ssf/static_corr_2");
                    for (size_t i = 0; i < Cin.size(); ++i) Cin.v[i] += c2.v[i];
                } else {
                    std::fprintf(stderr, "Warning: %s has no static_corr_2; dropping var_intra\n", files[f].c_str());
                    have_intra = false;
                }
            }
        }
    }

    Array e_mean = E, e2_mean = E2, var = E, heat = E;
    for (size_t i = 0; i < E.size(); ++i) {
        e_mean.v[i] = E.v[i] / n.v[i];
        e2_mean.v[i] = E2.v[i] / n.v[i];
        var.v[i] = e2_mean.v[i] - e_mean.v[i] * e_mean.v[i];
        heat.v[i] = var.v[i] / (T.v[i] * T.v[i]);
    }
    // per-seed sample count n_T / K broadcast along axis 1 of [n_corr, n_T, ...]
    if (C.dims.size() < 2 || C.dims[1] != ns.size()) throw std::runtime_error("static_corr has no T axis");
    size_t inner = 1;
    for (size_t d = 2; d < C.dims.size(); ++d) inner *= size_t(C.dims[d]);
    const size_t nT = ns.size();
    Array var_inter = C, var_intra = C;
    // S_mean: the real parts of static_corr over the per-T sample count
    Array s_mean;
    s_mean.dims.assign(C.dims.begin(), C.dims.end() - 1);
    s_mean.v.resize(C.size() / 2);
    for (size_t i = 0; i < s_mean.v.size(); ++i) {
        const size_t t = (i / (inner / 2)) % nT;
        s_mean.v[i] = C.v[2 * i] / ns.v[t];
    }
    for (size_t i = 0; i < C.size(); ++i) {
        const size_t t = (i / inner) % nT;
        const double nps = ns.v[t] / double(K);
        const double m = C.v[i] / (double(K) * nps);
        var_inter.v[i] = C2.v[i] / (double(K) * nps * nps) - m * m;
        if (have_intra) var_intra.v[i] = Cin.v[i] / nps - C2.v[i] / (nps * nps);
    }

    const std::string out = merged_name(files.front(), K);
    H5 o(out, true);
    hid_t g = make_group(o.id, "energy");
    write_d(g, "T_list", T);
    write_d(g, "E", E);
    write_d(g, "E2", E2);
    write_u64(g, "n_samples", n);
    write_d(g, "e_mean", e_mean);
    write_d(g, "e2_mean", e2_mean);
    write_d(g, "var", var);
    write_d(g, "C", heat);
    const unsigned long long kk = K;
    write(g, "n_seeds", H5T_NATIVE_ULLONG, {}, &kk);
    H5Gclose(g);
    g = make_group(o.id, "geometry");
    std::vector<long long> ic(idx_cell.v.begin(), idx_cell.v.end());
    write(g, "index_cell", H5T_NATIVE_LLONG, idx_cell.dims, ic.data());
    write_d(g, "recip_vectors", recip);
    H5Gclose(g);
    g = make_group(o.id, "ssf");
    write_d(g, "T_list", Ts);
    write_strings(g, "corr_lookup", lookup);
    write_u64(g, "n_samples", ns);
    write_d(g, "static_corr", C);
    write_d(g, "static_corr_2", C2);
    write_d(g, "var_inter", var_inter);
    write_d(g, "S_mean", s_mean);
    if (have_intra) write_d(g, "var_intra", var_intra);
    std::vector<long long> sp(sl_pos.v.begin(), sl_pos.v.end());
    write(g, "sl_positions", H5T_NATIVE_LLONG, sl_pos.dims, sp.data());
    write_int_attr(g, "n_spins", n_spins, true);
    write_int_attr(g, "k_dims", k_dims, false);
    H5Gclose(g);
    std::printf("%zu seed(s) merged -> %s\n", K, out.c_str());
    return out;
}

}  // namespace

int main(int argc, char* argv[]) {
    ArgParser prog("acc_runs");
    auto& paths = prog.add({"paths"}, "Directories and/or .out.h5 files to process");
    paths.required = true;
    paths.nargs = -1;
    prog.add({"--delete-avg"}, "Delete pre-existing .avg.h5 files among the inputs first").flag = true;
    try {
        prog.parse(argc, argv);
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    try {
        std::set<std::string> outs, avgs;
        auto ends = [](const std::string& s, const std::string& x) {
            return s.size() >= x.size() && s.compare(s.size() - x.size(), x.size(), x) == 0;
        };
        for (const auto& p : prog.raw("paths")) {
            if (fs::is_directory(p)) {
                for (const auto& e : fs::directory_iterator(p)) {
                    const std::string s = e.path().string();
                    if (ends(s, ".out.h5")) outs.insert(s);
                    if (ends(s, ".avg.h5")) avgs.insert(s);
                }
            } else if (ends(p, ".out.h5")) {
                outs.insert(p);
            } else if (ends(p, ".avg.h5")) {
                avgs.insert(p);
            } else {
                std::fprintf(stderr, "Warning: ignoring unrecognized path: %s\n", p.c_str());
            }
        }
        if (!avgs.empty()) {
            std::printf("Found %zu pre-existing .avg.h5 file(s):\n", avgs.size());
            for (const auto& f : avgs) std::printf("  %s\n", f.c_str());
            if (prog.flag("--delete-avg")) {
                for (const auto& f : avgs) fs::remove(f);
                std::printf("Deleted %zu file(s).\n", avgs.size());
            } else {
                std::printf("Leaving pre-existing .avg.h5 files in place.\n");
            }
        }
        if (outs.empty()) {
            std::fprintf(stderr, "No .out.h5 files found.\n");
            return 1;
        }
        std::map<std::string, std::vector<std::string>> groups;
        for (const auto& f : outs) {
            const auto tags = parse_tags(f);
            if (!tags.count("seed")) {
                std::fprintf(stderr, "Warning: no seed= tag in %s, skipping\n", fs::path(f).filename().c_str());
                continue;
            }
            groups[group_key(tags)].push_back(f);
        }
        if (groups.empty()) {
            std::fprintf(stderr, "No seed-tagged run files to merge.\n");
            return 1;
        }
        std::printf("Found %zu run file(s) in %zu group(s).\n", outs.size(), groups.size());
        for (auto& [key, files] : groups) {
            std::sort(files.begin(), files.end());
            merge_group(files);
        }
    } catch (const std::exception& e) {
        std::cerr << "acc_runs: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
