// Model flags shared by the anneal and temper drivers (declare_LJ123,
// name_LJ123 and the J3/Q resolution of cli_bits.hpp:60-127).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <sstream>
#include <string>

#include "cli.hpp"
#include "pyro_host.hpp"

inline void declare_physical(ArgParser& p) {
    p.add({"L"}, "Linear dimension of the cubic supercell").required = true;
    p.add({"--J1"}, "Nearest-neighbour Heisenberg coupling strength").required = true;
    p.add({"--J2"}, "Second-nearest-neighbour coupling").def = {"0"};
    p.add({"--J3"}, "Third-nearest-neighbour coupling (exclusive with --Q)").def = {"0"};
    p.add({"--K"}, "Biquadratic coupling (accepted, unused, as in the reference)").def = {"0"};
    p.add({"--Q"}, "Spiral wavevector in units of 2pi/a; sets J3 (exclusive with --J3)");
    p.add({"--spiral_axis", "-x"}, "Axis (0,1,2) of --Q").def = {"2"};
    p.add({"--Delta"}, "Nearest-neighbour XXZ anisotropy in the local [111] frame").def = {"1.0"};
    auto& b = p.add({"--external_field", "-B"}, "Global magnetic field (3 values)");
    b.nargs = 3;
    b.def = {"0", "0", "0"};
}

inline double resolve_J3(const ArgParser& p) {
    if (p.is_used("--J3") && p.is_used("--Q")) throw std::runtime_error("--J3 and --Q are mutually exclusive");
    if (p.is_used("--Q")) {
        const int L = int(p.integer("L"));
        const double Qz = round_Qz_to_supercell(p.dbl("--Q"), L);
        const double dQ = std::abs(p.dbl("--Q") - Qz), tol = std::min(1e-10, 1e-3 * std::abs(p.dbl("--Q")));
        if (dQ > tol) std::fprintf(stderr, "WARNING: Qz=%.10g rounded to %.10g (diff=%.3e > tol=%.3e)\n", p.dbl("--Q"), Qz, dQ, tol);
        const double J1 = p.dbl("--J1");
        const double J3 = J3_from_Qz(p.dbl("--J2") / std::abs(J1), Qz) * std::abs(J1);
        std::cout << "Calculated J3=" << J3 << std::endl;
        return J3;
    }
    return p.dbl("--J3");
}

inline std::string name_LJ123(const ArgParser& p, double J3) {
    std::stringstream n;
    const int L = int(p.integer("L"));
    n << "L=" << L << DELIM << "J1=" << p.dbl("--J1") << DELIM << "J2=" << p.dbl("--J2") << DELIM << "J3=" << J3
      << DELIM << "Delta=" << p.dbl("--Delta") << DELIM;
    if (p.is_used("--Q")) {
        n << "Q=" << round_Qz_to_supercell(p.dbl("--Q"), L) << DELIM;
        const int axis = int(p.integer("--spiral_axis"));
        if (axis != 2) n << "spiral_axis=" << axis << DELIM;
    }
    return n.str();
}

