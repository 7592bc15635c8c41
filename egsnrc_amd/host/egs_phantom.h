// egs_phantom.h -- .egsphant voxel phantoms and CT ramps for the host
// application (egs++/geometry/egs_nd_geometry/egs_nd_geometry.cpp:303-772).
#pragma once
#include <string>
#include <vector>

struct EgsPhantom {
    std::vector<std::string> media;     // names in the header
    std::vector<double> xb, yb, zb;     // plane positions
    std::vector<int> medium;            // 0-based index into media, x fastest
    std::vector<double> density;        // g/cm^3, x fastest
};

struct CtRampEntry {
    std::string name;
    double rho_min, rho_max, rho_def;
};

// Plain or gzip-compressed .egsphant; throws std::runtime_error.
EgsPhantom read_egsphant(const std::string &path);

// "name rho_min rho_max rho_default" per line.
std::vector<CtRampEntry> read_ct_ramp(const std::string &path);
