// gpu.conf reader with the reference's semantics (options.cu:12-53): lines
// are trimmed of " \t\n\r\f\v"; a line that starts with '#' or has no '='
// is skipped; key and value are split at the first '=' and trimmed; unknown
// keys are ignored; a missing file throws std::runtime_error("Can't open
// <name>").  Integer keys go through std::stoi (which throws on a malformed
// value).  nyz and nxyz are derived at the end.
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>

#include "options.hpp"

namespace {

const char* kSpace = " \t\n\r\f\v";

std::string strip(const std::string& s)
{
    const auto b = s.find_first_not_of(kSpace);
    if (b == std::string::npos) return std::string();
    const auto e = s.find_last_not_of(kSpace);
    return s.substr(b, e - b + 1);
}

bool split_line(const std::string& raw, std::string& key, std::string& value)
{
    const std::string line = strip(raw);
    if (line.empty() || line[0] == '#') return false;
    const auto eq = line.find('=');
    if (eq == std::string::npos) return false;
    key = strip(line.substr(0, eq));
    value = strip(line.substr(eq + 1));
    return true;
}

}  // namespace

void load_options(Options& opts, const std::string& fname)
{
    std::ifstream in(fname.c_str());
    if (!in) throw std::runtime_error("Can't open " + fname);
    std::string raw, key, value;
    while (std::getline(in, raw)) {
        if (!split_line(raw, key, value)) continue;
        if (key == "nx") opts.nx = std::stoi(value);
        else if (key == "ny") opts.ny = std::stoi(value);
        else if (key == "nz") opts.nz = std::stoi(value);
        else if (key == "npx") opts.npx = std::stoi(value);
        else if (key == "vmin") opts.vmin = std::stod(value);
        else if (key == "vmax") opts.vmax = std::stod(value);
        else if (key == "R") opts.R = std::stod(value);
        else if (key == "method") opts.method = (value == "korobov" || value == "1") ? 1 : 0;
        else if (key == "tables") opts.tables = (value == "fp64" || value == "0") ? 0 : 1;
        else if (key == "correction") opts.correction = std::stoi(value);
        else if (key == "korobov_M") opts.korobov_M = (unsigned)std::stoul(value);
        else if (key == "seed") opts.seed = std::stoull(value);
        else if (key == "s0") opts.s0 = std::stod(value);
        else if (key == "steps") opts.steps = std::stoi(value);
        else if (key == "courant") opts.courant = std::stod(value);
    }
    opts.nyz = opts.ny * opts.nz;
    opts.nxyz = opts.nx * opts.nyz;
}

void print_options(const Options& opts)
{
    std::cout << "N_X: \t" << opts.nx << "\n"
              << "N_Y: \t" << opts.ny << "\n"
              << "N_Z: \t" << opts.nz << "\n"
              << "N_YZ: \t" << opts.nyz << "\n"
              << "N_XYZ: \t" << opts.nxyz << "\n"
              << "N_PX: \t" << opts.npx << std::endl;
}
