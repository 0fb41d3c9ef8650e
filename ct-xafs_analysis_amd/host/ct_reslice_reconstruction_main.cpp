// ct_reslice_reconstruction -- command-line drop-in for the reference's
// Reslice_CTreconst program (Reslice_CTreconst/main.cpp): mu-t projections of
// every energy (or fitting-parameter map) are resliced into sinograms and
// reconstructed on the device, one output directory per energy.
//
//   ct_reslice_reconstruction <input.inp> [-ip dir] [-enr a-b] [-fpn a,b,..] [-op dir] [-ar a-b]
//   ct_reslice_reconstruction --print-params <input.inp> [flags]
//
// The flags override the parameter file as in the reference; a missing value
// is an error instead of an interactive prompt.
#include "ctr_host.hpp"

#include <cstdio>
#include <iostream>
#include <sstream>
#include <string>

namespace {

std::string json_str(const std::string &s)
{
    std::string o = "\"";
    for (unsigned char ch : s) {
        if (ch == '"' || ch == '\\') { o += '\\'; o += (char)ch; }
        else if (ch < 0x20) { char b[8]; std::snprintf(b, sizeof b, "\\u%04x", ch); o += b; }
        else o += (char)ch;
    }
    return o + "\"";
}

// output_flag (share source/CTXAFS_analysis_share.cpp:29-41): the argument
// after `flag`.
std::string flag_value(int argc, char **argv, int from, const std::string &flag)
{
    for (int i = from; i < argc - 1; ++i)
        if (flag == argv[i]) return argv[i + 1];
    return std::string();
}

bool parse_range(const std::string &s, int &a, int &b)
{
    std::istringstream iss(s);
    if (!(iss >> a)) return false;
    iss.ignore();
    return (bool)(iss >> b);
}

}  // namespace

int main(int argc, char **argv)
{
    const bool print_only = argc >= 3 && std::string(argv[1]) == "--print-params";
    const int file_arg = print_only ? 2 : 1;
    if (argc <= file_arg) {
        std::cerr << "usage: " << argv[0] << " <input.inp> [-ip dir] [-enr a-b] [-fpn a,b] [-op dir] [-ar a-b]\n       "
                  << argv[0] << " --print-params <input.inp> [flags]\n";
        return 2;
    }
    ctrhost::EnergyParams p;
    std::string err;
    if (!ctrhost::parse_energy_inp(argv[file_arg], p, &err)) {
        std::cerr << "ERROR: " << err << std::endl;
        return 1;
    }
    // Reslice_CTreconst/main.cpp:72-164
    std::string v;
    if (!(v = flag_value(argc, argv, file_arg + 1, "-ip")).empty()) p.input_dir = v;
    if (!(v = flag_value(argc, argv, file_arg + 1, "-enr")).empty() && !parse_range(v, p.start_energy, p.end_energy)) {
        std::cerr << "ERROR: bad energy range " << v << std::endl;
        return 1;
    }
    if (!(v = flag_value(argc, argv, file_arg + 1, "-fpn")).empty()) {
        p.names.clear();
        std::istringstream iss(v);
        std::string item;
        while (std::getline(iss, item, ',')) p.names.push_back(item);
    }
    if (!(v = flag_value(argc, argv, file_arg + 1, "-op")).empty()) p.output_dir = v;
    if (!(v = flag_value(argc, argv, file_arg + 1, "-ar")).empty() && !parse_range(v, p.start_angle, p.end_angle)) {
        std::cerr << "ERROR: bad angle range " << v << std::endl;
        return 1;
    }
    if (print_only) {
        std::cout << "{\"input_dir\": " << json_str(p.input_dir) << ", \"output_dir\": " << json_str(p.output_dir)
                  << ", \"output_base\": " << json_str(p.output_base) << ", \"angles\": [" << p.start_angle << ", "
                  << p.end_angle << "], \"energies\": [" << p.start_energy << ", " << p.end_energy << "], \"names\": [";
        for (size_t i = 0; i < p.names.size(); ++i) std::cout << (i ? ", " : "") << json_str(p.names[i]);
        std::cout << "], \"image\": [" << p.image_x << ", " << p.image_y << "], \"g_px\": " << p.ct.px
                  << ", \"g_pa\": " << p.ct.pa << ", \"g_mode\": " << p.ct.mode << ", \"g_st\": " << p.ct.st
                  << ", \"g_num\": " << p.ct.num << ", \"CSlambda\": " << p.ct.CSlambda << ", \"g_devList\": "
                  << json_str(p.ct.devList) << ", \"numParallel\": " << p.ct.numParallel << "}" << std::endl;
        return 0;
    }
    return ctrhost::run_energies(p, std::cout);
}
