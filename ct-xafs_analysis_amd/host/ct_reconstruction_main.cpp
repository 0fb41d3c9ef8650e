// ct_reconstruction -- command-line drop-in for the reference's
// CT_reconstruction program (CT_reconstruction/main.cpp): reads an input
// parameter file, checks the input directory, loads the angle log and runs
// the reconstruction on the listed devices through libctr_hip.so.
//
//   ct_reconstruction <input.inp>                 run
//   ct_reconstruction --print-params <input.inp>  print the parsed parameters as JSON
#include "ctr_host.hpp"

#include <sys/stat.h>

#include <cstdio>
#include <iostream>
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

void print_params(const ctrhost::Params &p)
{
    std::cout << "{\"g_d1\": " << json_str(p.d1) << ", \"g_f1\": " << json_str(p.f1) << ", \"g_d2\": " << json_str(p.d2)
              << ", \"g_f2\": " << json_str(p.f2) << ", \"g_f3\": " << json_str(p.f3) << ", \"g_f4\": " << json_str(p.f4)
              << ", \"g_d5\": " << json_str(p.d5) << ", \"g_fst\": " << p.fst << ", \"g_ini\": " << p.ini
              << ", \"g_px\": " << p.px << ", \"g_pa\": " << p.pa << ", \"g_nx\": " << p.nx << ", \"g_ny\": " << p.ny
              << ", \"g_ox\": " << p.ox << ", \"g_oy\": " << p.oy << ", \"g_mode\": " << p.mode << ", \"g_it\": " << p.it
              << ", \"g_num\": " << p.num << ", \"g_st\": " << p.st << ", \"g_wt1\": " << p.wt1 << ", \"g_wt2\": " << p.wt2
              << ", \"g_ss\": " << p.ss << ", \"g_zp\": " << p.zp << ", \"g_devList\": " << json_str(p.devList)
              << ", \"numParallel\": " << p.numParallel << ", \"correctionMode\": " << p.correctionMode
              << ", \"amp\": " << p.amp << ", \"baseupOrder\": " << p.baseupOrder
              << ", \"CSitBool\": " << (p.CSitBool ? "true" : "false") << ", \"CSlambda\": " << p.CSlambda
              << ", \"CSoverlap\": " << p.CSoverlap << ", \"method_subsets\": " << ctrhost::method_subsets(p) << "}"
              << std::endl;
}

}  // namespace

int main(int argc, char **argv)
{
    const bool print_only = argc == 3 && std::string(argv[1]) == "--print-params";
    if (!(argc == 2 || print_only)) {
        std::cerr << "usage: " << argv[0] << " <input.inp>\n       " << argv[0] << " --print-params <input.inp>\n";
        return 2;
    }
    ctrhost::Params p;
    std::string err;
    if (!ctrhost::parse_inp(argv[argc - 1], p, &err)) {
        std::cerr << "ERROR: " << err << std::endl;
        return 1;
    }
    if (print_only) {
        print_params(p);
        return 0;
    }
    // main.cpp:69-105: the first layer's sinogram must be there
    struct stat sb;
    if (::stat(p.d1.c_str(), &sb) != 0 || !S_ISDIR(sb.st_mode)) {
        std::cerr << "ERROR: directory not found: " << p.d1 << std::endl;
        return 1;
    }
    const std::string first = p.d1 + "/" + ctrhost::layer_name(p.f1, p.st);
    if (::stat(first.c_str(), &sb) != 0) {
        std::cerr << "ERROR: no raw file found: " << first << std::endl;
        return 1;
    }
    std::cout << "raw sinogram file found: " << first << "\n\n";
    return ctrhost::run(p, std::cout);
}
