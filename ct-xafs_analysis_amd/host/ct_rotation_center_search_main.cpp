// ct_rotation_center_search -- command-line drop-in for the reference's
// RotationCenterSearch program (RotationCenterSearch/main.cpp,
// rotationCenterSearch_ocl.cpp:419-551): .his frames of one energy (dark, I0,
// It) -> mu-t -> the sinogram of one layer -> every candidate rotation-centre
// shift reconstructed by OS-EM and scored by its focus index -> the
// candidates' images, result.dat and the estimated shift.
//
//   ct_rotation_center_search <input.inp> [-ip dir] [-op dir] [-ob base] [-re E] [-ly layer]
//                             [-rss start] [-rsn count] [-rsst step]
//   ct_rotation_center_search --print-params <input.inp> [flags]
#include "ctr_host.hpp"
#include "ctr.h"

#include <dirent.h>
#include <hip/hip_runtime_api.h>
#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

struct SearchParams {
    ctrhost::Params ct;
    std::string input_dir, output_dir, output_base;
    int start_angle = -1, end_angle = -1;
    int image_x = 2048, image_y = 2048;
    int target_energy = -1;   // "#Target energy number for rotation center search" / -re
    int scan_n = 30;          // "#Number of scan for I0 and dark image"
    int layer = 0;            // "#Target layer number for rotation center search" / -ly (1-based)
    float shift_start = 0.0f; // "#Start shift for rotation center search" / -rss
    int shift_n = 0;          // "#Number of shift step for rotation center search" / -rsn
    float shift_step = NAN;   // "#Shift step for rotation center search" / -rsst
};

template <typename T>
bool next_value(std::istream &in, T &v)
{
    while (in) {
        const std::string line = ctrhost::getline_cr(in);
        if (line.find_first_not_of(" \t") == std::string::npos) continue;
        std::istringstream iss(line);
        return (bool)(iss >> v);
    }
    return false;
}

bool next_pair(std::istream &in, int &a, int &b)
{
    while (in) {
        const std::string line = ctrhost::getline_cr(in);
        if (line.find_first_not_of(" \t") == std::string::npos) continue;
        std::istringstream iss(line);
        if (!(iss >> a)) return false;
        iss.ignore();
        return (bool)(iss >> b);
    }
    return false;
}

// The keys of share source/input_parameter_class.cpp:318-470 and
// input_parameter_reslice_class.cpp:150-240 this program uses, plus the
// CT_reconstruction keys of the same file (RotationCenterSearch/main.cpp:13-49
// defaults: correctionMode 1).
bool parse(const std::string &path, SearchParams &p, std::string *err)
{
    p.ct.correctionMode = 1;
    p.ct.CSlambda = 1.0e-8f;
    if (!ctrhost::parse_inp(path, p.ct, err)) return false;
    std::ifstream in(path);
    bool first = true;
    while (in) {
        std::string line = ctrhost::getline_cr(in);
        if (first) {
            if (line.compare(0, 3, "\xEF\xBB\xBF") == 0) line.erase(0, 3);
            first = false;
        }
        bool ok = true;
        if (line == "#Input directory path") p.input_dir = ctrhost::getline_cr(in);
        else if (line == "#Output directory path") p.output_dir = ctrhost::getline_cr(in);
        else if (line == "#Output file base name") p.output_base = ctrhost::getline_cr(in);
        else if (line == "#Angle number range") ok = next_pair(in, p.start_angle, p.end_angle);
        else if (line == "#Image pixel size") ok = next_pair(in, p.image_x, p.image_y);
        else if (line == "#Target energy number for rotation center search") ok = next_value(in, p.target_energy);
        else if (line == "#Number of scan for I0 and dark image") ok = next_value(in, p.scan_n);
        else if (line == "#Target layer number for rotation center search") ok = next_value(in, p.layer);
        else if (line == "#Start shift for rotation center search") ok = next_value(in, p.shift_start);
        else if (line == "#Number of shift step for rotation center search") ok = next_value(in, p.shift_n);
        else if (line == "#Shift step for rotation center search") ok = next_value(in, p.shift_step);
        if (!ok) {
            if (err) *err = "bad value for key " + line + " in " + path;
            return false;
        }
    }
    return true;
}

std::string flag_value(int argc, char **argv, int from, const std::string &flag)
{
    for (int i = from; i < argc - 1; ++i)
        if (flag == argv[i]) return argv[i + 1];
    return std::string();
}

std::string tag4(const std::string &pre, int n, const std::string &post)
{
    char b[32];
    std::snprintf(b, sizeof b, "%04d", n);
    return pre + b + post;
}

int die(const std::string &msg)
{
    std::cerr << "ERROR: " << msg << std::endl;
    return 1;
}

struct Dev {
    void *p = nullptr;
    ~Dev()
    {
        if (p) (void)hipFree(p);
    }
    float *f() { return static_cast<float *>(p); }
};

}  // namespace

int main(int argc, char **argv)
{
    const bool print_only = argc >= 3 && std::string(argv[1]) == "--print-params";
    const int fa = print_only ? 2 : 1;
    if (argc <= fa) {
        std::cerr << "usage: " << argv[0]
                  << " <input.inp> [-ip dir] [-op dir] [-ob base] [-re E] [-ly layer] [-rss start] [-rsn count]"
                     " [-rsst step]\n";
        return 2;
    }
    SearchParams p;
    std::string err, v;
    if (!parse(argv[fa], p, &err)) return die(err);
    // RotationCenterSearch/main.cpp:84-211.  -rsn sets the number of shifts
    // (the reference passes its value to the start-shift setter).
    try {
        if (!(v = flag_value(argc, argv, fa + 1, "-ip")).empty()) p.input_dir = v;
        if (!(v = flag_value(argc, argv, fa + 1, "-op")).empty()) p.output_dir = v;
        if (!(v = flag_value(argc, argv, fa + 1, "-ob")).empty()) p.output_base = v;
        if (!(v = flag_value(argc, argv, fa + 1, "-re")).empty()) p.target_energy = std::stoi(v);
        if (!(v = flag_value(argc, argv, fa + 1, "-ly")).empty()) p.layer = std::stoi(v);
        if (!(v = flag_value(argc, argv, fa + 1, "-rss")).empty()) p.shift_start = std::stof(v);
        if (!(v = flag_value(argc, argv, fa + 1, "-rsn")).empty()) p.shift_n = std::stoi(v);
        if (!(v = flag_value(argc, argv, fa + 1, "-rsst")).empty()) p.shift_step = std::stof(v);
    } catch (const std::exception &) {
        return die("bad flag value " + v);
    }
    if (print_only) {
        std::cout << "{\"input_dir\": \"" << p.input_dir << "\", \"output_dir\": \"" << p.output_dir
                  << "\", \"output_base\": \"" << p.output_base << "\", \"angles\": [" << p.start_angle << ", "
                  << p.end_angle << "], \"image\": [" << p.image_x << ", " << p.image_y << "], \"target_energy\": "
                  << p.target_energy << ", \"scan_n\": " << p.scan_n << ", \"layer\": " << p.layer
                  << ", \"shift_start\": " << p.shift_start << ", \"shift_n\": " << p.shift_n << ", \"shift_step\": "
                  << (std::isnan(p.shift_step) ? std::string("null") : std::to_string(p.shift_step))
                  << ", \"correctionMode\": " << p.ct.correctionMode << ", \"g_ss\": " << p.ct.ss << "}" << std::endl;
        return 0;
    }
    // checks the reference makes through its prompts, as errors
    if (p.input_dir.empty()) return die("no input directory (#Input directory path, -ip)");
    if (p.output_dir.empty()) return die("no output directory (#Output directory path, -op)");
    if (p.target_energy < 0) return die("no target energy (#Target energy number for rotation center search, -re)");
    if (p.start_angle < 1 || p.end_angle < p.start_angle) return die("no angle number range (#Angle number range)");
    if (p.layer < 1 || p.layer > p.image_y) return die("the target layer must be 1.." + std::to_string(p.image_y));
    if (p.shift_n < 1 || std::isnan(p.shift_step)) return die("need a shift count >= 1 and a shift step");
    if (p.scan_n < 1) return die("the dark / I0 scan count must be >= 1");
    const int A = p.end_angle - p.start_angle + 1, W = p.image_x, H = p.image_y;
    p.ct.pa = A;   // main.cpp:218
    if (p.ct.px != W) return die("#カメラピクセル数 " + std::to_string(p.ct.px) + " differs from the camera width " +
                                 std::to_string(W));
    if (p.ct.ss < 1 || A % p.ct.ss != 0)
        return die("the projection count " + std::to_string(A) + " is not a multiple of the subset count " +
                   std::to_string(p.ct.ss));
    if (p.ct.ox != p.ct.nx || p.ct.oy != p.ct.ny)
        return die("output size must equal the reconstruction size (ROI extension is not provided)");
    // OSEM uses the first A angles of the log; setMask the angles of the range ends
    // (rotationCenterSearch_ocl.cpp:356-357)
    std::vector<float> ang;
    if (!ctrhost::read_log(p.ct.f3, std::max(A, p.end_angle), ang, &err)) return die(err);
    const float ang_ini = ang[p.start_angle - 1], ang_fin = ang[p.end_angle - 1];
    ang.resize(A);

    // main.cpp:114-135: <dir>/<base>dark.his names the data set
    DIR *d = ::opendir(p.input_dir.c_str());
    if (!d) return die("directory not found: " + p.input_dir);
    std::vector<std::string> names;
    for (dirent *e = ::readdir(d); e; e = ::readdir(d)) names.push_back(e->d_name);
    ::closedir(d);
    std::sort(names.begin(), names.end());
    std::string base;
    for (const std::string &n : names)
        if (n.find("dark.his") != std::string::npos && n.size() >= 8) {
            base = p.input_dir + "/" + n.substr(0, n.size() - 8);
            std::cout << "his file found: " << n << "\n";
            break;
        }
    if (base.empty()) return die("no his file found in " + p.input_dir);
    const std::string etag = ctrhost::energy_tag(p.target_energy);

    // his_data_input (rotationCenterSearch_ocl.cpp:16-103), for the two camera
    // rows the layer's sinogram samples (row layer-1 and the zero-weight row after it)
    const int row0 = p.layer - 1, nrows = std::min(2, H - row0);
    std::vector<unsigned short> dark_raw, i0_raw, it_raw;
    if (!ctrhost::read_his_rows(base + "dark.his", 1, p.scan_n, W, H, row0, nrows, dark_raw, &err) ||
        !ctrhost::read_his_rows(base + etag + "_I0.his", 1, p.scan_n, W, H, row0, nrows, i0_raw, &err) ||
        !ctrhost::read_his_rows(base + etag + ".his", p.start_angle, p.end_angle, W, H, row0, nrows, it_raw, &err))
        return die(err);

    const int dev = 0;
    ctr_ctx *ctx = nullptr;
    if (ctr_create(dev, &ctx) != CTR_OK) return die("cannot create a context on device 1");
    struct Guard {
        ctr_ctx *c;
        ~Guard() { ctr_destroy(c); }
    } guard{ctx};
    const size_t m = (size_t)nrows * W, nimg = (size_t)p.ct.nx * p.ct.ny;
    const int N = p.shift_n;
    Dev d_dark_raw, d_i0_raw, d_it_raw, d_dark, d_i0, d_mt, d_prj, d_rec, d_f;
    if (hipMalloc(&d_dark_raw.p, 2 * dark_raw.size()) != hipSuccess || hipMalloc(&d_i0_raw.p, 2 * i0_raw.size()) != hipSuccess ||
        hipMalloc(&d_it_raw.p, 2 * it_raw.size()) != hipSuccess || hipMalloc(&d_dark.p, 4 * m) != hipSuccess ||
        hipMalloc(&d_i0.p, 4 * m) != hipSuccess || hipMalloc(&d_mt.p, 4 * m * A) != hipSuccess ||
        hipMalloc(&d_prj.p, 4 * (size_t)A * W) != hipSuccess || hipMalloc(&d_rec.p, 4 * nimg * N) != hipSuccess ||
        hipMalloc(&d_f.p, 4 * (size_t)N) != hipSuccess)
        return die("out of device memory");
    if (hipMemcpy(d_dark_raw.p, dark_raw.data(), 2 * dark_raw.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_i0_raw.p, i0_raw.data(), 2 * i0_raw.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_it_raw.p, it_raw.data(), 2 * it_raw.size(), hipMemcpyHostToDevice) != hipSuccess)
        return die("host-to-device copy failed");
    auto ok = [&](ctr_status s) { return s == CTR_OK; };
    // merge_rawhisdata of the scans, mt_conversion of every angle, reslice of the layer
    if (!ok(ctr_merge_frames(ctx, static_cast<unsigned short *>(d_dark_raw.p), p.scan_n, W, nrows, (long long)m, d_dark.f())) ||
        !ok(ctr_merge_frames(ctx, static_cast<unsigned short *>(d_i0_raw.p), p.scan_n, W, nrows, (long long)m, d_i0.f())) ||
        !ok(ctr_mt_conversion(ctx, d_dark.f(), d_i0.f(), static_cast<unsigned short *>(d_it_raw.p), A, W, nrows,
                              (long long)m, d_mt.f())) ||
        !ok(ctr_reslice(ctx, d_mt.f(), A, nrows, W, nullptr, nullptr, 0.0f, 0, 0, 1, d_prj.f())))
        return die(ctr_last_error(ctx));
    ctr_geometry g{};
    g.nx = p.ct.nx; g.ny = p.ct.ny; g.ox = p.ct.ox; g.oy = p.ct.oy;
    g.px = W; g.pa = A; g.ss = p.ct.ss; g.zp = p.ct.zp;
    g.mode = 6; g.correction_mode = p.ct.correctionMode;
    g.wt1 = (float)p.ct.wt1; g.wt2 = (float)p.ct.wt2; g.cs_lambda = p.ct.CSlambda;
    if (!ok(ctr_set_geometry(ctx, &g, ang.data(), nullptr))) return die(ctr_last_error(ctx));
    std::cout << "Reconstructing CT images..." << std::endl;
    // reconstRotCntShift_thread: g_ini start, OSEM_execution(it = 1, correction)
    if (!ok(ctr_rotation_center_search(ctx, d_prj.f(), N, p.shift_start, p.shift_step, ang_ini, ang_fin, 1,
                                       (float)p.ct.ini, 1, d_rec.f(), d_f.f())) ||
        !ok(ctr_synchronize(ctx)))
        return die(ctr_last_error(ctx));
    std::vector<float> rec(nimg * N), findex(N);
    if (hipMemcpy(rec.data(), d_rec.p, 4 * rec.size(), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(findex.data(), d_f.p, 4 * (size_t)N, hipMemcpyDeviceToHost) != hipSuccess)
        return die("device-to-host copy failed");

    // output (rotationCenterSearch_ocl.cpp:511-549)
    if (::mkdir(p.output_dir.c_str(), 0755) != 0 && errno != EEXIST) return die("failed to create " + p.output_dir);
    float fmin = 3.402823466e+38f;
    int imin = 0;
    for (int i = 0; i < N; ++i) {
        if (!ctrhost::write_raw(p.output_dir, tag4(p.output_base, i + 1, ".raw"), rec.data() + nimg * i, nimg, &err))
            return die(err);
        if (fmin > findex[i]) {
            fmin = findex[i];
            imin = i;
        }
    }
    std::ofstream ofs(p.output_dir + "/result.dat", std::ios::out | std::ios::trunc);
    if (!ofs) return die("cannot write " + p.output_dir + "/result.dat");
    ofs << "Rotation center shift\tFocus Index\tfilename" << std::endl;
    std::cout << "Result" << std::endl;
    for (int i = 0; i < N; ++i) {
        const float shift = p.shift_start + p.shift_step * (float)i;
        std::cout << shift << " " << findex[i] << "\t(" << tag4(p.output_base, i + 1, ".raw)") << std::endl;
        ofs << shift << "\t" << findex[i] << "\t" << tag4(p.output_base, i + 1, ".raw") << std::endl;
    }
    const float best = p.shift_start + (float)imin * p.shift_step;
    std::cout << "\nEstimated rotation center shift is " << best << "(" << tag4(p.output_base, imin + 1, ".raw)")
              << std::endl;
    return 0;
}
