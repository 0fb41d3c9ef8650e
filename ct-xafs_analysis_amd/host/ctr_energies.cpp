// ctr_energies.cpp -- the multi-energy driver of the CT reconstruction host
// (Reslice_CTreconst/): mu-t projections per energy -> sinograms resliced on
// the device -> reconstruction on the device -> image files, driving
// libctr_hip.so.  The sinograms never leave the device.
#include "ctr_host.hpp"
#include "ctr.h"

#include <dirent.h>
#include <hip/hip_runtime_api.h>
#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cstdio>
#include <fstream>
#include <future>
#include <iomanip>
#include <mutex>
#include <ostream>
#include <sstream>
#include <thread>

namespace ctrhost {

namespace {

void set_err(std::string *err, const std::string &msg)
{
    if (err) *err = msg;
}

// "a-b" (any one separator character) on the next non-empty line, read as
// the reference does with `in >> a; in.ignore() >> b`.
bool read_pair(std::istream &in, int &a, int &b)
{
    while (in) {
        const std::string line = getline_cr(in);
        if (line.find_first_not_of(" \t") == std::string::npos) continue;
        std::istringstream iss(line);
        if (!(iss >> a)) return false;
        iss.ignore();
        return (bool)(iss >> b);
    }
    return false;
}

std::vector<std::string> split_names(const std::string &s)
{
    std::vector<std::string> out;
    std::istringstream iss(s);
    std::string item;
    while (std::getline(iss, item, ',')) out.push_back(item);
    return out;
}

std::string tag4(int n)
{
    char b[32];
    std::snprintf(b, sizeof b, "%04d", n);
    return b;
}

// Pinned host buffer (hipHostMalloc) for asynchronous copies.
struct Pinned {
    float *p = nullptr;
    explicit Pinned(size_t n)
    {
        if (hipHostMalloc(reinterpret_cast<void **>(&p), sizeof(float) * std::max<size_t>(n, 1), hipHostMallocDefault) !=
            hipSuccess)
            p = nullptr;
    }
    ~Pinned()
    {
        if (p) (void)hipHostFree(p);
    }
    Pinned(const Pinned &) = delete;
    Pinned &operator=(const Pinned &) = delete;
};

struct DevBuf {
    float *p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};

std::mutex g_log_mu;

}  // namespace

bool parse_energy_inp(const std::string &path, EnergyParams &p, std::string *err)
{
    p.ct.CSlambda = 0.001f;   // Reslice_CTreconst/main.cpp:48
    if (!parse_inp(path, p.ct, err)) return false;
    std::ifstream in(path);
    if (!in) {
        set_err(err, "input parameter file not found: " + path);
        return false;
    }
    bool first = true;
    while (in) {
        std::string line = getline_cr(in);
        if (first) {
            if (line.compare(0, 3, "\xEF\xBB\xBF") == 0) line.erase(0, 3);
            first = false;
        }
        bool ok = true;
        if (line == "#Input directory path") p.input_dir = getline_cr(in);
        else if (line == "#Output directory path") p.output_dir = getline_cr(in);
        else if (line == "#Output file base name") p.output_base = getline_cr(in);
        else if (line == "#Angle number range") ok = read_pair(in, p.start_angle, p.end_angle);
        else if (line == "#Energy number range") ok = read_pair(in, p.start_energy, p.end_energy);
        else if (line == "#Image pixel size") ok = read_pair(in, p.image_x, p.image_y);
        else if (line == "#Processing parameter name") p.names = split_names(getline_cr(in));
        if (!ok) {
            set_err(err, "bad value for key " + line + " in " + path);
            return false;
        }
    }
    return true;
}

std::string energy_tag(int energy)
{
    char b[32];
    std::snprintf(b, sizeof b, "%03d", energy);
    return b;
}

bool find_projection_base(const std::string &dir, std::string &base, std::string *err)
{
    DIR *d = ::opendir(dir.c_str());
    if (!d) {
        set_err(err, "directory not found: " + dir);
        return false;
    }
    std::vector<std::string> names;
    for (dirent *e = ::readdir(d); e; e = ::readdir(d)) names.push_back(e->d_name);
    ::closedir(d);
    std::sort(names.begin(), names.end());
    for (const std::string &n : names)
        if (n.find("0001.raw") != std::string::npos && n.size() >= 8) {
            base = n.substr(0, n.size() - 8);
            return true;
        }
    set_err(err, "no raw file found in " + dir);
    return false;
}

int run_energies(const EnergyParams &P, std::ostream &log)
{
    const Params &p = P.ct;
    auto fail = [&](const std::string &msg) {
        std::lock_guard<std::mutex> lk(g_log_mu);
        log << "ERROR: " << msg << std::endl;
        return 1;
    };
    std::string err;
    // energies, or parameter names (Reslice_CTreconst_ocl.cpp:282-298)
    std::vector<std::string> subs;
    if (P.start_energy > 0 && P.end_energy > 0) {
        for (int e = P.start_energy; e <= P.end_energy; ++e) subs.push_back(energy_tag(e));
    } else {
        subs = P.names;
    }
    if (subs.empty()) return fail("no energy range (#Energy number range, -enr) or parameter names (-fpn)");
    if (P.input_dir.empty()) return fail("no input directory (#Input directory path, -ip)");
    if (P.output_dir.empty()) return fail("no output directory (#Output directory path, -op)");
    if (P.start_angle < 0 || P.end_angle < P.start_angle) return fail("no angle number range (#Angle number range, -ar)");
    const int A = P.end_angle - P.start_angle + 1;
    if (A != p.pa)
        return fail("the angle number range holds " + std::to_string(A) + " projections but #投影数 is " +
                    std::to_string(p.pa));
    if (P.image_x != p.px)
        return fail("the camera width " + std::to_string(P.image_x) + " differs from #カメラピクセル数 " +
                    std::to_string(p.px));
    // Reslice_CTreconst_ocl.cpp:98-125: only MART, ML-EM, OS-EM and FBP run
    if (!(p.mode == 2 || p.mode == 5 || p.mode == 6 || p.mode == 7))
        return fail("reconstruction method " + std::to_string(p.mode) +
                    " is not run by the multi-energy driver (it runs 2, 5, 6 and 7)");
    if (!validate(p, &err)) return fail(err);
    // layers st .. st + num - 2 (Reslice_CTreconst_ocl.cpp:94); layer N is
    // camera row N, counted from 0 (OSEM_thread with BATCHEXE, GPU.cpp:414-416)
    const int first = p.st, last = p.st + p.num - 2;
    if (last < first) return fail("#レイヤー数 must be at least 2: the driver reconstructs layers st .. st + num - 2");
    if (first < 0 || last > P.image_y - 1)
        return fail("layers " + std::to_string(first) + ".." + std::to_string(last) +
                    " use camera rows of the same numbers, which must lie in 0.." + std::to_string(P.image_y - 1));
    std::vector<int> devs;
    if (!parse_device_list(p.devList, devs, &err)) return fail(err);
    std::vector<float> ang;
    if (!read_log(p.f3, p.pa, ang, &err)) return fail(err);
    std::string base;
    if (!find_projection_base(P.input_dir + "/" + subs[0], base, &err)) return fail(err);
    if (::mkdir(P.output_dir.c_str(), 0755) != 0 && errno != EEXIST)
        return fail("failed to create directory " + P.output_dir);

    const bool fbp = p.mode == 7, cs = p.CSitBool && !fbp;
    const int ss = fbp ? 1 : method_subsets(p);
    const int nL = last - first + 1;
    const int W = P.image_x, R = std::min(nL + 1, P.image_y - first);   // camera rows first .. first + R - 1
    const size_t mt_n = (size_t)A * R * W, prj_layer = (size_t)A * W, img_n = (size_t)p.nx * p.ny;
    const auto t0 = std::chrono::steady_clock::now();
    {
        static const char *names[] = {"", "", "MART", "", "", "ML-EM", "OS-EM", "FBP"};
        std::lock_guard<std::mutex> lk(g_log_mu);
        log << "Processing by " << names[p.mode] << " method" << (cs ? " with FISTA compressed sensing" : "") << ", "
            << subs.size() << " data set(s), layers " << first << ".." << last << "\n\n";
    }
    ctr_geometry g{};
    g.nx = p.nx; g.ny = p.ny; g.ox = p.ox; g.oy = p.oy;
    g.px = p.px; g.pa = p.pa; g.ss = ss; g.zp = p.zp;
    g.mode = p.mode; g.correction_mode = p.correctionMode;
    g.wt1 = (float)p.wt1; g.wt2 = (float)p.wt2; g.cs_lambda = p.CSlambda;

    std::vector<std::string> errors(devs.size());
    auto work = [&](size_t w) {
        std::string &e = errors[w];
        const int dev = devs[w] - 1;
        if (hipSetDevice(dev) != hipSuccess) { e = "device not found"; return; }
        int cus = 0;
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        const int dN = std::max(1, std::min(p.numParallel, cus));   // layers per OSEM_thread batch
        ctr_ctx *ctx = nullptr;
        if (ctr_create(dev, &ctx) != CTR_OK) { e = "cannot create a context"; return; }
        struct CtxGuard {
            ctr_ctx *c;
            ~CtxGuard() { ctr_destroy(c); }
        } guard{ctx};
        if (ctr_set_geometry(ctx, &g, ang.data(), nullptr) != CTR_OK) { e = ctr_last_error(ctx); return; }
        void *sv = nullptr;
        (void)ctr_get_stream(ctx, &sv);
        hipStream_t st = static_cast<hipStream_t>(sv);
        DevBuf d_mt, d_prj, d_img;
        if (hipMalloc(&d_mt.p, sizeof(float) * mt_n) != hipSuccess ||
            hipMalloc(&d_prj.p, sizeof(float) * prj_layer * nL) != hipSuccess ||
            hipMalloc(&d_img.p, sizeof(float) * img_n * nL) != hipSuccess) {
            e = "out of device memory";
            return;
        }
        Pinned h_mt0(mt_n), h_mt1(mt_n), h_img0(img_n * nL), h_img1(img_n * nL), h_init(fbp ? 0 : img_n * nL);
        if (!h_mt0.p || !h_mt1.p || !h_img0.p || !h_img1.p || !h_init.p) { e = "out of pinned host memory"; return; }
        float *h_mt[2] = {h_mt0.p, h_mt1.p}, *h_img[2] = {h_img0.p, h_img1.p};
        if (!fbp) {   // first_image of every layer (Reslice_CTreconst_ocl.cpp:91-92)
            if (!first_image(p, h_init.p, img_n, &e)) return;
            for (int l = 1; l < nL; ++l) std::copy(h_init.p, h_init.p + img_n, h_init.p + img_n * l);
        }
        std::vector<size_t> mine;
        for (size_t k = w; k < subs.size(); k += devs.size()) mine.push_back(k);
        if (mine.empty()) return;

        // reads projection rows first .. first + R - 1 of every angle
        auto read_set = [&](size_t k, float *dst) -> std::string {
            const std::string dir = P.input_dir + "/" + subs[k] + "/";
            const int nthreads = 8;
            std::vector<std::future<std::string>> parts;
            for (int t = 0; t < nthreads; ++t)
                parts.push_back(std::async(std::launch::async, [&, t]() -> std::string {
                    for (int a = t; a < A; a += nthreads) {
                        const std::string f = dir + base + tag4(P.start_angle + a) + ".raw";
                        std::ifstream in(f, std::ios::binary);
                        if (!in) return "file open error [" + f + "]";
                        in.seekg((std::streamoff)sizeof(float) * first * W);
                        const std::streamsize bytes = (std::streamsize)(sizeof(float) * (size_t)R * W);
                        in.read(reinterpret_cast<char *>(dst + (size_t)a * R * W), bytes);
                        if (in.gcount() != bytes) return "file " + f + " is shorter than " +
                                                          std::to_string(P.image_x) + " x " +
                                                          std::to_string(P.image_y) + " floats";
                    }
                    return std::string();
                }));
            std::string r;
            for (auto &f : parts) {
                const std::string s = f.get();
                if (r.empty()) r = s;
            }
            return r;
        };
        // output_thread (GPU.cpp:99-123) into <output_dir>/<sub>
        auto write_set = [&](size_t k, const float *src, int out_first, int out_last,
                             std::vector<std::pair<int, int>> ranges) -> std::string {
            const std::string dir = P.output_dir + "/" + subs[k];
            if (::mkdir(dir.c_str(), 0755) != 0 && errno != EEXIST) return "failed to create directory " + dir;
            std::string we;
            std::vector<float> roi;
            for (const auto &rg : ranges)
                for (int N = rg.first; N <= rg.second; ++N)
                    if (!write_raw(dir, P.output_base + tag4(N) + ".raw", output_layer(p, src + img_n * (N - first), roi),
                                   (size_t)p.ox * p.oy, &we))
                        return we;
            const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            std::lock_guard<std::mutex> lk(g_log_mu);
            log << "[" << subs[k] << "] layers " << out_first << ".." << out_last << " written, " << std::fixed
                << std::setprecision(1) << secs << " s elapsed\n";
            log.flush();
            return std::string();
        };

        std::future<std::string> reading = std::async(std::launch::async, read_set, mine[0], h_mt[0]);
        std::future<std::string> writing;
        for (size_t i = 0; i < mine.size(); ++i) {
            const size_t k = mine[i];
            float *hm = h_mt[i & 1], *hi = h_img[i & 1];
            e = reading.get();
            if (e.empty() && i + 1 < mine.size())
                reading = std::async(std::launch::async, read_set, mine[i + 1], h_mt[(i + 1) & 1]);   // read ahead
            if (!e.empty()) break;
            ctr_status s = CTR_OK;
            if (hipMemcpyAsync(d_mt.p, hm, sizeof(float) * mt_n, hipMemcpyHostToDevice, st) != hipSuccess ||
                (!fbp && hipMemcpyAsync(d_img.p, h_init.p, sizeof(float) * img_n * nL, hipMemcpyHostToDevice, st) !=
                             hipSuccess)) {
                e = "host-to-device copy failed";
                break;
            }
            // reslice_mtImg (no shifts, base-up 0, no correction): layer l = camera row first + l
            s = ctr_reslice(ctx, d_mt.p, A, R, W, nullptr, nullptr, 0.0f, 0, 0, nL, d_prj.p);
            std::vector<std::pair<int, int>> ranges;
            if (cs) {
                // OSEM_thread batches of dN layers and its output rule (GPU.cpp:453-457)
                for (int N = first; N <= last && s == CTR_OK; N += dN) {
                    const int end = std::min(N + dN - 1, last);
                    s = ctr_fista_osem_execution(ctx, d_img.p + img_n * (N - first), d_prj.p + prj_layer * (N - first),
                                                 end - N + 1, p.it, 1);
                    const int lo = N + (N > p.st ? p.CSoverlap : 0);
                    const int hi = (end < p.st + p.num - 1) ? end - p.CSoverlap : end;
                    if (lo <= hi) ranges.push_back({lo, hi});
                }
            } else {
                const int chunk = 256;   // layers are independent: any grouping gives the same images
                for (int N = first; N <= last && s == CTR_OK; N += chunk) {
                    const int cnt = std::min(chunk, last - N + 1);
                    float *img = d_img.p + img_n * (N - first), *prj = d_prj.p + prj_layer * (N - first);
                    s = fbp ? ctr_fbp_execution(ctx, img, prj, cnt, 1) : ctr_osem_execution(ctx, img, prj, cnt, p.it, 1);
                }
                ranges.push_back({first, last});
            }
            if (s == CTR_OK) s = ctr_synchronize(ctx);
            if (s != CTR_OK) {
                e = ctr_last_error(ctx);
                break;
            }
            if (writing.valid()) {
                e = writing.get();
                if (!e.empty()) break;
            }
            if (hipMemcpy(hi, d_img.p, sizeof(float) * img_n * nL, hipMemcpyDeviceToHost) != hipSuccess) {
                e = "device-to-host copy failed";
                break;
            }
            writing = std::async(std::launch::async, write_set, k, hi, first, last, ranges);
        }
        if (reading.valid()) reading.wait();
        if (writing.valid()) {
            const std::string we = writing.get();
            if (e.empty()) e = we;
        }
    };
    std::vector<std::thread> threads;
    for (size_t w = 0; w < devs.size(); ++w) threads.emplace_back(work, w);
    for (auto &t : threads) t.join();
    int rc = 0;
    for (size_t w = 0; w < devs.size(); ++w)
        if (!errors[w].empty()) rc = fail("device " + std::to_string(devs[w]) + ": " + errors[w]);
    return rc;
}

}  // namespace ctrhost
