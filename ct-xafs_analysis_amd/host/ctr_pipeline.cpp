// ctr_pipeline.cpp -- method dispatch and the per-device layer pipeline of the
// CT reconstruction host (GPU.cpp:23-97, 99-123 and the *_ocl / *_input_thread
// / *_thread functions of the reference), driving libctr_hip.so.
#include "ctr_host.hpp"
#include "ctr.h"

#include <hip/hip_runtime_api.h>
#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cstdio>
#include <future>
#include <iomanip>
#include <mutex>
#include <ostream>
#include <sstream>
#include <thread>

namespace ctrhost {

namespace {

enum class Method { OSEM, FISTA_OSEM, AART, FISTA, FBP, HYBRID };

// GPU.cpp:36-93: the method a mode runs.
Method method_of(const Params &p)
{
    switch (p.mode) {
    case 1:
    case 3: return p.CSitBool ? Method::FISTA : Method::AART;
    case 7: return Method::FBP;
    case 8: return Method::HYBRID;
    default: return p.CSitBool ? Method::FISTA_OSEM : Method::OSEM;   // 2, 4, 5, 6
    }
}

const char *method_name(int mode)
{
    static const char *names[] = {"", "AART", "MART", "ASIRT", "MSIRT", "ML-EM", "OS-EM", "FBP", "FBP+OS-EM hybrid"};
    return (mode >= 1 && mode <= 8) ? names[mode] : "?";
}

// One batch of consecutive layers [start, end].
struct Chunk {
    int start, end;
};

// Device buffers owned by one device worker.
struct DevMem {
    float *img = nullptr, *prj = nullptr, *l2 = nullptr;
    ~DevMem()
    {
        if (img) (void)hipFree(img);
        if (prj) (void)hipFree(prj);
        if (l2) (void)hipFree(l2);
    }
};

std::mutex g_log_mu;

struct HostBatch {
    Chunk c{};
    std::vector<float> imgs, prjs;
    std::string err;
    bool ok = false;
};

}  // namespace

int run(const Params &p, std::ostream &log)
{
    std::string err;
    auto fail = [&](const std::string &msg) {
        std::lock_guard<std::mutex> lk(g_log_mu);
        log << "ERROR: " << msg << std::endl;
        return 1;
    };
    if (!validate(p, &err)) return fail(err);
    std::vector<int> devs;
    if (!parse_device_list(p.devList, devs, &err)) return fail(err);
    std::vector<float> ang;
    if (!read_log(p.f3, p.pa, ang, &err)) return fail(err);
    if (::mkdir(p.d2.c_str(), 0755) != 0 && errno != EEXIST) return fail("failed to create directory " + p.d2);

    const Method m = method_of(p);
    const int ss = method_subsets(p);
    log << "Processing by " << method_name(p.mode) << " method"
        << ((p.CSitBool && m != Method::FBP) ? " with FISTA compressed sensing" : "") << "\n\n";
    const auto t0 = std::chrono::steady_clock::now();

    ctr_geometry g{};
    g.nx = p.nx; g.ny = p.ny; g.ox = p.ox; g.oy = p.oy;
    g.px = p.px; g.pa = p.pa; g.zp = p.zp;
    g.ss = (m == Method::FBP) ? 1 : ss;   // FBP has no subsets
    g.mode = p.mode;
    g.correction_mode = p.correctionMode;
    g.wt1 = (float)p.wt1; g.wt2 = (float)p.wt2;
    g.cs_lambda = p.CSlambda;

    // Layer batches in the reference's order: each device takes the next
    // dN layers; with the 3-D TV the batches overlap by 2 * CSoverlap layers
    // (GPU.cpp:634-648, 2251-2264).
    const int last = p.st + p.num - 1;
    const bool overlap = p.CSitBool && m != Method::FBP && m != Method::AART;
    const size_t img_n = (size_t)p.nx * p.ny, prj_n = (size_t)p.px * p.pa;

    // per-device set-up: context, batch size, L for FISTA
    struct Worker {
        int dev = 0, dN = 1;
        ctr_ctx *ctx = nullptr;
        DevMem mem;
        ~Worker()
        {
            if (ctx) ctr_destroy(ctx);
        }
    };
    std::vector<Worker> workers(devs.size());
    for (size_t w = 0; w < devs.size(); ++w) {
        Worker &W = workers[w];
        W.dev = devs[w] - 1;
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, W.dev) != hipSuccess)
            return fail("device " + std::to_string(devs[w]) + " not found");
        W.dN = std::max(1, std::min(p.numParallel, cus));
        if (overlap && W.dN <= 2 * p.CSoverlap && p.num > W.dN)
            return fail("layers per batch (" + std::to_string(W.dN) + ") must exceed twice the overlap");
        if (ctr_create(W.dev, &W.ctx) != CTR_OK) return fail("cannot create a context on device " + std::to_string(devs[w]));
        if (ctr_set_geometry(W.ctx, &g, ang.data(), nullptr) != CTR_OK) return fail(ctr_last_error(W.ctx));
        (void)hipSetDevice(W.dev);
        if (hipMalloc(&W.mem.img, sizeof(float) * img_n * W.dN) != hipSuccess ||
            hipMalloc(&W.mem.prj, sizeof(float) * prj_n * W.dN) != hipSuccess)
            return fail("out of device memory for " + std::to_string(W.dN) + " layers");
        if (m == Method::FISTA) {
            // powerIteration once per device (GPU.cpp:2244-2251)
            if (hipMalloc(&W.mem.l2, sizeof(float) * std::max(ss, 1)) != hipSuccess) return fail("out of device memory");
            if (ctr_power_iteration(W.ctx, W.mem.l2, 30) != CTR_OK || ctr_synchronize(W.ctx) != CTR_OK)
                return fail(ctr_last_error(W.ctx));
        }
    }

    // batches of the smallest per-device size, handed out in order to
    // whichever device is free
    int dN_plan = workers[0].dN;
    for (const Worker &W : workers) dN_plan = std::min(dN_plan, W.dN);
    std::vector<Chunk> chunks;
    for (int N = p.st; N <= last;) {
        chunks.push_back({N, std::min(N + dN_plan - 1, last)});
        if (chunks.back().end == last) break;
        N += dN_plan - (overlap ? 2 * p.CSoverlap : 0);
    }
    std::mutex q_mu;
    size_t next_chunk = 0;
    auto take = [&](Chunk &c) {
        std::lock_guard<std::mutex> lk(q_mu);
        if (next_chunk >= chunks.size()) return false;
        c = chunks[next_chunk++];
        return true;
    };

    auto read_batch = [&](Chunk c) {
        HostBatch b;
        b.c = c;
        const int dN = c.end - c.start + 1;
        b.prjs.resize(prj_n * dN);
        b.imgs.resize(img_n * dN);
        for (int i = 0; i < dN; ++i) {
            if (!read_raw(p.d1, layer_name(p.f1, c.start + i), &b.prjs[prj_n * i], prj_n, &b.err)) return b;
            if (m != Method::FBP && m != Method::HYBRID && !first_image(p, &b.imgs[img_n * i], img_n, &b.err))
                return b;
        }
        b.ok = true;
        return b;
    };

    // output_thread (GPU.cpp:99-123): writes layers [start + offset, end]
    auto write_batch = [&](HostBatch b, int offset, int out_end) -> std::string {
        std::string e;
        std::vector<float> roi;
        for (int i = b.c.start + offset; i <= out_end; ++i)
            if (!write_raw(p.d2, layer_name(p.f2, i), output_layer(p, &b.imgs[img_n * (i - b.c.start)], roi),
                           (size_t)p.ox * p.oy, &e))
                return e;
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::lock_guard<std::mutex> lk(g_log_mu);
        for (int i = b.c.start + offset; i <= out_end; ++i)
            log << "[Layer " << std::setfill('0') << std::setw(4) << i << "/" << last << "] " << std::fixed
                << std::setprecision(1) << secs << " s elapsed\n";
        log.flush();
        return e;
    };

    std::vector<std::string> errors(workers.size());
    auto work = [&](size_t w) {
        Worker &W = workers[w];
        (void)hipSetDevice(W.dev);
        Chunk c;
        if (!take(c)) return;
        std::future<HostBatch> next = std::async(std::launch::async, read_batch, c);
        std::future<std::string> writing;
        for (;;) {
            HostBatch b = next.get();
            if (!b.ok) { errors[w] = b.err; break; }
            Chunk cn;
            const bool more = take(cn);
            if (more) next = std::async(std::launch::async, read_batch, cn);   // read ahead
            const int dN = b.c.end - b.c.start + 1;
            ctr_status st = CTR_OK;
            if (hipMemcpy(W.mem.prj, b.prjs.data(), sizeof(float) * prj_n * dN, hipMemcpyHostToDevice) != hipSuccess ||
                (m != Method::FBP && m != Method::HYBRID &&
                 hipMemcpy(W.mem.img, b.imgs.data(), sizeof(float) * img_n * dN, hipMemcpyHostToDevice) != hipSuccess)) {
                errors[w] = "host-to-device copy failed";
                if (more) next.wait();
                break;
            }
            switch (m) {
            case Method::OSEM: st = ctr_osem_execution(W.ctx, W.mem.img, W.mem.prj, dN, p.it, 1); break;
            case Method::FISTA_OSEM: st = ctr_fista_osem_execution(W.ctx, W.mem.img, W.mem.prj, dN, p.it, 1); break;
            case Method::AART: st = ctr_aart_execution(W.ctx, W.mem.img, W.mem.prj, dN, p.it); break;
            case Method::FISTA: st = ctr_fista_execution(W.ctx, W.mem.l2, W.mem.img, W.mem.prj, dN, p.it, 1); break;
            case Method::FBP: st = ctr_fbp_execution(W.ctx, W.mem.img, W.mem.prj, dN, 1); break;
            case Method::HYBRID:
                st = ctr_hybrid_execution(W.ctx, W.mem.img, W.mem.prj, dN, p.it, 1, p.baseupOrder, p.CSitBool ? 1 : 0);
                break;
            }
            if (st == CTR_OK) st = ctr_synchronize(W.ctx);
            if (st != CTR_OK) {
                errors[w] = ctr_last_error(W.ctx);
                if (more) next.wait();
                break;
            }
            if (hipMemcpy(b.imgs.data(), W.mem.img, sizeof(float) * img_n * dN, hipMemcpyDeviceToHost) != hipSuccess) {
                errors[w] = "device-to-host copy failed";
                if (more) next.wait();
                break;
            }
            // with overlapping batches the shared layers are written by one batch
            const int offset = (overlap && b.c.start > p.st) ? p.CSoverlap : 0;
            const int out_end = (overlap && b.c.end < last) ? b.c.end - p.CSoverlap : b.c.end;
            if (writing.valid()) {
                const std::string e = writing.get();
                if (!e.empty()) { errors[w] = e; if (more) next.wait(); break; }
            }
            writing = std::async(std::launch::async, write_batch, std::move(b), offset, out_end);
            if (!more) break;
        }
        if (writing.valid()) {
            const std::string e = writing.get();
            if (errors[w].empty() && !e.empty()) errors[w] = e;
        }
    };
    std::vector<std::thread> threads;
    for (size_t w = 0; w < workers.size(); ++w) threads.emplace_back(work, w);
    for (auto &t : threads) t.join();
    int rc = 0;
    for (size_t w = 0; w < workers.size(); ++w) {
        if (!errors[w].empty()) rc = fail("device " + std::to_string(devs[w]) + ": " + errors[w]);
        ctr_destroy(workers[w].ctx);
        workers[w].ctx = nullptr;
    }
    return rc;
}

}  // namespace ctrhost
