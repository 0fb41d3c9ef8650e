// ctr_api.cpp -- C ABI of the MI355X CT projector hot path (declared in include/ctr.h).
//
// The execution drivers below follow the launch sequences of
// CT_reconstruction/GPU.cpp (cited per function).  Where the reference copies
// images between OpenCL image objects because OpenCL 1.2 cannot read and write
// one image in a kernel, the update here is done in place (each
// backprojection thread reads and writes only its own pixels).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <memory>
#include <string>
#include <vector>

#include "ctr.h"
#include "ctr_internal.h"

using namespace ctr;

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t ensure(size_t n)
    {
        if (n == 0) n = 4;
        if (n <= bytes) return hipSuccess;
        release();
        hipError_t e = hipMalloc(&p, n);
        if (e == hipSuccess) bytes = n;
        else p = nullptr;
        return e;
    }
    float *f() { return static_cast<float *>(p); }
};

int round_up(int v, int m) { return (v + m - 1) / m * m; }

// A set of slices with device descriptors and logical-block maps per grid kind.
// The descriptors and the six maps sit in one device block (dev, or a region
// the caller provides), d_sl / d_map pointing into it.
struct SliceSet {
    std::vector<SliceDev> h;
    std::vector<int> hmap[6];
    DevBuf dev;
    const SliceDev *d_sl = nullptr;
    const int *d_map[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int nb[6] = {0, 0, 0, 0, 0, 0};
    int fp_wpb = 4, fp_rays = 64, fp_tile = 60, bp_rows = 4, na = 1;
    // the maps in `dev` and what they were planned from (the schedule and
    // every slice's extents); own_at is null when dev holds no valid maps
    std::vector<int> map_sig;
    const void *own_at = nullptr;
    const int *own_map[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int own_nb[6] = {0, 0, 0, 0, 0, 0};
};

// Pinned host staging for host -> device uploads (slice descriptors, block
// maps, shift tables).  An upload copies its bytes into the next region of a
// pinned ring and queues an asynchronous copy from there on the context
// stream, so a call returns before its uploads complete.  A region is reused
// only after the event recorded behind the copy that read it has completed;
// regions are handed out in ring order, so the spans still in flight that a
// new region can overlap are the oldest ones.
struct Staging {
    struct Span {
        size_t off, len;
        hipEvent_t ev;
    };
    char *host = nullptr;
    size_t cap = 0, head = 0;
    std::deque<Span> busy;
    std::vector<hipEvent_t> idle;
    ~Staging() { release(); }
    void drain()
    {
        for (Span &sp : busy) {
            (void)hipEventSynchronize(sp.ev);
            idle.push_back(sp.ev);
        }
        busy.clear();
    }
    void release()
    {
        drain();
        for (hipEvent_t e : idle) (void)hipEventDestroy(e);
        idle.clear();
        if (host) (void)hipHostFree(host);
        host = nullptr;
        cap = head = 0;
    }
};

struct Layout {
    int nx, ny, px, ipitch, tpitch, rpitch, roff;
};

Layout make_layout(int nx, int ny, int px)
{
    Layout l;
    l.nx = nx; l.ny = ny; l.px = px;
    l.ipitch = round_up(nx + 2 * PAD, 16);
    l.tpitch = round_up(ny + 2 * PAD, 16);
    // Rows hold bin b at entry roff + b with zeros around: a pixel's bin is
    // within hypot(nx, ny)/2 + 1 of px/2, so roff zeros on either side cover
    // every bin the backprojectors can compute, and they read out-of-range
    // bins as 0 (the reference's border) without a clamp.
    l.roff = std::max(1, (int)std::ceil(0.5 * std::hypot((double)nx, (double)ny) - 0.5 * px) + 2);
    l.rpitch = round_up(px + 2 * l.roff, 16);
    return l;
}

size_t pimg_floats(const Layout &l) { return (size_t)(l.ny + 2 * PAD) * l.ipitch; }
size_t pimgT_floats(const Layout &l) { return (size_t)(l.nx + 2 * PAD) * l.tpitch; }

// A projector / backprojector schedule: the LDS-tiled projector (fp_rays = 0)
// with tiles of fp_tile samples, or the direct-gather one with fp_rays rays
// per wave and fp_wpb waves per block; bp_rows rows per backprojection thread.
struct Tune {
    int fp_wpb = 4, fp_rays = 0, fp_tile = 60, bp_rows = 8;   // the defaults
    bool operator==(const Tune &o) const
    {
        return fp_wpb == o.fp_wpb && fp_rays == o.fp_rays && fp_tile == o.fp_tile && bp_rows == o.bp_rows;
    }
};

// The LDS-tiled projector covers a ray of D samples with at most 64 tiles.
bool tile_fits(int tile, int side) { return (long long)tile * 64 >= side; }

// A batched stack: one slice set per distinct schedule (the slices of each
// side run the schedule chosen for that side), all in one set of buffers.
struct BatchPlan {
    std::vector<ctr_slice> key;
    int it_graph = -1;
    std::vector<std::unique_ptr<SliceSet>> sets;
    DevBuf pimg, pimgT, rows;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~BatchPlan()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

// Workspace of ctr_osem_execution_noalloc: a file-scope device array, one
// copy per device that the library runs on, held until the process ends.  Host
// code never uses its name as an address: the device address of each copy is
// resolved once per device (hipGetSymbolAddress, with that device current, at
// the first ctr_create on it) and kept in scratch_ptr.
static __device__ __align__(256) char scratch[CTR_NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
void *scratch_ptr[MAX_DEVICES] = {};
bool scratch_tried[MAX_DEVICES] = {};

// Resolves the current device's copy of `scratch` (call with `device`
// current).  Returns its address, or null if it could not be resolved.
void *resolve_scratch(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_tried[device]) {
        scratch_tried[device] = true;
        void *ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) == hipSuccess) scratch_ptr[device] = ptr;
        else (void)hipGetLastError();
    }
    return scratch_ptr[device];
}

}  // namespace

struct ctr_ctx {
    int device = 0;
    int cus = 256;           // compute units of the device
    bool adaptive_rows = true;   // small stacks run with fewer backprojection rows per thread
    hipStream_t own = nullptr, stream = nullptr;
    bool sync_each = false;
    std::string err;
    bool has_geom = false;
    ctr_geometry g{};
    std::vector<float> ang;
    std::vector<int> sub;
    DevBuf cs, cs_corr, tw;
    int tw_zp = 0;
    // stack workspace
    DevBuf w_pimg, w_pimgT, w_rows, w_pimg2, w_pimgT2, w_aux1, w_aux2, w_aux3, w_partial, w_l2, w_tmpfull, w_min;
    DevBuf w_rcs, w_shift;   // rotation-centre search: candidate sinograms, candidate shifts
    DevBuf w_xproj, w_stage; // sinogram formation: x-projections, sinograms before spike removal
    DevBuf w_gpad;           // alignment adjoint: slice-fastest projection rows with zero margins
    SliceSet set_a, set_b;
    // autotune table: n -> schedule
    std::map<int, Tune> tune;
    std::unique_ptr<BatchPlan> plan;
    Staging stage;
    SliceSet set_na;         // the no-allocation driver's slice set (descriptors in the scratch array)
    char *scratch = nullptr; // this device's copy of the scratch array (device address)
    // kernel timing
    bool timing = false;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    double fp_ms = 0.0, bp_ms = 0.0;
    int fp_n = 0, bp_n = 0;
};

namespace {

ctr_status fail(ctr_ctx *c, ctr_status st, const char *fmt, ...)
{
    if (c) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        c->err = buf;
    }
    return st;
}

ctr_status hipfail(ctr_ctx *c, const char *what, hipError_t e)
{
    return fail(c, e == hipErrorOutOfMemory ? CTR_OUT_OF_MEMORY : CTR_HIP_ERROR, "%s: %s", what,
                hipGetErrorString(e));
}

#define HIPTRY(c, call)                                          \
    do {                                                         \
        hipError_t e_ = (call);                                  \
        if (e_ != hipSuccess) return hipfail((c), #call, e_);    \
    } while (0)

// Called after every launch: reports launch errors, and mirrors GPU.cpp's
// finish() after each enqueue when sync_each is set.
ctr_status post(ctr_ctx *c, const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hipfail(c, what, e);
    if (c->sync_each) {
        e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return hipfail(c, what, e);
    }
    return CTR_OK;
}

#define POST(c, what)                            \
    do {                                         \
        ctr_status s_ = post((c), (what));       \
        if (s_ != CTR_OK) return s_;             \
    } while (0)

#define TRY(expr)                                \
    do {                                         \
        ctr_status s_ = (expr);                  \
        if (s_ != CTR_OK) return s_;             \
    } while (0)

Tune tuned(ctr_ctx *c, int n)
{
    auto it = c->tune.find(n);
    Tune t = it != c->tune.end() ? it->second : Tune{};
    if (t.fp_rays == 0 && !tile_fits(t.fp_tile, n)) t.fp_tile = 60;
    if (t.fp_rays == 0 && n > LDS_FP_MAX_SIDE) {
        t.fp_rays = 64;
        t.fp_wpb = 2;
    }
    return t;
}

// Copies the parts (host pointer, bytes) back to back, each at a 256-byte
// aligned offset, into a pinned staging region and queues one asynchronous
// copy of the region to dst on the context stream.  `may_grow` = 0 forbids
// allocating (a larger ring or a new event): the upload then waits for
// earlier ones instead, or fails when it cannot fit.
struct Part {
    const void *src;
    size_t bytes;
};

size_t parts_bytes(const Part *parts, int n)
{
    size_t t = 0;
    for (int i = 0; i < n; ++i) t = ((t + 255) & ~(size_t)255) + parts[i].bytes;
    return t;
}

ctr_status stage_upload(ctr_ctx *c, void *dst, const Part *parts, int n, bool may_grow)
{
    Staging &st = c->stage;
    const size_t total = std::max(parts_bytes(parts, n), (size_t)16);
    if (total > st.cap) {
        if (!may_grow)
            return fail(c, CTR_INVALID_ARGUMENT, "upload of %zu bytes exceeds the %zu-byte staging ring", total,
                        st.cap);
        st.drain();
        if (st.host) HIPTRY(c, hipHostFree(st.host));
        st.host = nullptr;
        st.cap = st.head = 0;
        const size_t want = std::max(total, (size_t)4 << 20);
        HIPTRY(c, hipHostMalloc((void **)&st.host, want, hipHostMallocDefault));
        st.cap = want;
    }
    size_t off = st.head;
    if (off + total > st.cap) off = 0;
    while (!st.busy.empty()) {
        const Staging::Span &f = st.busy.front();
        if (f.off >= off + total || f.off + f.len <= off) break;       // the oldest span is clear of the region
        HIPTRY(c, hipEventSynchronize(f.ev));
        st.idle.push_back(f.ev);
        st.busy.pop_front();
    }
    size_t o = 0;
    for (int i = 0; i < n; ++i) {
        o = (o + 255) & ~(size_t)255;
        if (parts[i].bytes) std::memcpy(st.host + off + o, parts[i].src, parts[i].bytes);
        o += parts[i].bytes;
    }
    HIPTRY(c, hipMemcpyAsync(dst, st.host + off, total, hipMemcpyHostToDevice, c->stream));
    hipEvent_t ev = nullptr;
    if (!st.idle.empty()) {
        ev = st.idle.back();
        st.idle.pop_back();
    } else if (may_grow) {
        HIPTRY(c, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    } else {
        // reuse the oldest span's event once its copy has completed
        if (st.busy.empty()) return fail(c, CTR_HIP_ERROR, "no staging event available");
        HIPTRY(c, hipEventSynchronize(st.busy.front().ev));
        ev = st.busy.front().ev;
        st.busy.pop_front();
    }
    HIPTRY(c, hipEventRecord(ev, c->stream));
    st.busy.push_back({off, total, ev});
    st.head = off + total;
    return CTR_OK;
}

// Logical blocks of slice d per grid kind (the projection's map is planned
// separately).
void block_counts(const SliceSet &S, const SliceDev &d, int cnt[6])
{
    cnt[0] = 0;
    cnt[G_BP] = bp_blocks(d.nx, d.ny, S.bp_rows);
    cnt[G_TR] = tr_blocks(d.nx, d.ny);
    cnt[G_PAD] = pad_blocks(d.ny, d.ipitch);
    cnt[G_EW] = ew_blocks(d.nx, d.ny);
    cnt[G_SPARE] = 0;
}

// Sets each slice's first logical block per kind, as plan_maps does (host only).
void block_offsets(SliceSet &S)
{
    int at[6] = {0, 0, 0, 0, 0, 0};
    for (SliceDev &d : S.h) {
        int cnt[6];
        block_counts(S, d, cnt);
        for (int k = 1; k < 6; ++k) {
            d.blk[k] = at[k];
            at[k] += cnt[k];
        }
        d.blk[G_FP] = 0;
    }
}

// Builds the per-kind logical-block maps of S (host only).
void plan_maps(SliceSet &S)
{
    const int n = (int)S.h.size();
    for (int k = 0; k < 6; ++k) S.hmap[k].clear();
    for (int s = 0; s < n; ++s) {
        SliceDev &d = S.h[s];
        int cnt[6];
        block_counts(S, d, cnt);
        for (int k = 1; k < 6; ++k) {
            d.blk[k] = (int)S.hmap[k].size();
            S.hmap[k].insert(S.hmap[k].end(), (size_t)cnt[k], s);
        }
        d.blk[G_FP] = 0;
    }
    // Forward projection: a block of a slice with D samples per ray takes time
    // a D + b, so equal block counts are not equal work.  Fitting the
    // LDS-tiled projector's rates at D = 256 and 1024 gives b / a ~ 93 samples
    // (the block's tile list, box loads and epilogue).  The block sequence is
    // cut into 8 segments of equal work (a slice may straddle two), logical
    // segment x runs on XCD x (xcd_remap), and the segments are padded to one
    // length with empty entries.  Entries are (slice, block within the slice)
    // pairs; slice -1 is empty.
    constexpr double BLOCK_COST_SAMPLES = 93.0;
    std::vector<int> cnt(n);
    double total = 0.0;
    for (int s = 0; s < n; ++s) {
        const SliceDev &d = S.h[s];
        cnt[s] = fp_blocks(d.px, S.na, S.fp_wpb, S.fp_rays);
        total += (double)cnt[s] * (std::max(d.nx, d.ny) + BLOCK_COST_SAMPLES);
    }
    std::vector<std::vector<std::pair<int, int>>> seg(8);
    double acc = 0.0;
    int x = 0;
    for (int s = 0; s < n; ++s) {
        const double w = std::max(S.h[s].nx, S.h[s].ny) + BLOCK_COST_SAMPLES;
        for (int b = 0; b < cnt[s]; ++b) {
            while (x < 7 && acc + 0.5 * w > total * (x + 1) / 8.0) ++x;
            seg[x].push_back({s, b});
            acc += w;
        }
    }
    size_t per = 0;
    for (const auto &v : seg) per = std::max(per, v.size());
    std::vector<int> &fm = S.hmap[G_FP];
    fm.assign(2 * 8 * per, -1);
    for (int k = 0; k < 8; ++k)
        for (size_t i = 0; i < seg[k].size(); ++i) {
            fm[2 * (k * per + i)] = seg[k][i].first;
            fm[2 * (k * per + i) + 1] = seg[k][i].second;
        }
    for (int k = 0; k < 6; ++k) {
        const size_t entries = S.hmap[k].size();
        S.nb[k] = (int)(k == G_FP ? entries / 2 : entries);
    }
}

// The parts of S's device block: descriptors, then the six maps.
int set_parts(const SliceSet &S, Part *parts)
{
    parts[0] = {S.h.data(), sizeof(SliceDev) * S.h.size()};
    for (int k = 0; k < 6; ++k) parts[1 + k] = {S.hmap[k].data(), sizeof(int) * S.hmap[k].size()};
    return 7;
}

// Points d_sl / d_map into the device block at `base`, laid out as parts_bytes lays it out.
void bind_set(SliceSet &S, char *base, const Part *parts)
{
    size_t o = 0;
    for (int i = 0; i < 7; ++i) {
        o = (o + 255) & ~(size_t)255;
        if (i == 0) S.d_sl = reinterpret_cast<const SliceDev *>(base + o);
        else S.d_map[i - 1] = reinterpret_cast<const int *>(base + o);
        o += parts[i].bytes;
    }
}

// Plans the maps of `S` and uploads descriptors and maps into its device
// block (grown when needed) or, when `dst` is given, into that region (which
// must hold set_bytes(S) bytes).  Asynchronous: the host vectors can change
// as soon as this returns.
size_t set_bytes(SliceSet &S)
{
    plan_maps(S);
    Part parts[7];
    return parts_bytes(parts, set_parts(S, parts));
}

// The inputs of plan_maps: the schedule and each slice's extents.
void map_signature(const SliceSet &S, std::vector<int> &sig)
{
    sig.clear();
    sig.reserve(5 + 4 * S.h.size());
    sig.insert(sig.end(), {S.na, S.fp_wpb, S.fp_rays, S.fp_tile, S.bp_rows});
    for (const SliceDev &d : S.h) sig.insert(sig.end(), {d.nx, d.ny, d.px, d.ipitch});
}

ctr_status upload_set(ctr_ctx *c, SliceSet &S, char *dst = nullptr)
{
    if (!dst) {
        // A stack called again with the same extents and schedule (an
        // iterative solver calling an operator every iteration) keeps the maps
        // already in its block and uploads only the descriptors, which hold
        // the buffer addresses: the maps of a large stack are megabytes.
        std::vector<int> sig;
        map_signature(S, sig);
        if (S.own_at && S.own_at == S.dev.p && sig == S.map_sig) {
            block_offsets(S);
            const Part p0 = {S.h.data(), sizeof(SliceDev) * S.h.size()};
            TRY(stage_upload(c, S.dev.p, &p0, 1, true));
            S.d_sl = static_cast<const SliceDev *>(S.dev.p);
            for (int k = 0; k < 6; ++k) {
                S.d_map[k] = S.own_map[k];
                S.nb[k] = S.own_nb[k];
            }
            return CTR_OK;
        }
        plan_maps(S);
        Part parts[7];
        const int np = set_parts(S, parts);
        S.own_at = nullptr;
        HIPTRY(c, S.dev.ensure(parts_bytes(parts, np)));
        TRY(stage_upload(c, S.dev.p, parts, np, true));
        bind_set(S, static_cast<char *>(S.dev.p), parts);
        S.map_sig.swap(sig);
        S.own_at = S.dev.p;
        for (int k = 0; k < 6; ++k) {
            S.own_map[k] = S.d_map[k];
            S.own_nb[k] = S.nb[k];
        }
        return CTR_OK;
    }
    plan_maps(S);
    Part parts[7];
    const int np = set_parts(S, parts);
    TRY(stage_upload(c, dst, parts, np, false));
    bind_set(S, dst, parts);
    return CTR_OK;
}

Launch launch_of(ctr_ctx *c, SliceSet &S, int kind)
{
    Launch L;
    L.sl = S.d_sl;
    L.bmap = S.d_map[kind];
    L.cs = static_cast<const float2 *>(c->cs.p);
    L.nblocks = S.nb[kind];
    return L;
}

ctr_status need_geom(ctr_ctx *c)
{
    if (!c) return CTR_INVALID_ARGUMENT;
    if (!c->has_geom) return fail(c, CTR_INVALID_ARGUMENT, "geometry not set (call ctr_set_geometry first)");
    return CTR_OK;
}

// Binds a uniform stack of nz layers of the context geometry.
struct StackBind {
    float *pimg = nullptr, *pimgT = nullptr, *rows = nullptr;
    const float *prj = nullptr, *src = nullptr;
    float *dst = nullptr, *aux = nullptr;
    int layer_base = 0;
};

void fill_stack(ctr_ctx *c, SliceSet &S, int nz, const StackBind &b);

// Rows per backprojection thread for the slices of S: the side's schedule
// (rows), halved while the launch would give the device fewer than 4 blocks
// per compute unit (unless the context runs every stack as scheduled).  One
// slice of 512^2 at 8 rows is 128 blocks; at 1 row, 1024.  Every row count
// gives each pixel the same sum in the same order, so only the launch shape
// changes.
int fill_rows(const ctr_ctx *c, const SliceSet &S, int rows)
{
    if (!c->adaptive_rows) return rows;
    const long long want = 4LL * c->cus;
    auto blocks = [&](int r) {
        long long n = 0;
        for (const SliceDev &d : S.h) n += bp_blocks(d.nx, d.ny, r);
        return n;
    };
    while (rows > 1 && blocks(rows) < want) rows /= 2;
    return rows;
}

ctr_status build_stack(ctr_ctx *c, SliceSet &S, int nz, const StackBind &b)
{
    fill_stack(c, S, nz, b);
    return upload_set(c, S);
}

// The descriptors of a uniform stack of nz layers of the context geometry
// (host only).
void fill_stack(ctr_ctx *c, SliceSet &S, int nz, const StackBind &b)
{
    const ctr_geometry &g = c->g;
    const Layout l = make_layout(g.nx, g.ny, g.px);
    const int na = g.pa / g.ss;
    const Tune t = tuned(c, std::max(g.nx, g.ny));
    S.fp_wpb = t.fp_wpb;
    S.fp_rays = t.fp_rays;
    S.fp_tile = t.fp_tile;
    S.bp_rows = t.bp_rows;
    S.na = na;
    S.h.assign(nz, SliceDev{});
    for (int z = 0; z < nz; ++z) {
        SliceDev &d = S.h[z];
        d.nx = g.nx; d.ny = g.ny; d.px = g.px;
        d.ipitch = l.ipitch; d.tpitch = l.tpitch; d.rpitch = l.rpitch; d.roff = l.roff;
        d.img = b.pimg ? b.pimg + (size_t)z * pimg_floats(l) : nullptr;
        d.imgT = b.pimgT ? b.pimgT + (size_t)z * pimgT_floats(l) : nullptr;
        d.rows = b.rows ? b.rows + (size_t)z * na * l.rpitch : nullptr;
        d.prj = b.prj ? b.prj + (size_t)z * g.pa * g.px : nullptr;
        d.src = b.src ? b.src + (size_t)z * g.nx * g.ny : nullptr;
        d.dst = b.dst ? b.dst + (size_t)z * g.nx * g.ny : nullptr;
        d.aux = b.aux ? b.aux + (size_t)z * g.nx * g.ny : nullptr;
        d.layer = b.layer_base + z;
    }
    S.bp_rows = fill_rows(c, S, t.bp_rows);
}

// Device workspace for an nz-layer stack of the current geometry.
ctr_status ensure_ws(ctr_ctx *c, int nz, bool second_image, int aux_planes)
{
    const ctr_geometry &g = c->g;
    const Layout l = make_layout(g.nx, g.ny, g.px);
    const int na = g.pa / g.ss;
    HIPTRY(c, c->w_pimg.ensure(sizeof(float) * nz * pimg_floats(l)));
    HIPTRY(c, c->w_pimgT.ensure(sizeof(float) * nz * pimgT_floats(l)));
    HIPTRY(c, c->w_rows.ensure(sizeof(float) * (size_t)nz * std::max(na, g.pa) * l.rpitch));
    if (second_image) {
        HIPTRY(c, c->w_pimg2.ensure(sizeof(float) * nz * pimg_floats(l)));
        HIPTRY(c, c->w_pimgT2.ensure(sizeof(float) * nz * pimgT_floats(l)));
    }
    const size_t plane = (size_t)nz * g.nx * g.ny * sizeof(float);
    if (aux_planes > 0) HIPTRY(c, c->w_aux1.ensure(plane));
    if (aux_planes > 1) HIPTRY(c, c->w_aux2.ensure(plane));
    if (aux_planes > 2) HIPTRY(c, c->w_aux3.ensure(plane));
    return CTR_OK;
}

ctr_status zero_rows(ctr_ctx *c, int nz, int nrows)
{
    const Layout l = make_layout(c->g.nx, c->g.ny, c->g.px);
    launch_fill(c->w_rows.f(), (size_t)nz * nrows * l.rpitch, 0.0f, c->stream);
    POST(c, "fill rows");
    return CTR_OK;
}

ctr_status check_stack_args(ctr_ctx *c, const void *img, const void *prj, int dN)
{
    TRY(need_geom(c));
    if (!img || !prj) return fail(c, CTR_INVALID_ARGUMENT, "null buffer");
    if (dN < 1) return fail(c, CTR_INVALID_ARGUMENT, "dN must be >= 1 (got %d)", dN);
    return CTR_OK;
}

ctr_status correct_in_place(ctr_ctx *c, float *prj, int dN)
{
    const ctr_geometry &g = c->g;
    launch_sino_correction(prj, prj, static_cast<const float2 *>(c->cs_corr.p), dN, g.pa, g.px,
                           g.correction_mode, 0.5f, 0.5f, c->stream);
    POST(c, "sinogramCorrection");
    return CTR_OK;
}

void time_begin(ctr_ctx *c, int which)
{
    if (c->timing) (void)hipEventRecord(c->ev[which * 2], c->stream);
}

void time_end(ctr_ctx *c, int which)
{
    if (!c->timing) return;
    (void)hipEventRecord(c->ev[which * 2 + 1], c->stream);
    (void)hipEventSynchronize(c->ev[which * 2 + 1]);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, c->ev[which * 2], c->ev[which * 2 + 1]);
    if (which == 0) { c->fp_ms += ms; c->fp_n++; }
    else { c->bp_ms += ms; c->bp_n++; }
}

ctr_status run_fp(ctr_ctx *c, SliceSet &S, int var, int sub, int sub_from_layer, int geo_power)
{
    FpParams p{sub, c->g.ss, c->g.pa / c->g.ss, sub_from_layer, geo_power, 0};
    time_begin(c, 0);
    launch_fp(var, S.fp_wpb, S.fp_rays, S.fp_tile, launch_of(c, S, G_FP), p, c->stream);
    POST(c, "forward projection");
    time_end(c, 0);
    return CTR_OK;
}

ctr_status run_bp(ctr_ctx *c, SliceSet &S, int var, int sub, int sub_from_layer, float alpha, const float *L,
                  int to_aux, int write_t = 0)
{
    BpParams p{sub, c->g.ss, c->g.pa / c->g.ss, c->g.pa, sub_from_layer, alpha, L, to_aux, 0, write_t};
    time_begin(c, 1);
    launch_bp(var, S.bp_rows, launch_of(c, S, G_BP), p, c->stream);
    POST(c, "backprojection");
    time_end(c, 1);
    return CTR_OK;
}

ctr_status run_kind(ctr_ctx *c, SliceSet &S, int kind)
{
    const Launch L = launch_of(c, S, kind);
    if (kind == G_TR) launch_transpose(L, c->stream);
    else if (kind == G_PAD) launch_pad_in(L, c->stream);
    else launch_unpad_out(L, c->stream);
    POST(c, kind == G_TR ? "transpose" : (kind == G_PAD ? "pad" : "unpad"));
    return CTR_OK;
}

ctr_status angle_tables(ctr_ctx *c)
{
    const int pa = c->g.pa;
    std::vector<float2> cs(pa), cc(pa);
    for (int a = 0; a < pa; ++a) {
        // angle_pr = angle[th]*PI/180.0f (double product, float result), cos/sin in float
        const float ap = (float)((double)c->ang[a] * M_PI / 180.0);
        cs[a] = make_float2(std::cos(ap), std::sin(ap));
        // sinogramCorrection: angle[th]/180*PI (float division first)
        const float ac = (float)((double)(c->ang[a] / 180.0f) * M_PI);
        cc[a] = make_float2(std::cos(ac), 0.0f);
    }
    HIPTRY(c, c->cs.ensure(sizeof(float2) * pa));
    HIPTRY(c, c->cs_corr.ensure(sizeof(float2) * pa));
    HIPTRY(c, hipMemcpy(c->cs.p, cs.data(), sizeof(float2) * pa, hipMemcpyHostToDevice));
    HIPTRY(c, hipMemcpy(c->cs_corr.p, cc.data(), sizeof(float2) * pa, hipMemcpyHostToDevice));
    return CTR_OK;
}

ctr_status twiddles(ctr_ctx *c, int zp)
{
    if (c->tw_zp == zp) return CTR_OK;
    std::vector<float2> tw(std::max(zp / 2, 1));
    for (int l = 0; l < zp / 2; ++l) {
        const double a = 2.0 * M_PI * (double)l / (double)zp;
        tw[l] = make_float2((float)std::cos(a), (float)-std::sin(a));
    }
    HIPTRY(c, c->tw.ensure(sizeof(float2) * tw.size()));
    HIPTRY(c, hipMemcpy(c->tw.p, tw.data(), sizeof(float2) * tw.size(), hipMemcpyHostToDevice));
    c->tw_zp = zp;
    return CTR_OK;
}

bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

}  // namespace

// ===========================================================================
// context

extern "C" const char *ctr_status_string(ctr_status s)
{
    switch (s) {
    case CTR_OK: return "ok";
    case CTR_INVALID_ARGUMENT: return "invalid argument";
    case CTR_INVALID_GEOMETRY: return "invalid geometry";
    case CTR_NO_DEVICE: return "no device";
    case CTR_OUT_OF_MEMORY: return "out of memory";
    case CTR_HIP_ERROR: return "HIP error";
    case CTR_IO_ERROR: return "I/O error";
    case CTR_NOT_SUPPORTED: return "not supported";
    }
    return "unknown status";
}

extern "C" ctr_status ctr_create(int device, ctr_ctx **out)
{
    if (!out) return CTR_INVALID_ARGUMENT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return CTR_NO_DEVICE;
    if (device < 0 || device >= n) return CTR_NO_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return CTR_NO_DEVICE;
    auto *c = new ctr_ctx();
    c->device = device;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        c->cus = cus;
    else
        (void)hipGetLastError();
    if (hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return CTR_HIP_ERROR;
    }
    c->stream = c->own;
    for (auto &e : c->ev) (void)hipEventCreate(&e);
    c->scratch = static_cast<char *>(resolve_scratch(device));
    // the staging ring and its events exist from the start, so uploads of up
    // to 4 MiB (every call of the no-allocation driver) allocate nothing
    constexpr size_t RING = (size_t)4 << 20;
    if (hipHostMalloc((void **)&c->stage.host, RING, hipHostMallocDefault) != hipSuccess) {
        c->stage.host = nullptr;
        (void)hipStreamDestroy(c->own);
        for (auto &e : c->ev) (void)hipEventDestroy(e);
        delete c;
        return CTR_OUT_OF_MEMORY;
    }
    c->stage.cap = RING;
    for (int i = 0; i < 16; ++i) {
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess) c->stage.idle.push_back(e);
    }
    *out = c;
    return CTR_OK;
}

extern "C" ctr_status ctr_destroy(ctr_ctx *c)
{
    if (!c) return CTR_INVALID_ARGUMENT;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    c->plan.reset();
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    hipStream_t own = c->own;
    delete c;  // frees device buffers
    if (own) (void)hipStreamDestroy(own);
    return CTR_OK;
}

extern "C" const char *ctr_last_error(const ctr_ctx *c) { return c ? c->err.c_str() : "null context"; }

extern "C" ctr_status ctr_set_stream(ctr_ctx *c, void *s)
{
    if (!c) return CTR_INVALID_ARGUMENT;
    c->stream = s ? static_cast<hipStream_t>(s) : c->own;
    return CTR_OK;
}

extern "C" ctr_status ctr_get_stream(const ctr_ctx *c, void **s)
{
    if (!c || !s) return CTR_INVALID_ARGUMENT;
    *s = (void *)c->stream;
    return CTR_OK;
}

extern "C" ctr_status ctr_synchronize(ctr_ctx *c)
{
    if (!c) return CTR_INVALID_ARGUMENT;
    HIPTRY(c, hipStreamSynchronize(c->stream));
    return CTR_OK;
}

extern "C" ctr_status ctr_set_sync_each_kernel(ctr_ctx *c, int enable)
{
    if (!c) return CTR_INVALID_ARGUMENT;
    c->sync_each = enable != 0;
    return CTR_OK;
}

// The reference's order (GPU.cpp:564-582): for ss = 2^m, position i holds
// the m-bit reversal of i, so consecutive subsets are as far apart in angle as
// the earlier ones allow (16 -> 0, 8, 4, 12, 2, ...); any other count runs
// 0 .. ss - 1 in order.
extern "C" ctr_status ctr_subset_order(int ss, int *sub)
{
    if (ss < 1 || !sub) return CTR_INVALID_ARGUMENT;
    const bool pow2 = (ss & (ss - 1)) == 0;
    int m = 0;
    while ((1 << m) < ss) ++m;
    for (int i = 0; i < ss; ++i) {
        int r = i;
        if (pow2) {
            r = 0;
            for (int k = 0; k < m; ++k) r |= ((i >> k) & 1) << (m - 1 - k);
        }
        sub[i] = r;
    }
    return CTR_OK;
}

extern "C" ctr_status ctr_query_layout(int nx, int ny, int px, ctr_layout *out)
{
    if (!out || nx < 1 || ny < 1 || px < 1 || nx > 16384 || ny > 16384 || px > 16384) return CTR_INVALID_ARGUMENT;
    const Layout l = make_layout(nx, ny, px);
    out->pad = PAD;
    out->ipitch = l.ipitch;
    out->tpitch = l.tpitch;
    out->rpitch = l.rpitch;
    out->roff = l.roff;
    return CTR_OK;
}

extern "C" ctr_status ctr_set_geometry(ctr_ctx *c, const ctr_geometry *g, const float *angles_deg, const int *sub)
{
    if (!c || !g || !angles_deg) return fail(c, CTR_INVALID_ARGUMENT, "null argument");
    if (g->nx < 2 || g->ny < 2 || g->px < 2 || g->pa < 1 || g->ss < 1)
        return fail(c, CTR_INVALID_GEOMETRY, "sizes must be positive (nx=%d ny=%d px=%d pa=%d ss=%d)", g->nx,
                    g->ny, g->px, g->pa, g->ss);
    if (g->nx > 16384 || g->ny > 16384 || g->px > 16384 || g->pa > (1 << 20))
        return fail(c, CTR_INVALID_GEOMETRY, "sizes beyond the supported maximum (16384 px, 2^20 angles)");
    if (g->ox < 1 || g->ox > g->nx || g->oy < 1 || g->oy > g->ny)
        return fail(c, CTR_INVALID_GEOMETRY, "output ROI (%d x %d) outside the grid (%d x %d)", g->ox, g->oy,
                    g->nx, g->ny);
    if (g->pa % g->ss != 0)
        return fail(c, CTR_INVALID_GEOMETRY, "angle count %d is not a multiple of the subset count %d", g->pa,
                    g->ss);
    if (g->correction_mode < 0 || g->correction_mode > 4)
        return fail(c, CTR_INVALID_GEOMETRY, "correction mode %d outside 0..4", g->correction_mode);
    for (int a = 0; a < g->pa; ++a)
        if (!std::isfinite(angles_deg[a])) return fail(c, CTR_INVALID_ARGUMENT, "angle %d is not finite", a);
    std::vector<int> order(g->ss);
    if (sub) {
        std::vector<char> seen(g->ss, 0);
        for (int k = 0; k < g->ss; ++k) {
            if (sub[k] < 0 || sub[k] >= g->ss || seen[sub[k]])
                return fail(c, CTR_INVALID_ARGUMENT, "subset order is not a permutation of 0..%d", g->ss - 1);
            seen[sub[k]] = 1;
            order[k] = sub[k];
        }
    } else {
        ctr_subset_order(g->ss, order.data());
    }
    HIPTRY(c, hipSetDevice(c->device));
    HIPTRY(c, hipStreamSynchronize(c->stream));
    c->g = *g;
    c->ang.assign(angles_deg, angles_deg + g->pa);
    c->sub = order;
    c->plan.reset();
    c->tw_zp = 0;
    TRY(angle_tables(c));
    c->has_geom = true;
    return CTR_OK;
}

// ===========================================================================
// execution drivers

// An output ROI smaller than the grid is an FBP feature only.  The iterative
// drivers' "extent reconst image" copies (GPU.cpp:199-211, 343-355,
// 1177-1189, 1866-1878) fill the image outside the centred ox x oy block only along the
// centred rows, and on the left only down to the last copy origin >= -ox/2;
// the rest of the device image is never written before the projector reads
// it.  Their OSEM2 / AART2 also update the top-left ox x oy block (an ox x oy
// range over full-grid coordinates) while the copy back takes the centred one.
static ctr_status require_full_grid(ctr_ctx *c)
{
    const ctr_geometry &g = c->g;
    if (g.ox != g.nx || g.oy != g.ny)
        return fail(c, CTR_INVALID_GEOMETRY,
                    "output ROI %d x %d smaller than the grid %d x %d: only FBP supports it", g.ox, g.oy, g.nx, g.ny);
    return CTR_OK;
}

// The OS-EM iterations over a bound stack: pad the images in, build the
// transposed copy, then per subset OSEM1 -> OSEM2 (each update also refreshes
// the transposed copy the next projection reads), and unpad the result.
static ctr_status osem_iterations(ctr_ctx *c, SliceSet &S, int it)
{
    TRY(run_kind(c, S, G_PAD));
    TRY(run_kind(c, S, G_TR));
    for (int i = 0; i < it; ++i)
        for (int k = 0; k < c->g.ss; ++k) {
            TRY(run_fp(c, S, FP_RATIO, c->sub[k], 0, 0));
            TRY(run_bp(c, S, BP_OSEM2, c->sub[k], 0, 0.0f, nullptr, 0, 1));
        }
    TRY(run_kind(c, S, G_EW));
    return CTR_OK;
}

// OSEM_execution, GPU.cpp:125-240: [sinogramCorrection] then for it, for k:
// OSEM1(sub[k]) -> OSEM2(sub[k]) -> image update.
extern "C" ctr_status ctr_osem_execution(ctr_ctx *c, float *img, float *prj, int dN, int it, int prj_correction)
{
    TRY(check_stack_args(c, img, prj, dN));
    TRY(require_full_grid(c));
    const ctr_geometry &g = c->g;
    HIPTRY(c, hipSetDevice(c->device));
    TRY(ensure_ws(c, dN, false, 0));
    TRY(zero_rows(c, dN, g.pa / g.ss));
    if (prj_correction && g.correction_mode > 0) TRY(correct_in_place(c, prj, dN));
    StackBind b;
    b.pimg = c->w_pimg.f(); b.pimgT = c->w_pimgT.f(); b.rows = c->w_rows.f();
    b.prj = prj; b.src = img; b.dst = img;
    SliceSet &S = c->set_a;
    TRY(build_stack(c, S, dN, b));
    return osem_iterations(c, S, it);
}

// ctr_osem_execution with its workspace in the device array `scratch`
// (CTR_NOALLOC_SCRATCH_BYTES per device) instead of context buffers: the
// slice descriptors and block maps, then the padded images, their transposed
// copies and the subset rows of every layer.  Allocates nothing: the
// descriptors go through the staging ring created with the context.
extern "C" ctr_status ctr_osem_execution_noalloc(ctr_ctx *c, float *img, float *prj, int dN, int it,
                                                 int prj_correction)
{
    TRY(check_stack_args(c, img, prj, dN));
    TRY(require_full_grid(c));
    if (it < 0) return fail(c, CTR_INVALID_ARGUMENT, "it must be >= 0 (got %d)", it);
    if (!c->scratch) return fail(c, CTR_HIP_ERROR, "the scratch array's device address could not be resolved");
    const ctr_geometry &g = c->g;
    const Layout l = make_layout(g.nx, g.ny, g.px);
    const int na = g.pa / g.ss;
    const size_t per_layer = sizeof(float) * (pimg_floats(l) + pimgT_floats(l) + (size_t)na * l.rpitch);
    // the descriptor block's size does not depend on the buffer addresses
    SliceSet &S = c->set_na;
    StackBind none;
    fill_stack(c, S, dN, none);
    const size_t desc = (set_bytes(S) + 255) & ~(size_t)255;
    if (desc > CTR_NOALLOC_SCRATCH_BYTES || (size_t)dN > (CTR_NOALLOC_SCRATCH_BYTES - desc) / per_layer)
        return fail(c, CTR_INVALID_ARGUMENT,
                    "%d layers of %d x %d with %d bins need %zu bytes of scratch; the no-allocation driver has %zu",
                    dN, g.nx, g.ny, g.px, desc + (size_t)dN * per_layer, (size_t)CTR_NOALLOC_SCRATCH_BYTES);
    HIPTRY(c, hipSetDevice(c->device));
    char *base = c->scratch + desc;
    StackBind b;
    b.pimg = reinterpret_cast<float *>(base);
    b.pimgT = b.pimg + (size_t)dN * pimg_floats(l);
    b.rows = b.pimgT + (size_t)dN * pimgT_floats(l);
    b.prj = prj; b.src = img; b.dst = img;
    fill_stack(c, S, dN, b);
    // rows start at zero on every call (the array keeps the last call's contents)
    launch_fill(b.rows, (size_t)dN * na * l.rpitch, 0.0f, c->stream);
    POST(c, "fill rows");
    if (prj_correction && g.correction_mode > 0) TRY(correct_in_place(c, prj, dN));
    TRY(upload_set(c, S, c->scratch));
    return osem_iterations(c, S, it);
}

// FISTA_OSEM_execution, GPU.cpp:243-405.  L2norm is filled with 0.5
// (GPU.cpp:326), so lambda = CSlambda.  First sub-iteration: ISTA3D of the
// pre-update image replaces the OS-EM result (GPU.cpp:329-330, 369-377);
// later: x <- prox of the OS-EM update (FISTA3D's x_new, GPU.cpp:337, 383).
// FISTA3D's w and beta outputs never reach the returned image and are not
// computed.
extern "C" ctr_status ctr_fista_osem_execution(ctr_ctx *c, float *img, float *prj, int dN, int it,
                                               int prj_correction)
{
    TRY(check_stack_args(c, img, prj, dN));
    TRY(require_full_grid(c));
    const ctr_geometry &g = c->g;
    HIPTRY(c, hipSetDevice(c->device));
    TRY(ensure_ws(c, dN, false, 1));
    HIPTRY(c, c->w_l2.ensure(sizeof(float) * g.ss));
    TRY(zero_rows(c, dN, g.pa / g.ss));
    launch_fill(c->w_l2.f(), g.ss, 0.5f, c->stream);
    POST(c, "fill L2norm");
    if (prj_correction && g.correction_mode > 0) TRY(correct_in_place(c, prj, dN));
    const Layout l = make_layout(g.nx, g.ny, g.px);
    StackBind b;
    b.pimg = c->w_pimg.f(); b.pimgT = c->w_pimgT.f(); b.rows = c->w_rows.f();
    b.prj = prj; b.src = img; b.dst = img; b.aux = c->w_aux1.f();
    SliceSet &S = c->set_a;
    TRY(build_stack(c, S, dN, b));
    TRY(run_kind(c, S, G_PAD));
    bool first = true;
    for (int i = 0; i < it; ++i)
        for (int k = 0; k < g.ss; ++k) {
            TRY(run_kind(c, S, G_TR));
            TRY(run_fp(c, S, FP_RATIO, c->sub[k], 0, 0));
            TRY(run_bp(c, S, BP_OSEM2, c->sub[k], 0, 0.0f, nullptr, 1));
            // the prox reads an unpadded stack: the initial image (first step,
            // still in the caller's buffer) or the OS-EM update (aux)
            const float *v = first ? img : c->w_aux1.f();
            launch_tv(first ? 0 : 1, nullptr, v, nullptr, nullptr, c->w_pimg.f(), nullptr, 1, l.ipitch, dN, g.nx,
                      g.ny, c->w_l2.f(), c->sub[k], (double)g.cs_lambda, c->stream);
            POST(c, first ? "ISTA3D" : "FISTA3D");
            first = false;
        }
    TRY(run_kind(c, S, G_EW));
    return CTR_OK;
}

// AART_execution, GPU.cpp:1088-1216: [sinogramCorrection when correctionMode >
// 0], the image is replaced by backProjection of subset sub[0]
// (GPU.cpp:1167-1176), then for it, for k: AART1 -> AART2.
extern "C" ctr_status ctr_aart_execution(ctr_ctx *c, float *img, float *prj, int dN, int it)
{
    TRY(check_stack_args(c, img, prj, dN));
    TRY(require_full_grid(c));
    const ctr_geometry &g = c->g;
    HIPTRY(c, hipSetDevice(c->device));
    TRY(ensure_ws(c, dN, false, 0));
    TRY(zero_rows(c, dN, g.pa / g.ss));
    if (g.correction_mode > 0) TRY(correct_in_place(c, prj, dN));
    const float alpha = (g.mode == 3) ? g.wt2 : g.wt1;   // GPU.cpp:1161-1165
    StackBind b;
    b.pimg = c->w_pimg.f(); b.pimgT = c->w_pimgT.f(); b.rows = c->w_rows.f();
    b.prj = prj; b.src = img; b.dst = img;
    SliceSet &S = c->set_a;
    TRY(build_stack(c, S, dN, b));
    TRY(run_kind(c, S, G_PAD));   // zero borders (the interior is overwritten next)
    TRY(run_bp(c, S, BP_INIT, c->sub[0], 0, 0.0f, nullptr, 0));
    for (int i = 0; i < it; ++i)
        for (int k = 0; k < g.ss; ++k) {
            TRY(run_kind(c, S, G_TR));
            TRY(run_fp(c, S, FP_RESID, c->sub[k], 0, 0));
            TRY(run_bp(c, S, BP_AART2, c->sub[k], 0, alpha, nullptr, 0));
        }
    TRY(run_kind(c, S, G_EW));
    return CTR_OK;
}

// FBP_execution, GPU.cpp:665-819: [sinogramCorrection], ramp filter of every
// row (FBP.cl:34-160), backProjectionFBP (FBP.cl:164-188).  Every pixel of the
// grid is written; with an output ROI the reference computes only the
// top-left ox x oy block (GPU.cpp:704), the block FBP_thread reads back
// (GPU.cpp:848-870), and its values are the same as here.
extern "C" ctr_status ctr_fbp_execution(ctr_ctx *c, float *img, float *prj, int dN, int prj_correction)
{
    TRY(check_stack_args(c, img, prj, dN));
    const ctr_geometry &g = c->g;
    if (!is_pow2(g.zp) || g.zp < g.px || g.zp > 16384)
        return fail(c, CTR_INVALID_GEOMETRY, "zero-padding size %d must be a power of two in [px, 16384]", g.zp);
    HIPTRY(c, hipSetDevice(c->device));
    TRY(twiddles(c, g.zp));
    const Layout l = make_layout(g.nx, g.ny, g.px);
    HIPTRY(c, c->w_rows.ensure(sizeof(float) * (size_t)dN * g.pa * l.rpitch));
    launch_fill(c->w_rows.f(), (size_t)dN * g.pa * l.rpitch, 0.0f, c->stream);
    POST(c, "fill rows");
    if (prj_correction && g.correction_mode > 0) TRY(correct_in_place(c, prj, dN));
    launch_fbp_filter(prj, c->w_rows.f(), dN, g.pa, g.px, l.rpitch, l.roff, g.zp,
                      static_cast<const float2 *>(c->tw.p), c->stream);
    POST(c, "FBP filter");
    launch_fbp_bp(c->w_rows.f(), img, static_cast<const float2 *>(c->cs.p), dN, g.nx, g.ny, g.pa, g.px, l.rpitch,
                  l.roff, c->stream);
    POST(c, "backProjectionFBP");
    return CTR_OK;
}

// hybrid_thread, GPU.cpp:1459-1580 (mode 8): FBP_execution with the
// projection correction, the base-up of the FBP image (FBPcorrection = 1:
// findMinimumX/Y + baseUp, GPU.cpp:1538-1566), then OSEM_execution or
// FISTA_OSEM_execution from that image without a second correction.  The
// reference's minimum covers floor(n / work-group size) * work-group size
// columns and rows, the whole layer whenever the sizes are multiples of its
// work-group size; here it is always the whole layer.
extern "C" ctr_status ctr_hybrid_execution(ctr_ctx *c, float *img, float *prj, int dN, int it, int prj_correction,
                                           int baseup_order, int cs)
{
    TRY(check_stack_args(c, img, prj, dN));
    TRY(require_full_grid(c));
    if (baseup_order < 0 || baseup_order > 64) return fail(c, CTR_INVALID_ARGUMENT, "base-up order %d", baseup_order);
    TRY(ctr_fbp_execution(c, img, prj, dN, prj_correction));
    const ctr_geometry &g = c->g;
    const size_t plane = (size_t)g.nx * g.ny;
    HIPTRY(c, c->w_partial.ensure(sizeof(float) * l2_partial_count(dN)));
    HIPTRY(c, c->w_min.ensure(sizeof(float) * (size_t)dN));
    launch_layer_min(img, dN, plane, c->w_partial.f(), c->w_min.f(), c->stream);
    POST(c, "findMinimumX/Y");
    launch_baseup(img, c->w_min.f(), dN, plane, baseup_order, c->stream);
    POST(c, "baseUp");
    return cs ? ctr_fista_osem_execution(c, img, prj, dN, it, 0) : ctr_osem_execution(c, img, prj, dN, it, 0);
}

// ===========================================================================
// sinogram formation (Reslice/, the mt conversion of ImageReg share source/)

namespace {

ctr_status need_ctx(ctr_ctx *c)
{
    if (!c) return CTR_INVALID_ARGUMENT;
    HIPTRY(c, hipSetDevice(c->device));
    return CTR_OK;
}

ctr_status check_frames(ctr_ctx *c, int n, int W, int H, long long stride)
{
    if (n < 1 || W < 1 || H < 1) return fail(c, CTR_INVALID_ARGUMENT, "need n, W, H >= 1 (got %d, %d, %d)", n, W, H);
    if (stride < (long long)W * H)
        return fail(c, CTR_INVALID_ARGUMENT, "frame stride %lld is below W*H = %lld", stride, (long long)W * H);
    return CTR_OK;
}

ctr_status check_stack(ctr_ctx *c, int A, int H, int W)
{
    if (A < 1 || H < 1 || W < 1) return fail(c, CTR_INVALID_ARGUMENT, "need A, H, W >= 1 (got %d, %d, %d)", A, H, W);
    return CTR_OK;
}

}  // namespace

// merge_rawhisdata, imageregistration_kernel_src.cl:868-881.
extern "C" ctr_status ctr_merge_frames(ctr_ctx *c, const unsigned short *raw, int n, int W, int H,
                                       long long frame_stride, float *out)
{
    TRY(need_ctx(c));
    if (!raw || !out) return fail(c, CTR_INVALID_ARGUMENT, "null buffer");
    TRY(check_frames(c, n, W, H, frame_stride));
    launch_merge_frames(raw, out, n, (size_t)W * H, (size_t)frame_stride, c->stream);
    POST(c, "merge_rawhisdata");
    return CTR_OK;
}

// mt_conversion, imageregistration_kernel_src.cl:161-223.
extern "C" ctr_status ctr_mt_conversion(ctr_ctx *c, const float *dark, const float *I0, const unsigned short *It,
                                        int n, int W, int H, long long frame_stride, float *mt)
{
    TRY(need_ctx(c));
    if (!dark || !I0 || !It || !mt) return fail(c, CTR_INVALID_ARGUMENT, "null buffer");
    TRY(check_frames(c, n, W, H, frame_stride));
    launch_mt_conv(dark, I0, It, mt, n, (size_t)W * H, (size_t)frame_stride, c->stream);
    POST(c, "mt_conversion");
    return CTR_OK;
}

// xprojection, reslice.cl:72-88, for every angle (Z_Correction's loop,
// reslice_ocl.cpp:121-133).
extern "C" ctr_status ctr_xprojection(ctr_ctx *c, const float *mt, int A, int H, int W, int start_x, int end_x,
                                      float *xproj)
{
    TRY(need_ctx(c));
    if (!mt || !xproj) return fail(c, CTR_INVALID_ARGUMENT, "null buffer");
    TRY(check_stack(c, A, H, W));
    launch_xproj(mt, xproj, A, H, W, start_x, end_x, c->stream);
    POST(c, "xprojection");
    return CTR_OK;
}

// zcorrection, reslice.cl:90-219 (Z_Correction, reslice_ocl.cpp:136-148).
extern "C" ctr_status ctr_zcorrection(ctr_ctx *c, const float *xproj, int A, int H, int start_z, int end_z,
                                      float *yshift)
{
    TRY(need_ctx(c));
    if (!xproj || !yshift) return fail(c, CTR_INVALID_ARGUMENT, "null buffer");
    if (A < 1 || H < 1) return fail(c, CTR_INVALID_ARGUMENT, "need A, H >= 1 (got %d, %d)", A, H);
    if (H > ZCORR_MAX_H) return fail(c, CTR_NOT_SUPPORTED, "x-projection length %d exceeds %d", H, ZCORR_MAX_H);
    launch_zcorr(xproj, yshift, A, H, start_z, end_z, c->stream);
    POST(c, "zcorrection");
    return CTR_OK;
}

// reslice, reslice.cl:52-70 (host reslice, reslice_ocl.cpp:153-208).
extern "C" ctr_status ctr_reslice(ctr_ctx *c, const float *mt, int A, int H, int W, const float *xshift,
                                  const float *yshift, float baseup, int correction, int y0, int n_layers,
                                  float *prj)
{
    TRY(need_ctx(c));
    if (!mt || !prj) return fail(c, CTR_INVALID_ARGUMENT, "null buffer");
    TRY(check_stack(c, A, H, W));
    if (y0 < 0 || n_layers < 1 || (long long)y0 + n_layers > H)
        return fail(c, CTR_INVALID_ARGUMENT, "layers [%d, %d) outside [0, %d)", y0, y0 + n_layers, H);
    launch_reslice(mt, prj, A, H, W, xshift, yshift, baseup, correction, y0, n_layers, c->stream);
    POST(c, "reslice");
    return CTR_OK;
}

// denoiseSinogram, reslice.cl:222-248 (host reslice_ocl.cpp:211-255).
extern "C" ctr_status ctr_denoise_sinogram(ctr_ctx *c, const float *src, float *dst, int nz, int A, int W,
                                           float threshold)
{
    TRY(need_ctx(c));
    if (!src || !dst) return fail(c, CTR_INVALID_ARGUMENT, "null buffer");
    if (src == dst) return fail(c, CTR_INVALID_ARGUMENT, "denoiseSinogram needs src != dst");
    TRY(check_stack(c, A, nz, W));
    launch_denoise(src, dst, nz, A, W, threshold, c->stream);
    POST(c, "denoiseSinogram");
    return CTR_OK;
}

// reslice_thread, reslice_ocl.cpp:258-350: [Z_Correction] -> reslice ->
// [denoiseSinogram], with the x shifts at zero as the reference leaves them.
extern "C" ctr_status ctr_reslice_execution(ctr_ctx *c, const float *mt, int A, int H, int W,
                                            const ctr_reslice_params *p, float *yshift, float *prj)
{
    TRY(need_ctx(c));
    if (!mt || !p || !yshift || !prj) return fail(c, CTR_INVALID_ARGUMENT, "null argument");
    TRY(check_stack(c, A, H, W));
    if (p->y0 < 0 || p->n_layers < 1 || (long long)p->y0 + p->n_layers > H)
        return fail(c, CTR_INVALID_ARGUMENT, "layers [%d, %d) outside [0, %d)", p->y0, p->y0 + p->n_layers, H);
    launch_fill(yshift, (size_t)A, 0.0f, c->stream);   // enqueueFillBuffer(zshift_buff, 0), reslice_ocl.cpp:293-294
    POST(c, "fill zshift");
    if (p->z_correction) {
        HIPTRY(c, c->w_xproj.ensure(sizeof(float) * (size_t)A * H));
        TRY(ctr_xprojection(c, mt, A, H, W, p->start_x, p->end_x, c->w_xproj.f()));
        TRY(ctr_zcorrection(c, c->w_xproj.f(), A, H, p->start_z, p->end_z, yshift));
    }
    float *out = prj;
    if (p->denoise) {
        HIPTRY(c, c->w_stage.ensure(sizeof(float) * (size_t)p->n_layers * A * W));
        out = c->w_stage.f();
    }
    TRY(ctr_reslice(c, mt, A, H, W, nullptr, yshift, p->baseup, 0, p->y0, p->n_layers, out));
    if (p->denoise) TRY(ctr_denoise_sinogram(c, out, prj, p->n_layers, A, W, p->denoise_threshold));
    return CTR_OK;
}

// ===========================================================================
// rotation-centre search (RotationCenterSearch/)

namespace {

// Candidate i moves the detector by start + step * i, formed in float as the
// reference's host loop and setMask form it (rotationCenterSearch_ocl.cpp:139,
// rotCenterSearch.cl:56).
ctr_status upload_shifts(ctr_ctx *c, int n, float start, float step)
{
    std::vector<float> h((size_t)n);
    for (int i = 0; i < n; ++i) h[i] = start + step * (float)i;
    HIPTRY(c, c->w_shift.ensure(sizeof(float) * (size_t)n));
    const Part part{h.data(), sizeof(float) * (size_t)n};
    return stage_upload(c, c->w_shift.p, &part, 1, true);
}

// setMask's float2 angles {a*PI/180, a*PI/180 - PI_2} are formed in double
// and stored as float, then cos() is taken in float (rotCenterSearch.cl:59-66).
RcsAngles rcs_angles(float ang_min, float ang_max)
{
    const double pi = 3.14159265358979323846, pi_2 = 1.57079632679489661923;
    RcsAngles e;
    e.cmin = std::cos((float)((double)ang_min * pi / 180.0));
    e.smin = std::cos((float)((double)ang_min * pi / 180.0 - pi_2));
    e.cmax = std::cos((float)((double)ang_max * pi / 180.0));
    e.smax = std::cos((float)((double)ang_max * pi / 180.0 - pi_2));
    return e;
}

}  // namespace

// rotCenterShift, rotCenterSearch.cl:35-48, for every candidate of the host
// loop at rotationCenterSearch_ocl.cpp:106-147.
extern "C" ctr_status ctr_op_rot_center_shift(ctr_ctx *c, const float *prj, float *out, int n_shift,
                                              float shift_start, float shift_step)
{
    TRY(need_geom(c));
    if (!prj || !out) return fail(c, CTR_INVALID_ARGUMENT, "null buffer");
    if (n_shift < 1) return fail(c, CTR_INVALID_ARGUMENT, "n_shift must be >= 1 (got %d)", n_shift);
    HIPTRY(c, hipSetDevice(c->device));
    TRY(upload_shifts(c, n_shift, shift_start, shift_step));
    launch_rot_shift(prj, out, c->g.pa, c->g.px, n_shift, c->w_shift.f(), c->stream);
    POST(c, "rotCenterShift");
    return CTR_OK;
}

// setMask + imgFocusIndex, rotCenterSearch.cl:50-92, 154-227, as
// imgFocusIndex_estimation runs them (rotationCenterSearch_ocl.cpp:181-217).
extern "C" ctr_status ctr_op_focus_index(ctr_ctx *c, const float *img, int n, float shift_start, float shift_step,
                                         float ang_min, float ang_max, float *findex)
{
    TRY(need_geom(c));
    if (!img || !findex) return fail(c, CTR_INVALID_ARGUMENT, "null buffer");
    if (n < 1) return fail(c, CTR_INVALID_ARGUMENT, "n must be >= 1 (got %d)", n);
    HIPTRY(c, hipSetDevice(c->device));
    TRY(upload_shifts(c, n, shift_start, shift_step));
    HIPTRY(c, c->w_partial.ensure(sizeof(float2) * l2_partial_count(n)));
    launch_focus_index(img, n, c->g.nx, c->g.ny, c->w_shift.f(), rcs_angles(ang_min, ang_max),
                       static_cast<float2 *>(c->w_partial.p), findex, c->stream);
    POST(c, "imgFocusIndex");
    return CTR_OK;
}

// rotationCenterSearch from the resliced layer on (rotationCenterSearch_ocl.cpp:
// 489-551): shifted candidates, reconstRotCntShift_thread's g_ini fill and
// OSEM_execution(it, prjCorrection = true) (rotationCenterSearch_ocl.cpp:
// 281-331), then the focus indices.  The reference reconstructs the
// candidates in batches of numParallel layers; OS-EM treats every layer on its
// own, so one stack of all candidates gives the same images.
extern "C" ctr_status ctr_rotation_center_search(ctr_ctx *c, const float *prj, int n_shift, float shift_start,
                                                 float shift_step, float ang_min, float ang_max, int it,
                                                 float init, int prj_correction, float *recon, float *findex)
{
    TRY(check_stack_args(c, recon, prj, n_shift));
    if (!findex) return fail(c, CTR_INVALID_ARGUMENT, "null findex");
    if (it < 0) return fail(c, CTR_INVALID_ARGUMENT, "it must be >= 0 (got %d)", it);
    const ctr_geometry &g = c->g;
    HIPTRY(c, hipSetDevice(c->device));
    HIPTRY(c, c->w_rcs.ensure(sizeof(float) * (size_t)n_shift * g.pa * g.px));
    TRY(ctr_op_rot_center_shift(c, prj, c->w_rcs.f(), n_shift, shift_start, shift_step));
    launch_fill(recon, (size_t)n_shift * g.nx * g.ny, init, c->stream);
    POST(c, "fill g_ini");
    TRY(ctr_osem_execution(c, recon, c->w_rcs.f(), n_shift, it, prj_correction));
    return ctr_op_focus_index(c, recon, n_shift, shift_start, shift_step, ang_min, ang_max, findex);
}

// powerIteration, GPU.cpp:2079-2188: ss layers, layer s = subset s, filled
// with 1.0; each trial powerIter1 -> powerIter2 -> imageL2AbsX/Y -> powerIter3.
extern "C" ctr_status ctr_power_iteration(ctr_ctx *c, float *L2norm, int trials)
{
    TRY(need_geom(c));
    if (!L2norm || trials < 1) return fail(c, CTR_INVALID_ARGUMENT, "need L2norm and trials >= 1");
    const ctr_geometry &g = c->g;
    const int S = g.ss;
    HIPTRY(c, hipSetDevice(c->device));
    TRY(ensure_ws(c, S, false, 2));
    HIPTRY(c, c->w_partial.ensure(sizeof(float) * l2_partial_count(S)));
    TRY(zero_rows(c, S, g.pa / g.ss));
    const size_t plane = (size_t)g.nx * g.ny;
    launch_fill(c->w_aux2.f(), plane * S, 1.0f, c->stream);
    POST(c, "fill ones");
    const Layout l = make_layout(g.nx, g.ny, g.px);
    StackBind b;
    b.pimg = c->w_pimg.f(); b.pimgT = c->w_pimgT.f(); b.rows = c->w_rows.f();
    b.src = c->w_aux2.f(); b.aux = c->w_aux1.f();
    SliceSet &St = c->set_a;
    TRY(build_stack(c, St, S, b));
    TRY(run_kind(c, St, G_PAD));
    for (int t = 0; t < trials; ++t) {
        TRY(run_kind(c, St, G_TR));
        TRY(run_fp(c, St, FP_RAW, 0, 1, 1));
        TRY(run_bp(c, St, BP_POWER, 0, 1, 0.0f, nullptr, 1));
        launch_l2(c->w_aux1.f(), S, plane, c->w_partial.f(), L2norm, c->stream);
        POST(c, "imageL2Abs");
        launch_power3(c->w_aux1.f(), L2norm, c->w_pimg.f(), S, g.nx, g.ny, l.ipitch, c->stream);
        POST(c, "powerIter3");
    }
    return CTR_OK;
}

// FISTA_execution, GPU.cpp:1755-1924, with the reference's argument bindings
// (GPU.cpp:1827-1849, 1898-1899): the first step projects x0 (backProjection
// of sub[0]) and ISTA3D writes x1 into the returned image; later steps project
// w, and FISTA3D(x = x1, v, beta) updates w and beta, its x_new output being
// overwritten by beta (both bound to beta_dest_img).
extern "C" ctr_status ctr_fista_execution(ctr_ctx *c, const float *L2norm, float *img, float *prj, int dN, int it,
                                          int prj_correction)
{
    TRY(check_stack_args(c, img, prj, dN));
    TRY(require_full_grid(c));
    if (!L2norm) return fail(c, CTR_INVALID_ARGUMENT, "null L2norm");
    const ctr_geometry &g = c->g;
    HIPTRY(c, hipSetDevice(c->device));
    TRY(ensure_ws(c, dN, true, 2));
    TRY(zero_rows(c, dN, g.pa / g.ss));
    if (prj_correction && g.correction_mode > 0) TRY(correct_in_place(c, prj, dN));
    const Layout l = make_layout(g.nx, g.ny, g.px);
    const size_t plane = (size_t)g.nx * g.ny * dN;
    launch_fill(c->w_aux2.f(), plane, 0.0f, c->stream);   // beta = 0 (GPU.cpp:1799)
    POST(c, "fill beta");
    // set_a: R (x) in pimg; set_b: W in pimg2.  aux1 = v, aux2 = beta.
    StackBind br;
    br.pimg = c->w_pimg.f(); br.pimgT = c->w_pimgT.f(); br.rows = c->w_rows.f();
    br.prj = prj; br.src = img; br.dst = img; br.aux = c->w_aux1.f();
    TRY(build_stack(c, c->set_a, dN, br));
    StackBind bw = br;
    bw.pimg = c->w_pimg2.f(); bw.pimgT = c->w_pimgT2.f();
    TRY(build_stack(c, c->set_b, dN, bw));
    // x0 = backProjection(y, sub[0]) into R; W = R (GPU.cpp:1852-1864)
    TRY(run_kind(c, c->set_a, G_PAD));
    TRY(run_bp(c, c->set_a, BP_INIT, c->sub[0], 0, 0.0f, nullptr, 0));
    launch_copy(c->w_pimg.f(), c->w_pimg2.f(), (size_t)dN * pimg_floats(l), c->stream);
    POST(c, "copy x0 -> w");
    TRY(run_kind(c, c->set_a, G_EW));   // caller's buffer holds x0 (FISTA3D's x input later)
    bool first = true;
    for (int i = 0; i < it; ++i)
        for (int k = 0; k < g.ss; ++k) {
            SliceSet &S = first ? c->set_a : c->set_b;
            TRY(run_kind(c, S, G_TR));
            TRY(run_fp(c, S, FP_RESID, c->sub[k], 0, 0));
            TRY(run_bp(c, S, BP_FISTA, c->sub[k], 0, g.wt1, L2norm, 1));
            if (first) {
                // ISTA3D(v) -> the returned image x1
                launch_tv(0, nullptr, c->w_aux1.f(), nullptr, nullptr, img, nullptr, 0, l.ipitch, dN, g.nx, g.ny,
                          L2norm, c->sub[k], (double)g.cs_lambda, c->stream);
                POST(c, "ISTA3D");
                first = false;
            } else {
                // FISTA3D(x = x1, v, beta) -> w (padded, next projection source), beta in place
                launch_tv(1, img, c->w_aux1.f(), c->w_aux2.f(), c->w_pimg2.f(), nullptr, c->w_aux2.f(), 1,
                          l.ipitch, dN, g.nx, g.ny, L2norm, c->sub[k], (double)g.cs_lambda, c->stream);
                POST(c, "FISTA3D");
            }
        }
    return CTR_OK;
}

// ===========================================================================
// single operators

namespace {

ctr_status op_prologue(ctr_ctx *c, const void *a, const void *b, int dN)
{
    TRY(need_geom(c));
    if (!a || !b) return fail(c, CTR_INVALID_ARGUMENT, "null buffer");
    if (dN < 1) return fail(c, CTR_INVALID_ARGUMENT, "dN must be >= 1");
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, CTR_HIP_ERROR, "hipSetDevice failed");
    return CTR_OK;
}

ctr_status check_sub(ctr_ctx *c, int sub)
{
    if (sub < 0 || sub >= c->g.ss) return fail(c, CTR_INVALID_ARGUMENT, "subset %d outside 0..%d", sub, c->g.ss - 1);
    return CTR_OK;
}

// pad `img` (may be null for zero) into the workspace and bind a stack set
ctr_status op_stack(ctr_ctx *c, int dN, const float *img, const float *prj, float *aux, bool pad, int layer_base = 0)
{
    TRY(ensure_ws(c, dN, false, 0));
    TRY(zero_rows(c, dN, std::max(c->g.pa / c->g.ss, 1)));
    StackBind b;
    b.pimg = c->w_pimg.f(); b.pimgT = c->w_pimgT.f(); b.rows = c->w_rows.f();
    b.prj = prj; b.src = img; b.aux = aux; b.layer_base = layer_base;
    TRY(build_stack(c, c->set_a, dN, b));
    if (pad) {
        TRY(run_kind(c, c->set_a, G_PAD));
        TRY(run_kind(c, c->set_a, G_TR));
    }
    return CTR_OK;
}

ctr_status op_fp(ctr_ctx *c, int var, const float *img, const float *prj, float *out_full, int dN, int sub)
{
    TRY(op_prologue(c, img, out_full, dN));
    TRY(check_sub(c, sub));
    TRY(op_stack(c, dN, img, prj, nullptr, true));
    TRY(run_fp(c, c->set_a, var, sub, 0, 0));
    const Layout l = make_layout(c->g.nx, c->g.ny, c->g.px);
    launch_rows_to_full(c->w_rows.f(), out_full, dN, c->g.pa / c->g.ss, l.rpitch, l.roff, c->g.pa, c->g.px, sub,
                        c->g.ss, c->stream);
    POST(c, "rows to full");
    return CTR_OK;
}

ctr_status op_bp(ctr_ctx *c, int var, const float *img, const float *rows_full, float *dst, int dN, int sub,
                 float alpha, const float *L)
{
    TRY(op_prologue(c, rows_full, dst, dN));
    TRY(check_sub(c, sub));
    const bool raw = (var == BP_INIT);
    TRY(op_stack(c, dN, img, raw ? rows_full : nullptr, dst, img != nullptr));
    if (!raw) {
        const Layout l = make_layout(c->g.nx, c->g.ny, c->g.px);
        launch_full_to_rows(rows_full, c->w_rows.f(), dN, c->g.pa / c->g.ss, l.rpitch, l.roff, c->g.pa, c->g.px, sub,
                            c->g.ss, c->stream);
        POST(c, "full to rows");
    }
    TRY(run_bp(c, c->set_a, var, sub, 0, alpha, L, 1));
    return CTR_OK;
}

}  // namespace

extern "C" ctr_status ctr_op_osem1(ctr_ctx *c, const float *img, const float *prj, float *rprj, int dN, int sub)
{
    if (!prj) return fail(c, CTR_INVALID_ARGUMENT, "null prj");
    return op_fp(c, FP_RATIO, img, prj, rprj, dN, sub);
}

extern "C" ctr_status ctr_op_aart1(ctr_ctx *c, const float *img, const float *prj, float *dprj, int dN, int sub)
{
    if (!prj) return fail(c, CTR_INVALID_ARGUMENT, "null prj");
    return op_fp(c, FP_RESID, img, prj, dprj, dN, sub);
}

extern "C" ctr_status ctr_op_osem2(ctr_ctx *c, const float *img, const float *rprj, float *img_dst, int dN, int sub)
{
    if (!img) return fail(c, CTR_INVALID_ARGUMENT, "null img");
    return op_bp(c, BP_OSEM2, img, rprj, img_dst, dN, sub, 0.0f, nullptr);
}

extern "C" ctr_status ctr_op_aart2(ctr_ctx *c, const float *img, const float *dprj, float *img_dst, int dN, int sub,
                                   float alpha)
{
    if (!img) return fail(c, CTR_INVALID_ARGUMENT, "null img");
    return op_bp(c, BP_AART2, img, dprj, img_dst, dN, sub, alpha, nullptr);
}

extern "C" ctr_status ctr_op_backprojection(ctr_ctx *c, const float *prj, float *img_dst, int dN, int sub)
{
    return op_bp(c, BP_INIT, nullptr, prj, img_dst, dN, sub, 0.0f, nullptr);
}

extern "C" ctr_status ctr_op_fista_backprojection(ctr_ctx *c, const float *img, const float *dprj, const float *L,
                                                  float *img_dst, int dN, int sub, float alpha)
{
    if (!img || !L) return fail(c, CTR_INVALID_ARGUMENT, "null img or L2norm");
    return op_bp(c, BP_FISTA, img, dprj, img_dst, dN, sub, alpha, L);
}

extern "C" ctr_status ctr_op_sinogram_correction(ctr_ctx *c, const float *src, float *dst, int dN, int mode)
{
    TRY(op_prologue(c, src, dst, dN));
    if (mode < 0 || mode > 4) return fail(c, CTR_INVALID_ARGUMENT, "mode %d outside 0..4", mode);
    launch_sino_correction(src, dst, static_cast<const float2 *>(c->cs_corr.p), dN, c->g.pa, c->g.px, mode, 0.5f,
                           0.5f, c->stream);
    POST(c, "sinogramCorrection");
    return CTR_OK;
}

extern "C" ctr_status ctr_op_fbp_filter(ctr_ctx *c, const float *prj, float *fprj, int dN)
{
    TRY(op_prologue(c, prj, fprj, dN));
    const ctr_geometry &g = c->g;
    if (!is_pow2(g.zp) || g.zp < g.px || g.zp > 16384)
        return fail(c, CTR_INVALID_GEOMETRY, "zero-padding size %d must be a power of two in [px, 16384]", g.zp);
    TRY(twiddles(c, g.zp));
    const Layout l = make_layout(g.nx, g.ny, g.px);
    HIPTRY(c, c->w_rows.ensure(sizeof(float) * (size_t)dN * g.pa * l.rpitch));
    launch_fill(c->w_rows.f(), (size_t)dN * g.pa * l.rpitch, 0.0f, c->stream);
    POST(c, "fill rows");
    launch_fbp_filter(prj, c->w_rows.f(), dN, g.pa, g.px, l.rpitch, l.roff, g.zp,
                      static_cast<const float2 *>(c->tw.p), c->stream);
    POST(c, "FBP filter");
    // all rows are "subset 0 of 1": copy the padded rows back out
    launch_rows_to_full(c->w_rows.f(), fprj, dN, g.pa, l.rpitch, l.roff, g.pa, g.px, 0, 1, c->stream);
    POST(c, "rows to full");
    return CTR_OK;
}

extern "C" ctr_status ctr_op_fbp_backprojection(ctr_ctx *c, const float *fprj, float *img, int dN)
{
    TRY(op_prologue(c, fprj, img, dN));
    const ctr_geometry &g = c->g;
    const Layout l = make_layout(g.nx, g.ny, g.px);
    HIPTRY(c, c->w_rows.ensure(sizeof(float) * (size_t)dN * g.pa * l.rpitch));
    launch_fill(c->w_rows.f(), (size_t)dN * g.pa * l.rpitch, 0.0f, c->stream);
    POST(c, "fill rows");
    launch_full_to_rows(fprj, c->w_rows.f(), dN, g.pa, l.rpitch, l.roff, g.pa, g.px, 0, 1, c->stream);
    POST(c, "full to rows");
    launch_fbp_bp(c->w_rows.f(), img, static_cast<const float2 *>(c->cs.p), dN, g.nx, g.ny, g.pa, g.px, l.rpitch,
                  l.roff, c->stream);
    POST(c, "backProjectionFBP");
    return CTR_OK;
}

extern "C" ctr_status ctr_op_power_iter1(ctr_ctx *c, const float *img, float *prjc)
{
    TRY(op_prologue(c, img, prjc, 1));
    const ctr_geometry &g = c->g;
    TRY(op_stack(c, g.ss, img, nullptr, nullptr, true));
    TRY(run_fp(c, c->set_a, FP_RAW, 0, 1, 1));
    // compact rows of layer s go to rows 0..na-1 of layer s (FISTA.cl:61)
    const Layout l = make_layout(g.nx, g.ny, g.px);
    launch_rows_to_full(c->w_rows.f(), prjc, g.ss, g.pa / g.ss, l.rpitch, l.roff, g.pa, g.px, 0, 1, c->stream);
    POST(c, "rows to full");
    return CTR_OK;
}

extern "C" ctr_status ctr_op_power_iter2(ctr_ctx *c, const float *prjc, float *cnd)
{
    TRY(op_prologue(c, prjc, cnd, 1));
    const ctr_geometry &g = c->g;
    TRY(op_stack(c, g.ss, nullptr, nullptr, cnd, false));
    const Layout l = make_layout(g.nx, g.ny, g.px);
    launch_full_to_rows(prjc, c->w_rows.f(), g.ss, g.pa / g.ss, l.rpitch, l.roff, g.pa, g.px, 0, 1, c->stream);
    POST(c, "full to rows");
    TRY(run_bp(c, c->set_a, BP_POWER, 0, 1, 0.0f, nullptr, 1));
    return CTR_OK;
}

extern "C" ctr_status ctr_op_ista3d(ctr_ctx *c, const float *v, float *x_new, int dN, int sub, const float *L)
{
    TRY(op_prologue(c, v, x_new, dN));
    TRY(check_sub(c, sub));
    if (!L) return fail(c, CTR_INVALID_ARGUMENT, "null L2norm");
    launch_tv(0, nullptr, v, nullptr, nullptr, x_new, nullptr, 0, 0, dN, c->g.nx, c->g.ny, L, sub,
              (double)c->g.cs_lambda, c->stream);
    POST(c, "ISTA3D");
    return CTR_OK;
}

extern "C" ctr_status ctr_op_fista3d(ctr_ctx *c, const float *x, const float *v, const float *b, float *w_new,
                                     float *x_new, float *b_new, int dN, int sub, const float *L)
{
    TRY(op_prologue(c, x, v, dN));
    TRY(check_sub(c, sub));
    if (!b || !L) return fail(c, CTR_INVALID_ARGUMENT, "null beta or L2norm");
    if (x_new && x_new == b_new) x_new = nullptr;   // reference write order: beta remains
    launch_tv(1, x, v, b, w_new, x_new, b_new, 0, 0, dN, c->g.nx, c->g.ny, L, sub, (double)c->g.cs_lambda,
              c->stream);
    POST(c, "FISTA3D");
    return CTR_OK;
}

// The per-slice TV prox of the joint alignment model (no reference
// counterpart; the 2-D form of ISTA3D, FISTA.cl:409-493).
extern "C" ctr_status ctr_tv_prox_2d(ctr_ctx *c, const float *v, float *x_new, int dN, float lambda, float lower)
{
    TRY(op_prologue(c, v, x_new, dN));
    if (!(lambda >= 0.0f) || !std::isfinite(lambda)) return fail(c, CTR_INVALID_ARGUMENT, "lambda must be finite and >= 0");
    if (v == x_new) return fail(c, CTR_INVALID_ARGUMENT, "the TV prox needs v != x_new");
    launch_tv2d(v, x_new, dN, c->g.nx, c->g.ny, lambda, lower, c->stream);
    POST(c, "TV prox 2-D");
    return CTR_OK;
}

// ===========================================================================
// batched uneven stacks

namespace {

bool same_key(const std::vector<ctr_slice> &a, const ctr_slice *b, int n)
{
    if ((int)a.size() != n) return false;
    for (int i = 0; i < n; ++i)
        if (a[i].n != b[i].n || a[i].img != b[i].img || a[i].prj != b[i].prj) return false;
    return true;
}

// One OS-EM iteration over the whole plan: per subset a projection and a
// backprojection launch, each covering every slice; the backprojection also
// writes the transposed copy the next projection reads.
ctr_status batched_iteration(ctr_ctx *c, BatchPlan &P)
{
    for (int k = 0; k < c->g.ss; ++k) {
        for (auto &S : P.sets) TRY(run_fp(c, *S, FP_RATIO, c->sub[k], 0, 0));
        for (auto &S : P.sets) TRY(run_bp(c, *S, BP_OSEM2, c->sub[k], 0, 0.0f, nullptr, 0, 1));
    }
    return CTR_OK;
}

}  // namespace

extern "C" ctr_status ctr_osem_batched(ctr_ctx *c, const ctr_slice *slices, int count, int it, int use_graph)
{
    TRY(need_geom(c));
    if (!slices || count < 1 || it < 0) return fail(c, CTR_INVALID_ARGUMENT, "need slices, count >= 1, it >= 0");
    for (int i = 0; i < count; ++i)
        if (slices[i].n < 2 || slices[i].n > 16384 || !slices[i].img || !slices[i].prj)
            return fail(c, CTR_INVALID_ARGUMENT, "slice %d: side %d or null buffer", i, slices[i].n);
    HIPTRY(c, hipSetDevice(c->device));
    const int na = c->g.pa / c->g.ss;
    if (!c->plan || !same_key(c->plan->key, slices, count)) {
        HIPTRY(c, hipStreamSynchronize(c->stream));
        c->plan.reset(new BatchPlan());
        BatchPlan &P = *c->plan;
        P.key.assign(slices, slices + count);
        // largest slices first: their blocks start earliest, evening out the tail
        std::vector<int> order(count);
        for (int i = 0; i < count; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return slices[a].n > slices[b].n; });
        size_t fi = 0, ft = 0, fr = 0;
        std::vector<Layout> lay(count);
        for (int i = 0; i < count; ++i) {
            lay[i] = make_layout(slices[i].n, slices[i].n, slices[i].n);
            fi += pimg_floats(lay[i]);
            ft += pimgT_floats(lay[i]);
            fr += (size_t)na * lay[i].rpitch;
        }
        HIPTRY(c, P.pimg.ensure(sizeof(float) * fi));
        HIPTRY(c, P.pimgT.ensure(sizeof(float) * ft));
        HIPTRY(c, P.rows.ensure(sizeof(float) * fr));
        launch_fill(P.rows.f(), fr, 0.0f, c->stream);
        POST(c, "fill rows");
        // one set per distinct schedule, in the order of each schedule's largest side
        std::vector<Tune> tunes;
        std::vector<int> set_of(count);
        for (int r = 0; r < count; ++r) {
            const Tune t = tuned(c, slices[order[r]].n);
            int k = 0;
            while (k < (int)tunes.size() && !(tunes[k] == t)) ++k;
            if (k == (int)tunes.size()) tunes.push_back(t);
            set_of[r] = k;
        }
        for (const Tune &t : tunes) {
            auto S = std::make_unique<SliceSet>();
            S->fp_wpb = t.fp_wpb;
            S->fp_rays = t.fp_rays;
            S->fp_tile = t.fp_tile;
            S->bp_rows = t.bp_rows;
            S->na = na;
            P.sets.push_back(std::move(S));
        }
        size_t oi = 0, ot = 0, orr = 0;
        for (int r = 0; r < count; ++r) {
            const int i = order[r];
            const Layout &l = lay[i];
            SliceDev d{};
            d.nx = d.ny = d.px = slices[i].n;
            d.ipitch = l.ipitch; d.tpitch = l.tpitch; d.rpitch = l.rpitch; d.roff = l.roff;
            d.img = P.pimg.f() + oi;
            d.imgT = P.pimgT.f() + ot;
            d.rows = P.rows.f() + orr;
            d.prj = slices[i].prj;
            d.src = slices[i].img;
            d.dst = slices[i].img;
            d.layer = r;
            P.sets[set_of[r]]->h.push_back(d);
            oi += pimg_floats(l);
            ot += pimgT_floats(l);
            orr += (size_t)na * l.rpitch;
        }
        for (auto &S : P.sets) {
            S->bp_rows = fill_rows(c, *S, S->bp_rows);
            TRY(upload_set(c, *S));
        }
    }
    BatchPlan &P = *c->plan;
    for (auto &S : P.sets) {
        TRY(run_kind(c, *S, G_PAD));
        TRY(run_kind(c, *S, G_TR));
    }
    const bool graph = use_graph && !c->sync_each && !c->timing;
    if (graph && it > 0 && !P.exec) {
        HIPTRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        ctr_status st = batched_iteration(c, P);
        hipGraph_t gph = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &gph);
        if (st != CTR_OK) {
            if (gph) (void)hipGraphDestroy(gph);
            return st;
        }
        if (e != hipSuccess) return hipfail(c, "hipStreamEndCapture", e);
        P.graph = gph;
        HIPTRY(c, hipGraphInstantiate(&P.exec, P.graph, nullptr, nullptr, 0));
    }
    for (int i = 0; i < it; ++i) {
        if (graph) {
            HIPTRY(c, hipGraphLaunch(P.exec, c->stream));
        } else {
            TRY(batched_iteration(c, P));
        }
    }
    for (auto &S : P.sets) TRY(run_kind(c, *S, G_EW));
    return CTR_OK;
}

// ===========================================================================
// timing + autotune

extern "C" ctr_status ctr_set_kernel_timing(ctr_ctx *c, int enable)
{
    if (!c) return CTR_INVALID_ARGUMENT;
    c->timing = enable != 0;
    c->fp_ms = c->bp_ms = 0.0;
    c->fp_n = c->bp_n = 0;
    return CTR_OK;
}

extern "C" ctr_status ctr_kernel_timing(ctr_ctx *c, double *fp_ms, double *bp_ms, int *fp_n, int *bp_n)
{
    if (!c) return CTR_INVALID_ARGUMENT;
    if (fp_ms) *fp_ms = c->fp_n ? c->fp_ms / c->fp_n : 0.0;
    if (bp_ms) *bp_ms = c->bp_n ? c->bp_ms / c->bp_n : 0.0;
    if (fp_n) *fp_n = c->fp_n;
    if (bp_n) *bp_n = c->bp_n;
    return CTR_OK;
}

namespace {

bool schedule_ok(const ctr_schedule &s, int n)
{
    const bool lds = s.fp_rays == 0;
    return n >= 2 && n <= 16384 &&
           (lds ? (s.fp_waves == 4 && (s.fp_tile == 32 || s.fp_tile == 44 || s.fp_tile == 60) &&
                   tile_fits(s.fp_tile, n) && n <= LDS_FP_MAX_SIDE)
                : ((s.fp_rays == 64 || s.fp_rays == 32 || s.fp_rays == 16 || s.fp_rays == 8) &&
                   (s.fp_waves == 2 || s.fp_waves == 4))) &&
           (s.bp_rows == 1 || s.bp_rows == 2 || s.bp_rows == 4 || s.bp_rows == 8 || s.bp_rows == 16);
}

void to_schedule(const Tune &t, ctr_schedule *s)
{
    s->fp_rays = t.fp_rays;
    s->fp_waves = t.fp_wpb;
    s->fp_tile = t.fp_tile;
    s->bp_rows = t.bp_rows;
}

}  // namespace

extern "C" ctr_status ctr_set_schedule(ctr_ctx *c, int n, const ctr_schedule *s)
{
    if (!c || !s) return CTR_INVALID_ARGUMENT;
    if (!schedule_ok(*s, n))
        return fail(c, CTR_INVALID_ARGUMENT, "schedule (n=%d, rays=%d, waves=%d, tile=%d, rows=%d) not supported", n,
                    s->fp_rays, s->fp_waves, s->fp_tile, s->bp_rows);
    Tune t;
    t.fp_rays = s->fp_rays;
    t.fp_wpb = s->fp_waves;
    t.fp_tile = s->fp_rays == 0 ? s->fp_tile : 60;
    t.bp_rows = s->bp_rows;
    c->tune[n] = t;
    if (c->plan) {
        // the cached batched plan holds the old schedule: rebuild it, once
        // its graph and buffers are no longer in use
        HIPTRY(c, hipSetDevice(c->device));
        HIPTRY(c, hipStreamSynchronize(c->stream));
        c->plan.reset();
    }
    return CTR_OK;
}

extern "C" ctr_status ctr_set_adaptive_rows(ctr_ctx *c, int enable)
{
    if (!c) return CTR_INVALID_ARGUMENT;
    if (c->adaptive_rows != (enable != 0)) {
        c->adaptive_rows = enable != 0;
        if (c->plan) {
            // a cached batched plan holds its sets' row counts: rebuild it,
            // once its graph and buffers are no longer in use
            HIPTRY(c, hipSetDevice(c->device));
            HIPTRY(c, hipStreamSynchronize(c->stream));
            c->plan.reset();
        }
    }
    return CTR_OK;
}

extern "C" ctr_status ctr_get_schedule(ctr_ctx *c, int n, ctr_schedule *s)
{
    if (!c || !s || n < 2 || n > 16384) return CTR_INVALID_ARGUMENT;
    *s = ctr_schedule{};
    to_schedule(tuned(c, n), s);
    return CTR_OK;
}

// Times every projector and backprojector schedule for side n on the device
// and keeps the fastest: the LDS-tiled projector with tiles of 60, 44 and 32
// samples (those whose 64 tiles cover the side), the direct-gather projector
// with 64 / 32 / 16 / 8 rays per wave and 2 or 4 waves per block, then 2, 4,
// 8 or 16 backprojection rows per thread.  Each candidate is timed on a
// launch of batch size: copies of one slice adding up to about eight 1024^2
// slices of work (128 copies of a 256^2 slice), which fills the device at any
// side.  The reference fixes its work-group shape per compiled program
// (GPU.cpp:169-172, 538-547).
extern "C" ctr_status ctr_autotune(ctr_ctx *c, int n, ctr_schedule *best_out)
{
    TRY(need_geom(c));
    if (n < 2 || n > 16384) return fail(c, CTR_INVALID_ARGUMENT, "side %d out of range", n);
    HIPTRY(c, hipSetDevice(c->device));
    const int na = c->g.pa / c->g.ss;
    const int reps = 5;
    // Synthetic slices of ones, each with its own image, transposed image, rows
    // and sinogram, enough of them (1 GiB, 2..4096 slices) that a launch streams
    // its slices from HBM as a large batch does: copies that shared one image
    // would run from the caches and favour the schedules that re-read it most.
    // Each part starts on a 256-byte boundary, as in the workspaces.
    const Layout l = make_layout(n, n, n);
    auto up64 = [](size_t f) { return (f + 63) / 64 * 64; };
    const size_t f_img = up64(pimg_floats(l)), f_imgT = up64(pimgT_floats(l)),
                 f_rows = up64((size_t)na * l.rpitch), f_prj = up64((size_t)c->g.pa * n);
    const size_t f_slice = f_img + f_imgT + f_rows + f_prj;
    const size_t target = (size_t)1 << 30;
    const int copies = (int)std::max<size_t>(2, std::min<size_t>(4096, (target + 4 * f_slice - 1) / (4 * f_slice)));
    DevBuf mem;
    HIPTRY(c, mem.ensure(sizeof(float) * f_slice * copies));
    launch_fill(mem.f(), f_slice * copies, 1.0f, c->stream);
    POST(c, "autotune fill");
    SliceSet S;
    S.na = na;
    S.h.assign(copies, SliceDev{});
    for (int k = 0; k < copies; ++k) {
        SliceDev &d = S.h[k];
        float *base = mem.f() + f_slice * k;
        d.nx = d.ny = d.px = n;
        d.ipitch = l.ipitch; d.tpitch = l.tpitch; d.rpitch = l.rpitch; d.roff = l.roff;
        d.img = base; d.imgT = base + f_img; d.rows = base + f_img + f_imgT; d.prj = base + f_img + f_imgT + f_rows;
    }
    hipEvent_t e0, e1;
    HIPTRY(c, hipEventCreate(&e0));
    HIPTRY(c, hipEventCreate(&e1));
    auto timeit = [&](auto &&fn) -> double {
        fn();
        (void)hipEventRecord(e0, c->stream);
        for (int r = 0; r < reps; ++r) fn();
        (void)hipEventRecord(e1, c->stream);
        (void)hipEventSynchronize(e1);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        return ms / reps;
    };
    Tune def;                                    // the default schedule for this side
    if (n > LDS_FP_MAX_SIDE) {
        def.fp_rays = 64;
        def.fp_wpb = 2;
    }
    std::vector<Tune> fp_cands;
    for (int tile : {60, 44, 32})
        if (n <= LDS_FP_MAX_SIDE && tile_fits(tile, n)) {
            Tune t;
            t.fp_tile = tile;
            fp_cands.push_back(t);
        }
    for (int rays : {64, 32, 16, 8})
        for (int w : {2, 4}) {
            Tune t;
            t.fp_rays = rays;
            t.fp_wpb = w;
            fp_cands.push_back(t);
        }
    Tune best_t = def;
    double best = 1e30, def_fp = 0.0, def_bp = 0.0;
    for (const Tune &t : fp_cands) {
        S.fp_wpb = t.fp_wpb;
        S.fp_rays = t.fp_rays;
        S.fp_tile = t.fp_tile;
        S.bp_rows = 4;
        TRY(upload_set(c, S));
        FpParams p{0, c->g.ss, na, 0, 0, 0};
        const double ms = timeit([&] { launch_fp(FP_RATIO, t.fp_wpb, t.fp_rays, t.fp_tile, launch_of(c, S, G_FP), p,
                                                 c->stream); });
        if (t.fp_rays == def.fp_rays && t.fp_wpb == def.fp_wpb && t.fp_tile == def.fp_tile) def_fp = ms;
        if (ms < best) {
            best = ms;
            best_t.fp_wpb = t.fp_wpb;
            best_t.fp_rays = t.fp_rays;
            best_t.fp_tile = t.fp_tile;
        }
    }
    const double best_fp = best;
    best = 1e30;
    for (int r : {2, 4, 8, 16}) {
        S.fp_wpb = best_t.fp_wpb;
        S.fp_rays = best_t.fp_rays;
        S.fp_tile = best_t.fp_tile;
        S.bp_rows = r;
        TRY(upload_set(c, S));
        BpParams p{0, c->g.ss, na, c->g.pa, 0, 0.0f, nullptr, 0, 0};
        const double ms = timeit([&] { launch_bp(BP_OSEM2, r, launch_of(c, S, G_BP), p, c->stream); });
        if (r == def.bp_rows) def_bp = ms;
        if (ms < best) { best = ms; best_t.bp_rows = r; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPTRY(c, hipGetLastError());
    HIPTRY(c, hipStreamSynchronize(c->stream));
    c->tune[n] = best_t;
    c->plan.reset();   // plans pick up the new choice on rebuild
    if (best_out) {
        to_schedule(best_t, best_out);
        best_out->fp_ms = (float)best_fp;
        best_out->bp_ms = (float)best;
        best_out->fp_ms_default = (float)def_fp;
        best_out->bp_ms_default = (float)def_bp;
    }
    return CTR_OK;
}

// ===========================================================================
// differentiable alignment model

namespace {

ctr_status align_pad(ctr_ctx *c, const float *img, int dN)
{
    const ctr_geometry &g = c->g;
    TRY(ensure_ws(c, dN, false, 0));
    StackBind b;
    b.pimg = c->w_pimg.f(); b.pimgT = c->w_pimgT.f(); b.rows = c->w_rows.f(); b.src = img;
    TRY(build_stack(c, c->set_a, dN, b));
    TRY(run_kind(c, c->set_a, G_PAD));
    TRY(run_kind(c, c->set_a, G_TR));
    (void)g;
    return CTR_OK;
}

}  // namespace

namespace {

ctr_status align_forward(ctr_ctx *c, const float *img, const float *shifts, float center, float *prj, float *dprj,
                         int dN)
{
    TRY(op_prologue(c, img, prj, dN));
    if (!std::isfinite(center)) return fail(c, CTR_INVALID_ARGUMENT, "center is not finite");
    TRY(align_pad(c, img, dN));
    const ctr_geometry &g = c->g;
    const Layout l = make_layout(g.nx, g.ny, g.px);
    launch_align_forward(c->w_pimg.f(), c->w_pimgT.f(), l.ipitch, l.tpitch, static_cast<const float2 *>(c->cs.p),
                         shifts, center, prj, dprj, dN, g.nx, g.ny, g.pa, g.px, c->stream);
    POST(c, "align forward");
    return CTR_OK;
}

}  // namespace

extern "C" ctr_status ctr_align_forward(ctr_ctx *c, const float *img, const float *shifts, float center, float *prj,
                                        int dN)
{
    return align_forward(c, img, shifts, center, prj, nullptr, dN);
}

extern "C" ctr_status ctr_align_forward_deriv(ctr_ctx *c, const float *img, const float *shifts, float center,
                                              float *prj, float *dprj, int dN)
{
    if (c && !dprj) return fail(c, CTR_INVALID_ARGUMENT, "dprj is null");
    return align_forward(c, img, shifts, center, prj, dprj, dN);
}

extern "C" ctr_status ctr_align_adjoint(ctr_ctx *c, const float *prj, const float *shifts, float center, float *img,
                                        int dN)
{
    TRY(op_prologue(c, prj, img, dN));
    if (!std::isfinite(center)) return fail(c, CTR_INVALID_ARGUMENT, "center is not finite");
    const ctr_geometry &g = c->g;
    HIPTRY(c, c->w_gpad.ensure(sizeof(float) * align_adj_floats(dN, g.pa, g.px)));
    launch_align_adjoint(prj, c->w_gpad.f(), static_cast<const float2 *>(c->cs.p), shifts, center, img, dN, g.nx,
                         g.ny, g.pa, g.px, c->stream);
    POST(c, "align adjoint");
    return CTR_OK;
}

extern "C" ctr_status ctr_align_grad_shift(ctr_ctx *c, const float *img, const float *gprj, const float *shifts,
                                           float center, float *gshift, int dN)
{
    TRY(op_prologue(c, img, gprj, dN));
    if (!gshift) return fail(c, CTR_INVALID_ARGUMENT, "null gshift");
    if (!std::isfinite(center)) return fail(c, CTR_INVALID_ARGUMENT, "center is not finite");
    TRY(align_pad(c, img, dN));
    const ctr_geometry &g = c->g;
    const Layout l = make_layout(g.nx, g.ny, g.px);
    const size_t parts = (size_t)g.pa * dN * ((g.px + 63) / 64);
    HIPTRY(c, c->w_partial.ensure(sizeof(float) * parts));
    launch_align_grad(c->w_pimg.f(), c->w_pimgT.f(), l.ipitch, l.tpitch, static_cast<const float2 *>(c->cs.p), gprj,
                      shifts, center, c->w_partial.f(), gshift, dN, g.nx, g.ny, g.pa, g.px, c->stream);
    POST(c, "align grad");
    return CTR_OK;
}
