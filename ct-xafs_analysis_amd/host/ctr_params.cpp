// ctr_params.cpp -- input-parameter file, file formats and layer names of the
// CT reconstruction host (CT_reconstruction_input_parameter.cpp, fileIO.cpp,
// main.cpp of the reference).
#include "ctr_host.hpp"

#include <algorithm>
#include <cerrno>
#include <cstring>
#include <fstream>
#include <sstream>

namespace ctrhost {

namespace {

void set_err(std::string *err, const std::string &msg)
{
    if (err) *err = msg;
}

// Next non-empty line's first token (the reference reads these values with
// operator>>, which skips whitespace and line ends).
template <typename T>
bool read_value(std::istream &in, T &out)
{
    std::string line;
    while (in) {
        line = getline_cr(in);
        std::istringstream iss(line);
        if (iss >> out) return true;
        if (!line.empty() && line.find_first_not_of(" \t") != std::string::npos) return false;
    }
    return false;
}

}  // namespace

std::string getline_cr(std::istream &in)
{
    std::string s;
    std::getline(in, s);
    if (!s.empty() && s.back() == '\r') s.pop_back();
    return s;
}

bool parse_inp(const std::string &path, Params &p, std::string *err)
{
    std::ifstream in(path);
    if (!in) {
        set_err(err, "input parameter file not found: " + path);
        return false;
    }
    // Keys of CT_reconstruction_input_parameter.cpp:47-179, verbatim: each key
    // line is followed by its value on the next line.
    enum Kind { STR, INT, DBL, FLT, BOOL01, SIZE_ALL };
    struct Key {
        const char *text;
        Kind kind;
        void *dst;
    };
    int size_all = 0;
    const Key keys[] = {
        {"#Open CL Platform & Device No", STR, &p.devList},
        {"#Number of Parallel processing per one device", INT, &p.numParallel},
        {"#測定ファイルディレクトリパス", STR, &p.d1},
        {"#測定ファイル名", STR, &p.f1},
        {"#出力ファイルディレクトリパス", STR, &p.d2},
        {"#出力ファイル名", STR, &p.f2},
        {"#ログファイルパス", STR, &p.f3},
        {"#初期画像ファイル名", STR, &p.f4},
        {"#一時ファイルディレクトリパス", STR, &p.d5},
        {"#レイヤー数", INT, &p.num},
        {"#開始ナンバー", INT, &p.st},
        {"#初期画像の使用 1:Yes 2: No", INT, &p.fst},
        {"#カメラピクセル数", INT, &p.px},
        {"#投影数", INT, &p.pa},
        {"#再構成画像サイズ", SIZE_ALL, &size_all},
        {"#再構成画像サイズ X", INT, &p.nx},
        {"#再構成画像サイズ Y", INT, &p.ny},
        {"#出力画像サイズ X", INT, &p.ox},
        {"#出力画像サイズ Y", INT, &p.oy},
        {"#再構成法", INT, &p.mode},
        {"#反復回数", INT, &p.it},
        {"#AARTファクター", DBL, &p.wt1},
        {"#サブセット(OSEM)", INT, &p.ss},
        {"#ゼロパディングサイズ(FBP)", INT, &p.zp},
        {"#投影像補正 0:なし 1: x方向のみ 2:θ方向のみ 3:x,θ方向両方", INT, &p.correctionMode},
        {"#再構成像強度増幅因子", FLT, &p.amp},
        {"#Base up 減少速度次数 (Hybrid)", INT, &p.baseupOrder},
        {"#逐次圧縮センシング計算 (FBP適応外)", BOOL01, &p.CSitBool},
        {"#逐次圧縮センシング計算加算ファクター", FLT, &p.CSlambda},
        {"#逐次圧縮センシング計算重複レイヤー数", INT, &p.CSoverlap},
    };
    bool first = true;
    while (in) {
        std::string line = getline_cr(in);
        if (first) {
            // a UTF-8 byte-order mark before the first key
            if (line.compare(0, 3, "\xEF\xBB\xBF") == 0) line.erase(0, 3);
            first = false;
        }
        for (const Key &k : keys) {
            if (line != k.text) continue;
            bool ok = true;
            switch (k.kind) {
            case STR: *static_cast<std::string *>(k.dst) = getline_cr(in); break;
            case INT: ok = read_value(in, *static_cast<int *>(k.dst)); break;
            case DBL: ok = read_value(in, *static_cast<double *>(k.dst)); break;
            case FLT: ok = read_value(in, *static_cast<float *>(k.dst)); break;
            case BOOL01: {
                int v = 0;
                ok = read_value(in, v);
                *static_cast<bool *>(k.dst) = (v == 1);
                break;
            }
            case SIZE_ALL: {
                // one size for the grid and the output (input_parameter.cpp:127-134)
                int v = 0;
                ok = read_value(in, v);
                p.nx = p.ny = p.ox = p.oy = v;
                break;
            }
            }
            if (!ok) {
                set_err(err, std::string("bad value for key ") + k.text + " in " + path);
                return false;
            }
            break;
        }
    }
    return true;
}

std::string layer_name(const std::string &pattern, int layer)
{
    const size_t a = pattern.find_first_of('*');
    if (a == std::string::npos) return pattern;
    const size_t b = pattern.find_last_of('*');
    char num[32];
    std::snprintf(num, sizeof num, "%04d", layer);
    return pattern.substr(0, a) + num + pattern.substr(b + 1);
}

bool parse_device_list(const std::string &list, std::vector<int> &devices, std::string *err)
{
    devices.clear();
    std::stringstream ss(list);
    std::string item;
    while (std::getline(ss, item, ',')) {
        if (item.find_first_not_of(" \t") == std::string::npos) continue;
        int a = 0, b = 0;
        char dash = 0;
        std::istringstream is(item);
        if (!(is >> a)) {
            set_err(err, "bad device list: " + list);
            return false;
        }
        b = a;
        if (is >> dash) {
            if (dash != '-' || !(is >> b)) {
                set_err(err, "bad device list: " + list);
                return false;
            }
        }
        if (a < 1 || b < a) {
            set_err(err, "bad device list: " + list);
            return false;
        }
        for (int d = a; d <= b; ++d) {
            bool seen = false;
            for (int e : devices) seen = seen || e == d;
            if (!seen) devices.push_back(d);
        }
    }
    if (devices.empty()) {
        set_err(err, "empty device list");
        return false;
    }
    return true;
}

bool read_log(const std::string &path, int pa, std::vector<float> &ang, std::string *err)
{
    std::ifstream in(path);
    if (!in) {
        set_err(err, "file open error [" + path + "]");
        return false;
    }
    ang.assign(pa, 0.0f);
    for (int k = 0; k < pa; ++k) {
        std::istringstream iss(getline_cr(in));
        if (!(iss >> ang[k])) {
            set_err(err, "angle log " + path + " has fewer than " + std::to_string(pa) + " angles");
            return false;
        }
    }
    return true;
}

// fileIO.cpp:27 joins with '/'; an empty directory means the name as given
// (the reference would turn it into "/name").
static std::string join_path(const std::string &dir, const std::string &name)
{
    return dir.empty() ? name : dir + "/" + name;
}

bool read_raw(const std::string &dir, const std::string &name, float *data, size_t count, std::string *err)
{
    const std::string f = join_path(dir, name);
    std::ifstream in(f, std::ios::binary);
    if (!in) {
        set_err(err, "file open error [" + f + "]");
        return false;
    }
    in.read(reinterpret_cast<char *>(data), (std::streamsize)(sizeof(float) * count));
    if ((size_t)in.gcount() != sizeof(float) * count) {
        set_err(err, "file " + f + " holds fewer than " + std::to_string(count) + " floats");
        return false;
    }
    return true;
}

const float *output_layer(const Params &p, const float *img, std::vector<float> &tmp)
{
    if (p.ox == p.nx && p.oy == p.ny) return img;
    tmp.resize((size_t)p.ox * p.oy);
    for (int y = 0; y < p.oy; ++y)
        std::copy(img + (size_t)y * p.nx, img + (size_t)y * p.nx + p.ox, tmp.data() + (size_t)y * p.ox);
    return tmp.data();
}

bool write_raw(const std::string &dir, const std::string &name, const float *data, size_t count,
               std::string *err)
{
    const std::string f = join_path(dir, name);
    std::ofstream out(f, std::ios::binary | std::ios::trunc);
    if (!out) {
        set_err(err, "file open error [" + f + "]");
        return false;
    }
    out.write(reinterpret_cast<const char *>(data), (std::streamsize)(sizeof(float) * count));
    if (!out) {
        set_err(err, "write error [" + f + "]");
        return false;
    }
    return true;
}

bool read_his(const std::string &path, int start_frame, int end_frame, int W, int H, std::vector<unsigned short> &out,
              std::string *err)
{
    std::ifstream in(path, std::ios::binary);
    if (!in) {
        set_err(err, "failed to load his: " + path);
        return false;
    }
    unsigned char hdr[4];
    if (!in.read(reinterpret_cast<char *>(hdr), 4)) {
        set_err(err, "his file " + path + " has no header");
        return false;
    }
    const unsigned comment = (unsigned)hdr[2] | ((unsigned)hdr[3] << 8);   // little-endian
    const long long frame = (long long)W * H + 32;
    const long long skip = (long long)comment + 64 + (long long)sizeof(unsigned short) * frame * (start_frame - 1);
    in.seekg(skip, std::ios::cur);
    const long long n = frame * (end_frame - start_frame + 1);
    out.assign((size_t)n, 0);
    in.read(reinterpret_cast<char *>(out.data()), (std::streamsize)(sizeof(unsigned short) * n));
    // the last frame needs its W*H pixels, not the 32 after them
    if (in.gcount() < (std::streamsize)(sizeof(unsigned short) * (n - 32))) {
        set_err(err, "his file " + path + " holds fewer than " + std::to_string(end_frame) + " frames of " +
                         std::to_string(W) + " x " + std::to_string(H));
        return false;
    }
    return true;
}

bool read_his_rows(const std::string &path, int start_frame, int end_frame, int W, int H, int row0, int nrows,
                   std::vector<unsigned short> &out, std::string *err)
{
    if (row0 < 0 || nrows < 1 || row0 + nrows > H || start_frame < 1 || end_frame < start_frame) {
        set_err(err, "bad frame or row range for " + path);
        return false;
    }
    std::ifstream in(path, std::ios::binary);
    if (!in) {
        set_err(err, "failed to load his: " + path);
        return false;
    }
    unsigned char hdr[4];
    if (!in.read(reinterpret_cast<char *>(hdr), 4)) {
        set_err(err, "his file " + path + " has no header");
        return false;
    }
    const long long data0 = 4 + ((long long)hdr[2] | ((long long)hdr[3] << 8)) + 64;
    const long long frame = (long long)W * H + 32;
    const size_t per = (size_t)nrows * W;
    out.assign(per * (size_t)(end_frame - start_frame + 1), 0);
    for (int f = start_frame; f <= end_frame; ++f) {
        const long long off = data0 + (long long)sizeof(unsigned short) * (frame * (f - 1) + (long long)row0 * W);
        in.seekg(off, std::ios::beg);
        in.read(reinterpret_cast<char *>(out.data() + per * (f - start_frame)), (std::streamsize)(sizeof(unsigned short) * per));
        if (in.gcount() != (std::streamsize)(sizeof(unsigned short) * per)) {
            set_err(err, "his file " + path + " holds fewer than " + std::to_string(end_frame) + " frames of " +
                             std::to_string(W) + " x " + std::to_string(H));
            return false;
        }
    }
    return true;
}

bool first_image(const Params &p, float *img, size_t count, std::string *err)
{
    switch (p.fst) {
    case 1:
        return read_raw("", p.f4, img, count, err);
    case 2:
        for (size_t i = 0; i < count; ++i) img[i] = (float)p.ini;
        return true;
    default:
        // the reference leaves the buffer as allocated
        for (size_t i = 0; i < count; ++i) img[i] = 0.0f;
        return true;
    }
}

int method_subsets(const Params &p)
{
    switch (p.mode) {
    case 2: return p.pa;   // MART: OSEM_ocl(..., g_pa)
    case 3:                // ASIRT: AART_ocl / FISTA_ocl(..., 1)
    case 4:                // MSIRT: OSEM_ocl(..., 1)
    case 5: return 1;      // ML-EM: OSEM_ocl(..., 1)
    default: return p.ss;
    }
}

bool validate(const Params &p, std::string *err)
{
    if (p.mode < 1 || p.mode > 8) {
        set_err(err, "reconstruction method " + std::to_string(p.mode) + " is not 1-8");
        return false;
    }
    // an output ROI is read back from FBP only (the top-left ox x oy block,
    // GPU.cpp:848-870); the iterative methods' ROI extension is not provided
    if (p.ox < 1 || p.ox > p.nx || p.oy < 1 || p.oy > p.ny) {
        set_err(err, "output size must be within the reconstruction size");
        return false;
    }
    if ((p.ox != p.nx || p.oy != p.ny) && p.mode != 7) {
        set_err(err, "output size must equal the reconstruction size except for FBP (ROI extension is not provided)");
        return false;
    }
    if (p.nx < 2 || p.ny < 2 || p.px < 2 || p.pa < 1 || p.num < 1 || p.it < 0 || p.numParallel < 1) {
        set_err(err, "sizes, layer count and parallel count must be positive");
        return false;
    }
    const int ss = method_subsets(p);
    if (p.mode != 7 && (ss < 1 || p.pa % ss != 0)) {   // FBP alone has no subsets
        set_err(err, "the projection count " + std::to_string(p.pa) + " is not a multiple of the subset count " +
                         std::to_string(ss));
        return false;
    }
    if ((p.mode == 7 || p.mode == 8) && (p.zp < p.px || (p.zp & (p.zp - 1)) != 0)) {
        set_err(err, "the zero-padding size must be a power of two >= the detector width");
        return false;
    }
    if (p.mode == 8 && (p.baseupOrder < 0 || p.baseupOrder > 64)) {
        set_err(err, "base-up order must be 0-64");
        return false;
    }
    if (p.correctionMode < 0 || p.correctionMode > 4) {
        set_err(err, "projection correction mode must be 0-4");
        return false;
    }
    return true;
}

}  // namespace ctrhost
