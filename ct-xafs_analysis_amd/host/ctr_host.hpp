// ctr_host.hpp -- C++ host layer of the MI355X CT reconstruction path.
//
// Re-creates the reference's CT_reconstruction host program on top of the
// C ABI in include/ctr.h: the input-parameter file, the angle log and raw
// float32 file formats, the layer-batch pipeline (read -> reconstruct ->
// write, one pipeline per device) and the reconstruction-method dispatch.
// Reconstruction runs in libctr_hip.so; nothing here computes on the CPU.
#ifndef CTR_HOST_HPP
#define CTR_HOST_HPP

#include <iosfwd>
#include <string>
#include <vector>

namespace ctrhost {

// The run parameters: the globals of CT_reconstruction/main.cpp:12-41 with
// the same defaults.
struct Params {
    std::string d1 = "input";      // g_d1: input (sinogram) directory
    std::string f1 = "i****.raw";  // g_f1: input file pattern, **** = layer number
    std::string d2 = "output";     // g_d2: output directory
    std::string f2 = "r****.img";  // g_f2: output file pattern
    std::string f3 = "log.dat";    // g_f3: angle log (one angle in degrees per line)
    std::string d4 = ".";          // g_d4
    std::string f4 = "1st.img";    // g_f4: initial image file (g_fst == 1)
    std::string d5 = "temp";       // g_d5: temporary directory
    int fst = 2;                   // g_fst: 1 = initial image from f4, 2 = constant g_ini
    double ini = 1.0;              // g_ini
    int px = 2048, pa = 1600;      // g_px, g_pa
    int nx = 2048, ny = 2048;      // g_nx, g_ny
    int ox = 2048, oy = 2048;      // g_ox, g_oy
    int mode = 6;                  // g_mode: 1 AART 2 MART 3 ASIRT 4 MSIRT 5 ML-EM 6 OS-EM 7 FBP 8 hybrid
    int it = 1;                    // g_it
    int num = 1, st = 1;           // g_num layers starting at g_st
    double wt1 = 1.0, wt2 = 0.01;  // g_wt1 (AART), g_wt2 (ASIRT)
    int ss = 20;                   // g_ss
    int zp = 4096;                 // g_zp
    std::string devList = "1";     // g_devList: 1-based device numbers, "1,2-4"
    int numParallel = 10;          // layers per device batch (capped by the compute-unit count)
    int correctionMode = 0;        // projection correction 0-4
    float amp = 1.0f;              // parsed, unused by the reconstruction (as in the reference)
    int baseupOrder = 4;           // hybrid mode only
    bool CSitBool = false;         // FISTA-TV (compressed-sensing) iteration
    float CSlambda = 0.0001f;      // TV weight
    int CSoverlap = 1;             // overlapping layers between batches for the 3-D TV
};

// Reads one line and drops a trailing '\r' (share source/CTXAFS_analysis_share.cpp:14-27).
std::string getline_cr(std::istream &in);

// Applies the keys of an input-parameter file
// (CT_reconstruction_input_parameter.cpp:40-179) to `p`.  Unknown lines are
// ignored, like the reference.  Returns false (with a message) when the file
// cannot be opened or a value does not parse.
bool parse_inp(const std::string &path, Params &p, std::string *err);

// Layer file name: the text before the first '*', the 4-digit zero-padded
// layer number, the text after the last '*' (GPU.cpp:26-33, 101-103).
std::string layer_name(const std::string &pattern, int layer);

// Device numbers of a list like "1,3-4" (1-based, in order, no duplicates).
bool parse_device_list(const std::string &list, std::vector<int> &devices, std::string *err);

// fileIO.cpp:11-24: the first `pa` numbers of the angle log.
bool read_log(const std::string &path, int pa, std::vector<float> &ang, std::string *err);
// fileIO.cpp:25-34 / 35-44: raw little-endian float32, no header.
bool read_raw(const std::string &dir, const std::string &name, float *data, size_t count, std::string *err);
bool write_raw(const std::string &dir, const std::string &name, const float *data, size_t count,
               std::string *err);
// readHisFile_stream (share source/readHisRawFile.cpp:63-95): frames
// start_frame .. end_frame (1-based) of a .his file as 16-bit pixels, each
// frame W*H pixels followed by 32 more (the next frame's header), so frame k
// starts at out + (k - start_frame) * (W*H + 32).  The data begin 4 + c + 64
// bytes into the file, c being the 16-bit value at byte 2.
// With row0 / nrows only rows row0 .. row0 + nrows - 1 of each frame are
// read, packed as [frames][nrows][W].
bool read_his(const std::string &path, int start_frame, int end_frame, int W, int H, std::vector<unsigned short> &out,
              std::string *err);
bool read_his_rows(const std::string &path, int start_frame, int end_frame, int W, int H, int row0, int nrows,
                   std::vector<unsigned short> &out, std::string *err);

// fileIO.cpp:55-68: the initial image (file for fst == 1, constant ini for fst == 2).
bool first_image(const Params &p, float *img, size_t count, std::string *err);
// The output layer of an nx x ny image: the image itself, or with an output
// ROI (FBP only) its top-left ox x oy block, the region FBP_thread reads back
// from origin (0, 0) (GPU.cpp:848-870), copied into `tmp`.
const float *output_layer(const Params &p, const float *img, std::vector<float> &tmp);

// Effective subset count of each method (GPU.cpp:36-93: MART uses one
// subset per projection, ASIRT / MSIRT / ML-EM one subset).
int method_subsets(const Params &p);

// Checks the combination of parameters the library supports.
bool validate(const Params &p, std::string *err);

// Runs the whole reconstruction (GPU.cpp:23-97 and the *_ocl pipelines):
// every layer g_st .. g_st + g_num - 1 is read, reconstructed on the listed
// devices in batches of up to numParallel layers, and written.  Progress goes
// to `log`.  Returns 0 on success, non-zero after printing the error.
int run(const Params &p, std::ostream &log);

// ---- multi-energy driver (Reslice_CTreconst/) ------------------------------

// Parameters of the multi-energy driver: the CT_reconstruction keys of the
// same file (Reslice_CTreconst/main.cpp:13-48 defaults) and the shared keys it
// reads (share source/input_parameter_class.cpp:318-470,
// input_parameter_fitting_class.cpp:599-610).
struct EnergyParams {
    Params ct;
    std::string input_dir;               // "#Input directory path" / -ip
    std::string output_dir;              // "#Output directory path" / -op
    std::string output_base;             // "#Output file base name"
    int start_angle = -1, end_angle = -1;    // "#Angle number range" / -ar
    int start_energy = -1, end_energy = -1;  // "#Energy number range" / -enr
    std::vector<std::string> names;      // "#Processing parameter name" / -fpn
    int image_x = 2048, image_y = 2048;  // "#Image pixel size"
};

// Reads both key sets of `path`; false (with a message) when the file cannot
// be opened or a value does not parse.
bool parse_energy_inp(const std::string &path, EnergyParams &p, std::string *err);

// EnumTagString (share source/CTXAFS_analysis_share.cpp:85-96): 3-digit
// zero-padded energy number.
std::string energy_tag(int energy);

// The projection-file prefix in `dir` (Reslice_CTreconst/main.cpp:110-131):
// the first entry, in name order, that contains "0001.raw", without its last
// 8 characters.
bool find_projection_base(const std::string &dir, std::string &base, std::string *err);

// Runs reslice_CTreconst_ocl (Reslice_CTreconst_ocl.cpp:26-307): for every
// energy (or parameter name) the mu-t projections are read, resliced into
// sinograms on the device and reconstructed there, and the layers written to
// <output_dir>/<energy>/<output_base>NNNN.raw.  Returns 0 on success.
int run_energies(const EnergyParams &p, std::ostream &log);

}  // namespace ctrhost

#endif
