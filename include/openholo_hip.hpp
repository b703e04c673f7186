// openholo_hip.hpp -- the reference's C++ generator API, header-only, over
// the C ABI of oph_hip.h (link liboph_hip.so).
//
// A C++ caller of OpenHolo's generators includes this header instead of the
// reference's and keeps its calls:
//   * namespace openholo: the KETI point-cloud functions with the same names,
//     signatures, argument meaning and return values
//     (K/ophgen_pointcloud/ophgen_pointcloud.h:20-205, oph_common.h);
//   * openholo::HologramGenerator: the depth-map generator object
//     (DM/Hologram/HologramGenerator.h:122-150);
//   * openholo::OPHWRP: the WRP object (W/openwrp/WRP.h:17-60).
// Every generator runs on the GPU (gfx950).  The reference's CPU/GPU split
// (setMode, the _cuda entry point) therefore maps to one implementation; the
// GPU point-cloud entry point is named gpu::ophGenCghPointCloud_hip.
// Behaviour that differs from the reference, each documented at its function:
// no OpenMP races, the unzeroed KETI GPU output buffer is not reproduced,
// CWO's closed Fresnel convolution is replaced by the angular spectrum, and
// depth-map grids are powers of two.
#ifndef OPENHOLO_HIP_HPP
#define OPENHOLO_HIP_HPP

#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <sys/stat.h>
#include <dirent.h>

#include "oph_hip.h"

namespace openholo {

// ---- types (ophgen_pointcloud.h:20-70) ----
struct OphPointCloud {
    OphPointCloud() = default;
    OphPointCloud(const float x_, const float y_, const float z_) : x(x_), y(y_), z(z_) {}
    OphPointCloud(const float x_, const float y_, const float z_, const float amp, const float ph)
        : x(x_), y(y_), z(z_), amplitude(amp), phase(ph) {}
    float x = 0, y = 0, z = 0, amplitude = 0, phase = 0;
};
typedef OphPointCloud OphPC;

struct OphConfig {
    float pointCloudScaleX, pointCloudScaleY, pointCloudScaleZ;
    float offsetDepth;
    float samplingPitchX, samplingPitchY;
    int nx, ny;
    char* filterShapeFlag;  // text of field 8 (kept as the reference keeps it; owned by the caller)
    float filterXwidth, filterYwidth;
    float focalLengthLensIn, focalLengthLensOut, focalLengthLenEyepiece;
    float lambda;
    float tiltAngleX, tiltAngleY;
};
typedef OphConfig OphSpec;

namespace detail {

// One context per host thread (contexts are single-thread-affine), created on
// first use on device 0 and kept for the thread's lifetime, so the stateless
// reference functions keep their signatures.
struct ThreadContext {
    oph_ctx* ctx = nullptr;
    ~ThreadContext() {
        if (ctx) oph_destroy(ctx);
    }
};
inline oph_ctx* context() {
    static thread_local ThreadContext t;
    if (!t.ctx && oph_create(0, &t.ctx) != OPH_OK) t.ctx = nullptr;
    return t.ctx;
}

inline oph_spec to_spec(const OphSpec& c) {
    oph_spec s;
    std::memset(&s, 0, sizeof s);
    s.nx = c.nx;
    s.ny = c.ny;
    s.pitch_x = c.samplingPitchX;
    s.pitch_y = c.samplingPitchY;
    s.lambda = c.lambda;
    s.scale[0] = c.pointCloudScaleX;
    s.scale[1] = c.pointCloudScaleY;
    s.scale[2] = c.pointCloudScaleZ;
    s.offset_z = c.offsetDepth;
    s.tilt_deg[0] = c.tiltAngleX;
    s.tilt_deg[1] = c.tiltAngleY;
    return s;
}

// KETI fringe: dst += sum_n a_n cos(k r_n - k x sin tx - k y sin ty) (the
// point phase is not used by the reference), as float; returns seconds.
inline double fringe(const std::vector<double>& x, const std::vector<double>& y, const std::vector<double>& z,
                     const std::vector<double>& a, const OphSpec& config, float* dst) {
    const auto t0 = std::chrono::steady_clock::now();
    oph_ctx* ctx = context();
    if (!ctx || !dst) return -1.0;
    const oph_spec s = to_spec(config);
    oph_points p;
    p.x = x.data();
    p.y = y.data();
    p.z = z.data();
    p.amp = a.data();
    p.phase = nullptr;
    p.n = (int64_t)x.size();
    if (oph_pc_rs(ctx, &s, &p, dst, OPH_OUT_FRINGE | OPH_ACCUMULATE) != OPH_OK) return -1.0;
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

inline bool keti_points(const char* path, std::vector<double>& v, int64_t& n) {
    if (oph_load_keti_points(path, nullptr, 0, &n) != OPH_OK) return false;
    v.assign((size_t)n * 5, 0.0);
    return oph_load_keti_points(path, v.data(), n, &n) == OPH_OK;
}

inline int next_pow2(int n) {
    int p = 16;
    while (p < n) p <<= 1;
    return p;
}

inline bool make_dir(const std::string& d) { return mkdir(d.c_str(), 0755) == 0 || errno == EEXIST; }

// first file name (sorted) in dir starting with prefix, as QDir::entryList
// with the name filter "prefix*.*"
inline std::string first_match(const std::string& dir, const std::string& prefix) {
    DIR* d = opendir(dir.c_str());
    if (!d) return std::string();
    std::vector<std::string> names;
    while (dirent* e = readdir(d)) {
        const std::string n = e->d_name;
        if (n.compare(0, prefix.size(), prefix) == 0 && n.find('.', prefix.size()) != std::string::npos)
            names.push_back(n);
    }
    closedir(d);
    if (names.empty()) return std::string();
    std::sort(names.begin(), names.end());
    return dir + "/" + names[0];
}

}  // namespace detail

// ---- point-cloud loaders (ophgen_pointcloud.cpp:13-109) ----
// The point array is malloc'ed; the caller frees it.  Returns the point
// count, or -1 on a missing file or an index mismatch.
inline int ophLoadPointCloud(char* inputData, OphPointCloud*& model) {
    std::vector<double> v;
    int64_t n = 0;
    if (!detail::keti_points(inputData, v, n)) return -1;
    model = static_cast<OphPointCloud*>(std::malloc(sizeof(OphPointCloud) * (n > 0 ? n : 1)));
    for (int64_t i = 0; i < n; ++i)
        model[i] = OphPointCloud((float)v[5 * i], (float)v[5 * i + 1], (float)v[5 * i + 2], (float)v[5 * i + 4],
                                 (float)v[5 * i + 3]);
    return (int)n;
}

inline int ophLoadPointCloud(const std::string inputData, std::vector<OphPointCloud>& model) {
    std::vector<double> v;
    int64_t n = 0;
    if (!detail::keti_points(inputData.c_str(), v, n)) return -1;
    model.clear();
    for (int64_t i = 0; i < n; ++i)
        model.emplace_back((float)v[5 * i], (float)v[5 * i + 1], (float)v[5 * i + 2], (float)v[5 * i + 4],
                           (float)v[5 * i + 3]);
    return (int)n;
}

inline int ophLoadPointCloud(const std::string inputData, std::vector<float>& vertexArray,
                             std::vector<float>& phaseArray, std::vector<float>& amplitudeArray) {
    std::vector<double> v;
    int64_t n = 0;
    if (!detail::keti_points(inputData.c_str(), v, n)) return -1;
    vertexArray.clear();
    phaseArray.clear();
    amplitudeArray.clear();
    for (int64_t i = 0; i < n; ++i) {
        vertexArray.push_back((float)v[5 * i]);
        vertexArray.push_back((float)v[5 * i + 1]);
        vertexArray.push_back((float)v[5 * i + 2]);
        phaseArray.push_back((float)v[5 * i + 3]);
        amplitudeArray.push_back((float)v[5 * i + 4]);
    }
    return (int)n;
}

// ---- spec file and setters (ophgen_pointcloud.cpp:112-192) ----
// filterShapeFlag points to storage owned by the returned config's caller in
// the reference (a leaked malloc); here it is a malloc'ed copy as well.
inline bool ophLoadSpecConfig(char* inputData, OphSpec* config) {
    char v[17][64];
    if (!config || oph_keti_config_fields(inputData, v) != OPH_OK) return false;
    config->pointCloudScaleX = (float)std::atof(v[0]);
    config->pointCloudScaleY = (float)std::atof(v[1]);
    config->pointCloudScaleZ = (float)std::atof(v[2]);
    config->offsetDepth = (float)std::atof(v[3]);
    config->samplingPitchX = (float)std::atof(v[4]);
    config->samplingPitchY = (float)std::atof(v[5]);
    config->nx = std::atoi(v[6]);
    config->ny = std::atoi(v[7]);
    config->filterShapeFlag = static_cast<char*>(std::malloc(64));
    std::memcpy(config->filterShapeFlag, v[8], 64);
    config->filterXwidth = (float)std::atof(v[9]);
    config->filterYwidth = (float)std::atof(v[10]);
    config->focalLengthLensIn = (float)std::atof(v[11]);
    config->focalLengthLensOut = (float)std::atof(v[12]);
    config->focalLengthLenEyepiece = (float)std::atof(v[13]);
    config->lambda = (float)std::atof(v[14]);
    config->tiltAngleX = (float)std::atof(v[15]);
    config->tiltAngleY = (float)std::atof(v[16]);
    return true;
}
inline bool ophLoadSpecConfig(const std::string inputData, OphSpec& config) {
    return ophLoadSpecConfig(const_cast<char*>(inputData.c_str()), &config);
}

inline void ophSetScaleFactor(OphSpec& config, const float scaleX, const float scaleY, const float scaleZ) {
    config.pointCloudScaleX = scaleX;
    config.pointCloudScaleY = scaleY;
    config.pointCloudScaleZ = scaleZ;
}
inline void ophSetOffsetDepth(OphSpec& config, const float offsetDepth) { config.offsetDepth = offsetDepth; }
inline void ophSetSamplingPitch(OphSpec& config, const float pitchX, const float pitchY) {
    config.samplingPitchX = pitchX;
    config.samplingPitchY = pitchY;
}
inline void ophSetImageSize(OphSpec& config, const int n_x, const int n_y) {
    config.nx = n_x;
    config.ny = n_y;
}
inline void ophSetWaveLength(OphSpec& config, const float lambda) { config.lambda = lambda; }
inline void ophSetTiltAngle(OphSpec& config, const float tiltAngleX, const float tiltAngleY) {
    config.tiltAngleX = tiltAngleX;
    config.tiltAngleY = tiltAngleY;
}

// ---- fringe generation (ophgen_pointcloud.cpp:195-365) ----
// dst (nx*ny floats, zeroed by the caller as in the reference) accumulates
// the fringe.  Returns the elapsed seconds, or -1 on failure.
inline double ophGenCghPointCloud(const OphPointCloud* model, const int n_points, const OphSpec& config,
                                  float* dst) {
    std::vector<double> x(n_points), y(n_points), z(n_points), a(n_points);
    for (int i = 0; i < n_points; ++i) {
        x[i] = model[i].x;
        y[i] = model[i].y;
        z[i] = model[i].z;
        a[i] = model[i].amplitude;
    }
    return detail::fringe(x, y, z, a, config, dst);
}
inline double ophGenCghPointCloud(const std::vector<OphPointCloud>& model, const OphSpec& config, float* dst) {
    return ophGenCghPointCloud(model.data(), (int)model.size(), config, dst);
}
inline double ophGenCghPointCloud(const std::vector<float>& vertexArray, const std::vector<float>& amplitudeArray,
                                  const OphSpec& config, float* dst) {
    const size_t n = amplitudeArray.size();
    std::vector<double> x(n), y(n), z(n), a(n);
    for (size_t i = 0; i < n; ++i) {
        x[i] = vertexArray[3 * i];
        y[i] = vertexArray[3 * i + 1];
        z[i] = vertexArray[3 * i + 2];
        a[i] = amplitudeArray[i];
    }
    return detail::fringe(x, y, z, a, config, dst);
}

namespace gpu {
// gpu::ophGenCghPointCloud_cuda (ophgen_pointcloud_gpu.cpp:418-503) on HIP.
// The reference copies back a device buffer it never zeroed; here dst is
// overwritten with the fringe.
inline double ophGenCghPointCloud_hip(const std::vector<float>& vertexArray, const std::vector<float>& amplitudeArray,
                                      const OphSpec& config, float* dst) {
    if (!dst) return -1.0;
    std::memset(dst, 0, sizeof(float) * (size_t)config.nx * config.ny);
    return openholo::ophGenCghPointCloud(vertexArray, amplitudeArray, config, dst);
}
}  // namespace gpu

// ---- oph_common (oph_common.cpp:49-116) ----
inline void ophNormalize(float* src, unsigned char* dst, const int nx, const int ny) {
    oph_ctx* ctx = detail::context();
    if (ctx) oph_normalize_u8(ctx, src, dst, nx, ny, 0);
}
// writes "<file_name>.bmp" as the reference does
inline void ophSaveFileBmp(unsigned char* pixelbuffer, int pic_width, int pic_height, char* file_name) {
    oph_save_bmp8(pixelbuffer, pic_width, pic_height, (std::string(file_name) + ".bmp").c_str());
}

// ---- depth map (DM/Hologram/HologramGenerator.h:122-150) ----
// readConfig() reads config_openholo.txt from the working directory, as the
// reference; readConfig(path) reads another file.  GenerateHologram() runs
// every frame: it reads the intensity and depth BMPs (prefix match in the
// source folder, converted to 8-bit gray; they must already have the SLM
// pixel count -- the reference resamples with Qt), generates the fringe on a
// power-of-two transform grid (the scene itself when its sides are powers of
// two, as the reference's unpadded grid; otherwise the next power of two,
// zero-padded), and writes RESULT_FOLDER/RESULT_PREFIX<frame>.bmp as the
// reference's RGB888 image of width pn_x/3.  ReconstructImage() runs the eye
// model over the configured focus (or eye-y) steps and writes the
// simulation images.  last_error() names the failing step.
class HologramGenerator {
   public:
    HologramGenerator() = default;
    ~HologramGenerator() = default;

    void setMode(bool isCPU) { isCPU_ = isCPU; }  // both modes run the GPU path
    bool readConfig() { return readConfig("config_openholo.txt"); }
    bool readConfig(const std::string& path) {
        if (oph_dm_config_load(path.c_str(), &cfg_) != OPH_OK) return fail("cannot read " + path);
        have_cfg_ = true;
        return true;
    }
    void initialize() {
        fringe_.assign((size_t)cfg_.pn[0] * cfg_.pn[1], 0.0);
    }
    void GenerateHologram() {
        const int frames = cfg_.flag_static_image == 0 ? cfg_.number_of_frame : 1;
        for (int f = 0; f < frames; ++f) {
            if (!generate_frame(f)) continue;
            write_result(f);
        }
    }
    void ReconstructImage() {
        if (fringe_.empty()) {
            fail("no hologram data");
            return;
        }
        const int n = cfg_.sim_step_num;
        for (int v = (n > 0 ? 1 : 0); v <= (n > 0 ? n : 0); ++v) {
            double val = 0.0;
            if (n > 1) val = std::fmin(cfg_.sim_from, cfg_.sim_to) + (v - 1) * (cfg_.sim_to - cfg_.sim_from) / (n - 1);
            else if (n == 1) val = (cfg_.sim_to + cfg_.sim_from) / 2.0;
            oph_eye_spec e;
            e.pitch_x = cfg_.pp[0];
            e.pitch_y = cfg_.pp[1];
            e.lambda = cfg_.lambda;
            e.field_lens = cfg_.field_lens;
            e.eye_length = cfg_.eye_length;
            e.eye_pupil_diameter = cfg_.eye_pupil_diameter;
            e.eye_center_x = cfg_.eye_center_xy[0];
            e.eye_center_y = cfg_.eye_center_xy[1];
            e.focus_distance = cfg_.focus_distance;
            if (n > 0) {
                if (cfg_.sim_type == 0) e.focus_distance = val;
                else e.eye_center_y = val;
            }
            std::vector<double> sim(fringe_.size());
            if (!ok(oph_dm_reconstruct(ctx(), &e, fringe_.data(), cfg_.pn[0], cfg_.pn[1], sim.data(), 0),
                    "reconstruction"))
                return;
            sim_ = sim;
            char name[64];
            std::snprintf(name, sizeof name, "%d_%s%ld", n > 0 ? v : 0, cfg_.sim_type == 0 ? "FOCUS_" : "EYE_Y_",
                          (long)std::lround(val * 1000));
            write_rgb(sim_, std::string(cfg_.result_folder) + "/" + cfg_.sim_prefix + "_" + cfg_.result_prefix +
                                name + ".bmp");
        }
    }

    // results of the last frame / reconstruction
    const std::vector<double>& fringe() const { return fringe_; }
    const std::vector<double>& simulation() const { return sim_; }
    const oph_dm_config& config() const { return cfg_; }
    const std::string& last_error() const { return err_; }

   private:
    oph_dm_config cfg_{};
    bool have_cfg_ = false, isCPU_ = false;
    std::vector<double> fringe_, sim_;
    std::string err_;

    oph_ctx* ctx() { return detail::context(); }
    bool fail(const std::string& m) {
        err_ = m;
        std::fprintf(stderr, "HologramGenerator: %s\n", m.c_str());
        return false;
    }
    bool ok(int rc, const char* what) {
        if (rc == OPH_OK) return true;
        return fail(std::string(what) + ": " + (ctx() ? oph_last_error(ctx()) : "no device"));
    }

    bool generate_frame(int ftr) {
        if (!have_cfg_) return fail("no configuration");
        std::string dir = cfg_.source_folder;
        if (cfg_.flag_static_image == 0) {
            char num[32];
            if (cfg_.number_of_digit > 0)
                std::snprintf(num, sizeof num, "%0*d", cfg_.number_of_digit, ftr + cfg_.start_of_frame_numbering);
            else
                std::snprintf(num, sizeof num, "%d", ftr + cfg_.start_of_frame_numbering);
            dir += std::string("/") + num;
        }
        const std::string fi = detail::first_match(dir, cfg_.image_prefix);
        const std::string fd = detail::first_match(dir, cfg_.depth_prefix);
        if (fi.empty() || fd.empty()) return fail("source image or depth map missing in " + dir);
        const int w = cfg_.pn[0], h = cfg_.pn[1];
        std::vector<unsigned char> img((size_t)w * h), dep((size_t)w * h);
        int iw = 0, ih = 0, dw = 0, dh = 0;
        if (oph_load_bmp_gray8(fi.c_str(), &iw, &ih, nullptr, 0) != OPH_OK ||
            oph_load_bmp_gray8(fd.c_str(), &dw, &dh, nullptr, 0) != OPH_OK)
            return fail("cannot read " + fi + " / " + fd);
        if (iw != w || ih != h || dw != w || dh != h) return fail("input size differs from the SLM pixel count");
        oph_load_bmp_gray8(fi.c_str(), &iw, &ih, img.data(), (int64_t)img.size());
        oph_load_bmp_gray8(fd.c_str(), &dw, &dh, dep.data(), (int64_t)dep.size());
        oph_dm_spec s;
        std::memset(&s, 0, sizeof s);
        s.w = w;
        s.h = h;
        s.Nx = detail::next_pow2(w);
        s.Ny = detail::next_pow2(h);
        s.pitch_x = cfg_.pp[0];
        s.pitch_y = cfg_.pp[1];
        s.near_m = cfg_.near_depthmap;
        s.far_m = cfg_.far_depthmap;
        s.num_depth = cfg_.num_of_depth;
        s.quant_mode = cfg_.flag_change_depth_quantization ? 1 : 0;
        s.default_q = cfg_.default_depth_quantization;
        s.field_lens = cfg_.transform_method == 0 ? cfg_.field_lens : 0.0;
        s.rho_re = 1.0;
        s.rho_im = 0.0;
        const double lam = cfg_.lambda;
        fringe_.assign((size_t)w * h, 0.0);
        std::vector<double> part((size_t)w * h);
        // the rendered levels, in contiguous runs (the encode is linear, so
        // fringes of runs add)
        const int* r = cfg_.render_depth;
        for (int i = 0; i < cfg_.n_render_depth;) {
            int j = i;
            while (j + 1 < cfg_.n_render_depth && r[j + 1] == r[j] + 1) ++j;
            s.layer_lo = r[i];
            s.layer_hi = r[j];
            if (!ok(oph_depthmap(ctx(), &s, 1, &lam, img.data(), dep.data(), part.data(), OPH_DM_OUT_FRINGE),
                    "depth-map hologram"))
                return false;
            for (size_t k = 0; k < part.size(); ++k) fringe_[k] += part[k];
            i = j + 1;
        }
        return true;
    }

    void write_rgb(const std::vector<double>& v, const std::string& path) {
        std::vector<unsigned char> u8(v.size());
        if (!ok(oph_scale_u8(ctx(), v.data(), u8.data(), (int64_t)v.size(), 0), "encoding")) return;
        detail::make_dir(cfg_.result_folder);
        if (oph_save_bmp24(u8.data(), cfg_.pn[0] / 3, cfg_.pn[1], path.c_str()) != OPH_OK) fail("cannot write " + path);
    }
    void write_result(int ftr) {
        write_rgb(fringe_, std::string(cfg_.result_folder) + "/" + cfg_.result_prefix + std::to_string(ftr) + ".bmp");
    }
};

// ---- WRP (W/openwrp/WRP.h:17-60) ----
// readxyz / Setparameter / SingleWRP / Fresnelpropagation / SavetoImage with
// the reference's arguments.  SingleWRP uses the reference's window
// (WRP.cpp:111, mode 1) and no point amplitude (the reference's points carry
// none).  Fresnelpropagation propagates by the angular spectrum with at
// least 2x zero padding (the plane is embedded in a power-of-two grid when
// its side is not one); CWO's Fresnel convolution is closed and absent.
// SavetoImage writes the phase angle mapped linearly from [-pi, pi] to
// 0..255 as a gray 24-bit BMP (CWO_SAVE_AS_ARG; rows padded, any width).
class OPHWRP {
   public:
    std::vector<double> vec_xyz;
    int obj_num = 0;

    int readxyz(std::string filename) {
        int64_t n = 0;
        if (oph_load_xyz(filename.c_str(), nullptr, 0, &n) != OPH_OK) return -1;
        std::vector<double> v((size_t)n * 3);
        oph_load_xyz(filename.c_str(), v.data(), n, &n);
        vec_xyz.insert(vec_xyz.end(), v.begin(), v.end());  // the reference appends to vec_xyz
        obj_num = (int)(vec_xyz.size() / 3);
        return 0;
    }
    int Setparameter(float m_wavelength, float m_pixelpitch, float m_resolution) {
        lambda_ = m_wavelength;
        pitch_ = m_pixelpitch;
        n_ = (int)m_resolution;
        field_.assign((size_t)n_ * n_, std::complex<float>(0.f, 0.f));
        return 0;
    }
    void SingleWRP(float z) {
        const size_t n = (size_t)obj_num;
        std::vector<double> x(n), y(n), zz(n);
        for (size_t i = 0; i < n; ++i) {
            x[i] = vec_xyz[3 * i];
            y[i] = vec_xyz[3 * i + 1];
            zz[i] = vec_xyz[3 * i + 2];
        }
        oph_points p;
        p.x = x.data();
        p.y = y.data();
        p.z = zz.data();
        p.amp = nullptr;
        p.phase = nullptr;
        p.n = (int64_t)n;
        // the reference adds into the plane (AddPixel); a failed call leaves
        // the plane as it was and is reported on stderr and by status()
        status_ = oph_wrp(detail::context(), n_, n_, pitch_, lambda_, z, 1, &p,
                          reinterpret_cast<oph_c64*>(field_.data()), OPH_ACCUMULATE);
        if (status_ != OPH_OK) std::fprintf(stderr, "SingleWRP: %s\n", oph_last_error(detail::context()));
    }
    // OPH_OK, or the error code of the last SingleWRP call
    int status() const { return status_; }
    int Fresnelpropagation(float z) {
        const int m = detail::next_pow2(n_);
        if (m == n_) {
            return oph_asm_propagate(detail::context(), reinterpret_cast<oph_c64*>(field_.data()), n_, n_, pitch_,
                                     pitch_, lambda_, z, 2, 0, 0) == OPH_OK ? 0 : -1;
        }
        std::vector<std::complex<float>> big((size_t)m * m);
        const int o = m / 2 - n_ / 2;
        for (int r = 0; r < n_; ++r)
            for (int c = 0; c < n_; ++c) big[(size_t)(r + o) * m + c + o] = field_[(size_t)r * n_ + c];
        if (oph_asm_propagate(detail::context(), reinterpret_cast<oph_c64*>(big.data()), m, m, pitch_, pitch_,
                              lambda_, z, 2, 0, 0) != OPH_OK)
            return -1;
        for (int r = 0; r < n_; ++r)
            for (int c = 0; c < n_; ++c) field_[(size_t)r * n_ + c] = big[(size_t)(r + o) * m + c + o];
        return 0;
    }
    int SavetoImage(char* filename) {
        std::vector<unsigned char> rgb(field_.size() * 3);
        for (size_t i = 0; i < field_.size(); ++i) {
            const double a = std::arg(std::complex<double>(field_[i].real(), field_[i].imag()));
            rgb[3 * i] = rgb[3 * i + 1] = rgb[3 * i + 2] = (unsigned char)std::lround((a + M_PI) / (2.0 * M_PI) * 255.0);
        }
        return oph_save_bmp24(rgb.data(), n_, n_, filename) == OPH_OK ? 0 : -1;
    }

    // the complex plane (n x n, row index = y)
    std::vector<std::complex<float>>& field() { return field_; }

   private:
    double lambda_ = 0, pitch_ = 0;
    int n_ = 0;
    int status_ = OPH_OK;
    std::vector<std::complex<float>> field_;
};

}  // namespace openholo

#endif  // OPENHOLO_HIP_HPP
