// shim_check.cpp -- drives include/openholo_hip.hpp the way a C++ caller of
// the reference API would, and prints what it got as `key value...` lines
// for tests/test_cpp_shim.py and tests/test_gpu_cpp_shim.py to compare.
//   shim_check host <golden dir> <tmp dir>   loaders, spec, config, BMP (no GPU)
//   shim_check gpu  <golden dir> <tmp dir>   generators through the shim
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "openholo_hip.hpp"

using namespace openholo;

static void print_floats(const char* key, const float* v, size_t n) {
    std::printf("%s", key);
    for (size_t i = 0; i < n; ++i) std::printf(" %.9g", v[i]);
    std::printf("\n");
}

static int host(const std::string& golden, const std::string& tmp) {
    const std::string dat = golden + "/keti_TestPointCloud.dat";
    OphPointCloud* raw = nullptr;
    const int n1 = ophLoadPointCloud(const_cast<char*>(dat.c_str()), raw);
    std::vector<OphPointCloud> vec;
    const int n2 = ophLoadPointCloud(dat, vec);
    std::vector<float> xyz, ph, amp;
    const int n3 = ophLoadPointCloud(dat, xyz, ph, amp);
    std::printf("points %d %d %d\n", n1, n2, n3);
    const float first[5] = {raw[0].x, raw[0].y, raw[0].z, raw[0].amplitude, raw[0].phase};
    print_floats("first", first, 5);
    const float last[5] = {vec.back().x, vec.back().y, vec.back().z, amp.back(), ph.back()};
    print_floats("last", last, 5);
    std::free(raw);
    std::printf("missing %d\n", ophLoadPointCloud(tmp + "/none.dat", vec));

    OphSpec spec;
    const bool ok = ophLoadSpecConfig(golden + "/keti_TestSpec.config", spec);
    const float sv[14] = {spec.pointCloudScaleX, spec.pointCloudScaleY, spec.pointCloudScaleZ, spec.offsetDepth,
                          spec.samplingPitchX, spec.samplingPitchY, (float)spec.nx, (float)spec.ny,
                          spec.filterXwidth, spec.filterYwidth, spec.focalLengthLensIn, spec.lambda,
                          spec.tiltAngleX, spec.tiltAngleY};
    std::printf("spec_ok %d filter %s\n", (int)ok, spec.filterShapeFlag);
    print_floats("spec", sv, 14);
    ophSetWaveLength(spec, 633e-9f);
    ophSetImageSize(spec, 7, 9);
    std::printf("setters %.9g %d %d\n", spec.lambda, spec.nx, spec.ny);
    std::free(spec.filterShapeFlag);

    HologramGenerator g;
    std::printf("config_ok %d\n", (int)g.readConfig(golden + "/dm_config_openholo.txt"));
    const oph_dm_config& c = g.config();
    std::printf("config %d %d %.12g %.12g %d %d\n", c.pn[0], c.pn[1], c.pp[0], c.lambda, c.num_of_depth,
                c.n_render_depth);

    unsigned char px[64 * 4];
    for (int i = 0; i < 64 * 4; ++i) px[i] = (unsigned char)(i * 7);
    ophSaveFileBmp(px, 64, 4, const_cast<char*>((tmp + "/shim").c_str()));
    std::printf("bmp_written %s\n", (tmp + "/shim.bmp").c_str());

    FILE* f = std::fopen((tmp + "/pts.xyz").c_str(), "w");
    std::fprintf(f, "1e-3 2e-3 -0.1\n-1e-3 0 -0.1005\n");
    std::fclose(f);
    OPHWRP w;
    const int rc = w.readxyz(tmp + "/pts.xyz");
    std::printf("xyz %d %d %.9g %.9g\n", rc, w.obj_num, w.vec_xyz[2], w.vec_xyz[5]);
    return 0;
}

static int run_gpu(const std::string& golden, const std::string& tmp) {
    // KETI fringe of the shipped point cloud on a small grid, accumulated
    // into a zeroed buffer as the reference does
    std::vector<OphPointCloud> pc;
    ophLoadPointCloud(golden + "/keti_TestPointCloud.dat", pc);
    OphSpec spec;
    ophLoadSpecConfig(golden + "/keti_TestSpec.config", spec);
    ophSetImageSize(spec, 96, 64);
    std::vector<float> dst(96 * 64, 0.f);
    const double t = ophGenCghPointCloud(pc, spec, dst.data());
    std::printf("fringe_time %.6g\n", t);
    FILE* f = std::fopen((tmp + "/fringe.f32").c_str(), "wb");
    std::fwrite(dst.data(), 4, dst.size(), f);
    std::fclose(f);
    std::vector<float> xyz, ph, amp;
    ophLoadPointCloud(golden + "/keti_TestPointCloud.dat", xyz, ph, amp);
    std::vector<float> dst2(96 * 64, 123.f);
    gpu::ophGenCghPointCloud_hip(xyz, amp, spec, dst2.data());
    double diff = 0;
    for (size_t i = 0; i < dst.size(); ++i) diff = std::fmax(diff, std::fabs(dst[i] - dst2[i]));
    std::printf("gpu_entry_maxdiff %.9g\n", diff);
    std::free(spec.filterShapeFlag);

    std::vector<unsigned char> u8(96 * 64);
    ophNormalize(dst.data(), u8.data(), 96, 64);
    f = std::fopen((tmp + "/fringe.u8").c_str(), "wb");
    std::fwrite(u8.data(), 1, u8.size(), f);
    std::fclose(f);

    // WRP: two points, 64^2 plane, then propagation; and a non-power-of-two plane
    f = std::fopen((tmp + "/pts.xyz").c_str(), "w");
    std::fprintf(f, "1e-4 2e-4 -1e-3\n-1e-4 0 -1.5e-3\n");
    std::fclose(f);
    for (int n : {64, 50}) {
        OPHWRP w;
        w.readxyz(tmp + "/pts.xyz");
        w.Setparameter(532e-9f, 8e-6f, (float)n);
        w.SingleWRP(0.0f);
        f = std::fopen((tmp + "/wrp" + std::to_string(n) + ".c64").c_str(), "wb");
        std::fwrite(w.field().data(), 8, w.field().size(), f);
        std::fclose(f);
        const int rc = w.Fresnelpropagation(0.01f);
        f = std::fopen((tmp + "/prop" + std::to_string(n) + ".c64").c_str(), "wb");
        std::fwrite(w.field().data(), 8, w.field().size(), f);
        std::fclose(f);
        std::printf("wrp %d %d %d\n", n, rc, w.SavetoImage(const_cast<char*>((tmp + "/arg" + std::to_string(n) + ".bmp").c_str())));
    }

    // depth map end to end from a config in tmp (written by the test)
    HologramGenerator g;
    if (!g.readConfig(tmp + "/dm_config.txt")) return 1;
    g.initialize();
    g.GenerateHologram();
    f = std::fopen((tmp + "/dm_fringe.f64").c_str(), "wb");
    std::fwrite(g.fringe().data(), 8, g.fringe().size(), f);
    std::fclose(f);
    g.ReconstructImage();
    f = std::fopen((tmp + "/dm_sim.f64").c_str(), "wb");
    std::fwrite(g.simulation().data(), 8, g.simulation().size(), f);
    std::fclose(f);
    std::printf("dm_error '%s'\n", g.last_error().c_str());
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: shim_check host|gpu <golden dir> <tmp dir>\n");
        return 2;
    }
    const std::string mode = argv[1];
    return mode == "host" ? host(argv[2], argv[3]) : run_gpu(argv[2], argv[3]);
}
