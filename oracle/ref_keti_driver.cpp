// ref_keti_driver.cpp -- C entry points over the reference KETI point-cloud
// sources, compiled from where they lie in the reference tree (see Makefile).
//
// TEST INFRASTRUCTURE ONLY: the library built from this file goes to
// oracle/_ref/ and is loaded by the fixture script under tests/golden/ and by
// the CPU tests, never by the product.  It calls the reference functions
// unchanged; it restates nothing.
//   ophLoadPointCloud(char*, OphPointCloud*&)   ophgen_pointcloud.cpp:13-47
//   ophLoadSpecConfig(char*, OphSpec*)          ophgen_pointcloud.cpp:112-158
//   ophGenCghPointCloud(const OphPointCloud*,…) ophgen_pointcloud.cpp:195-249
//   ophNormalize(float*, uchar*, nx, ny)        oph_common.cpp:92-116
//   ophSaveFileBmp(uchar*, w, h, char*)         oph_common.cpp:49-90
#include "ophgen_pointcloud.h"
#include "oph_common.h"
#include <cstdlib>
#include <cstring>

extern "C" {

// Loads a .dat point cloud through the reference loader; copies up to cap
// points into out (x, y, z, amplitude, phase per point).  Returns the count
// or -1, as the reference.
int ref_keti_load(const char* path, float* out, int cap) {
    openholo::OphPointCloud* model = nullptr;
    int n = openholo::ophLoadPointCloud(const_cast<char*>(path), model);
    if (n < 0) return n;
    for (int i = 0; i < n && i < cap; ++i) {
        out[5 * i + 0] = model[i].x;
        out[5 * i + 1] = model[i].y;
        out[5 * i + 2] = model[i].z;
        out[5 * i + 3] = model[i].amplitude;
        out[5 * i + 4] = model[i].phase;
    }
    std::free(model);
    return n;
}

// Parses a 17-line spec through the reference parser into 13 floats in the
// order scale xyz, offsetDepth, pitch xy, nx, ny (as floats), lambda,
// tilt xy, filter widths xy.  Returns 1 on success, 0 on failure.
int ref_keti_spec(const char* path, float* out) {
    openholo::OphSpec s;
    std::memset(&s, 0, sizeof(s));
    if (!openholo::ophLoadSpecConfig(const_cast<char*>(path), &s)) return 0;
    const float v[13] = {s.pointCloudScaleX, s.pointCloudScaleY, s.pointCloudScaleZ,
                         s.offsetDepth, s.samplingPitchX, s.samplingPitchY,
                         (float)s.nx, (float)s.ny, s.lambda, s.tiltAngleX,
                         s.tiltAngleY, s.filterXwidth, s.filterYwidth};
    std::memcpy(out, v, sizeof(v));
    return 1;
}

// Fringe of n points (pts as ref_keti_load's layout) for the given spec,
// accumulated into dst (nx*ny floats) by the reference generator.
void ref_keti_fringe(const float* pts, int n, const float* spec13, float* dst) {
    openholo::OphSpec s;
    std::memset(&s, 0, sizeof(s));
    s.pointCloudScaleX = spec13[0]; s.pointCloudScaleY = spec13[1];
    s.pointCloudScaleZ = spec13[2]; s.offsetDepth = spec13[3];
    s.samplingPitchX = spec13[4]; s.samplingPitchY = spec13[5];
    s.nx = (int)spec13[6]; s.ny = (int)spec13[7]; s.lambda = spec13[8];
    s.tiltAngleX = spec13[9]; s.tiltAngleY = spec13[10];
    openholo::OphPointCloud* m =
        static_cast<openholo::OphPointCloud*>(std::malloc(sizeof(openholo::OphPointCloud) * n));
    for (int i = 0; i < n; ++i) {
        m[i].x = pts[5 * i]; m[i].y = pts[5 * i + 1]; m[i].z = pts[5 * i + 2];
        m[i].amplitude = pts[5 * i + 3]; m[i].phase = pts[5 * i + 4];
    }
    openholo::ophGenCghPointCloud(m, n, s, dst);
    std::free(m);
}

void ref_keti_normalize(float* src, unsigned char* dst, int nx, int ny) {
    openholo::ophNormalize(src, dst, nx, ny);
}

// Writes `<name>.bmp` through the reference writer.
void ref_keti_save_bmp(unsigned char* pixels, int w, int h, const char* name) {
    openholo::ophSaveFileBmp(pixels, w, h, const_cast<char*>(name));
}

}  // extern "C"
