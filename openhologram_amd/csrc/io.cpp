// io.cpp -- file formats around the hologram path: BMP output and input,
// the WRP point files, and the depth-map configuration file.  Host code only.
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "oph_hip.h"

namespace {

bool read_file(const char* path, std::string& out) {
    if (!path) return false;
    FILE* f = fopen(path, "rb");
    if (!f) return false;
    out.clear();
    char buf[1 << 16];
    size_t k;
    while ((k = fread(buf, 1, sizeof buf, f)) > 0) out.append(buf, k);
    fclose(f);
    return true;
}

void put16(std::vector<unsigned char>& b, unsigned v) {
    b.push_back(v & 0xFF);
    b.push_back((v >> 8) & 0xFF);
}
void put32(std::vector<unsigned char>& b, uint32_t v) {
    for (int i = 0; i < 4; ++i) b.push_back((v >> (8 * i)) & 0xFF);
}
uint32_t get32(const unsigned char* p) { return p[0] | (p[1] << 8) | (p[2] << 16) | ((uint32_t)p[3] << 24); }
unsigned get16(const unsigned char* p) { return p[0] | (p[1] << 8); }

// BITMAPFILEHEADER (14 bytes) + BITMAPINFOHEADER (40 bytes)
void bmp_headers(std::vector<unsigned char>& b, uint32_t file_size, uint32_t pixel_offset, int width, int height,
                 unsigned bpp, uint32_t image_size, uint32_t palette) {
    b.push_back('B');
    b.push_back('M');
    put32(b, file_size);
    put32(b, 0);
    put32(b, pixel_offset);
    put32(b, 40);
    put32(b, (uint32_t)width);
    put32(b, (uint32_t)height);
    put16(b, 1);       // planes
    put16(b, bpp);
    put32(b, 0);       // no compression
    put32(b, image_size);
    put32(b, 0x130B);  // pixels per metre, y then x: the reference header value
    put32(b, 0x130B);
    put32(b, palette);
    put32(b, 0);
}

bool write_all(const char* path, const std::vector<unsigned char>& head, const unsigned char* body, size_t n) {
    FILE* f = fopen(path, "wb");
    if (!f) return false;
    bool ok = fwrite(head.data(), 1, head.size(), f) == head.size();
    if (n) ok = ok && fwrite(body, 1, n, f) == n;
    return fclose(f) == 0 && ok;
}

// whitespace-token reader with the semantics of `std::istream >>` followed by
// `getline` (the reference reads its config that way)
struct Reader {
    const std::string& s;
    size_t pos = 0;
    explicit Reader(const std::string& text) : s(text) {}
    bool token(std::string& out) {
        while (pos < s.size() && isspace((unsigned char)s[pos])) ++pos;
        if (pos >= s.size()) return false;
        const size_t b = pos;
        while (pos < s.size() && !isspace((unsigned char)s[pos])) ++pos;
        out.assign(s, b, pos - b);
        return true;
    }
    // rest of the current line, newline consumed
    std::string line() {
        const size_t b = pos;
        while (pos < s.size() && s[pos] != '\n') ++pos;
        std::string r(s, b, pos - b);
        if (pos < s.size()) ++pos;
        if (!r.empty() && r.back() == '\r') r.pop_back();
        return r;
    }
    void skip_lines(int n) {
        for (int i = 0; i < n; ++i) line();
    }
};

// `>> double` then getline: the numeric prefix of the next token
bool field_double(Reader& r, double& v) {
    std::string t;
    if (!r.token(t)) return false;
    char* end = nullptr;
    errno = 0;
    v = strtod(t.c_str(), &end);
    r.line();
    return end != t.c_str() && errno == 0;
}
bool field_int(Reader& r, int& v) {
    std::string t;
    if (!r.token(t)) return false;
    char* end = nullptr;
    const long x = strtol(t.c_str(), &end, 10);
    r.line();
    v = (int)x;
    return end != t.c_str();
}
bool field_string(Reader& r, char* dst, size_t cap) {
    std::string t;
    if (!r.token(t) || t.size() + 1 > cap) return false;
    memcpy(dst, t.c_str(), t.size() + 1);
    r.line();
    return true;
}

// Arithmetic expression in double: numbers (with exponent), + - * /,
// parentheses and unary signs -- the subset of script the reference's
// config evaluates for its pixel-number and pitch fields.
struct Expr {
    const char* p;
    bool ok = true;
    void ws() {
        while (*p && isspace((unsigned char)*p)) ++p;
    }
    double primary() {
        ws();
        if (*p == '(') {
            ++p;
            const double v = sum();
            ws();
            if (*p == ')') ++p;
            else ok = false;
            return v;
        }
        if (*p == '+' || *p == '-') {
            const bool neg = *p == '-';
            ++p;
            const double v = primary();
            return neg ? -v : v;
        }
        char* end = nullptr;
        const double v = strtod(p, &end);
        if (end == p) {
            ok = false;
            return 0.0;
        }
        p = end;
        return v;
    }
    double product() {
        double v = primary();
        for (;;) {
            ws();
            if (*p == '*') {
                ++p;
                v *= primary();
            } else if (*p == '/') {
                ++p;
                v /= primary();
            } else {
                return v;
            }
        }
    }
    double sum() {
        double v = product();
        for (;;) {
            ws();
            if (*p == '+') {
                ++p;
                v += product();
            } else if (*p == '-') {
                ++p;
                v -= product();
            } else {
                return v;
            }
        }
    }
};

bool eval_expr(const std::string& text, double& v) {
    Expr e{text.c_str()};
    v = e.sum();
    e.ws();
    return e.ok && *e.p == 0 && std::isfinite(v);
}

bool field_expr(Reader& r, double& v) {
    std::string t;
    if (!r.token(t)) return false;
    r.line();
    return eval_expr(t, v);
}

// RENDER_DEPTH: text before "//", trimmed; "a:b" -> a..b, else every number
bool render_depth(const std::string& line, oph_dm_config* c) {
    std::string src = line.substr(0, line.find("//"));
    c->n_render_depth = 0;
    auto add = [&](long v) {
        if (c->n_render_depth < OPH_DM_MAX_RENDER) c->render_depth[c->n_render_depth++] = (int)v;
    };
    const size_t colon = src.find(':');
    if (colon != std::string::npos) {
        // first "digits:digits" occurrence
        for (size_t i = 0; i < src.size(); ++i) {
            if (!isdigit((unsigned char)src[i]) || (i > 0 && isdigit((unsigned char)src[i - 1]))) continue;
            size_t j = i;
            while (j < src.size() && isdigit((unsigned char)src[j])) ++j;
            if (j < src.size() && src[j] == ':' && j + 1 < src.size() && isdigit((unsigned char)src[j + 1])) {
                const long a = strtol(src.c_str() + i, nullptr, 10);
                const long b = strtol(src.c_str() + j + 1, nullptr, 10);
                for (long k = a; k <= b; ++k) add(k);
                break;
            }
        }
    } else {
        for (size_t i = 0; i < src.size();) {
            if (isdigit((unsigned char)src[i])) {
                char* end = nullptr;
                add(strtol(src.c_str() + i, &end, 10));
                i = end - src.c_str();
            } else {
                ++i;
            }
        }
    }
    return c->n_render_depth > 0;
}

bool has_ext(const char* path, const char* ext) {
    const char* dot = strrchr(path, '.');
    if (!dot) return false;
    ++dot;
    size_t n = strlen(ext);
    if (strlen(dot) != n) return false;
    for (size_t i = 0; i < n; ++i)
        if (tolower((unsigned char)dot[i]) != ext[i]) return false;
    return true;
}

}  // namespace

extern "C" int oph_save_bmp8(const unsigned char* pixels, int width, int height, const char* path) {
    if (!pixels || !path || width <= 0 || height <= 0) return OPH_EINVAL;
    const uint32_t bytes = (uint32_t)width * (uint32_t)height;
    const uint32_t head = 14 + 40 + 256 * 4;
    std::vector<unsigned char> b;
    b.reserve(head);
    bmp_headers(b, bytes + head, head, width, height, 8, bytes, 256);
    for (int i = 0; i < 256; ++i) {
        b.push_back((unsigned char)i);
        b.push_back((unsigned char)i);
        b.push_back((unsigned char)i);
        b.push_back(0);
    }
    return write_all(path, b, pixels, bytes) ? OPH_OK : OPH_EIO;
}

extern "C" int oph_save_bmp24(const unsigned char* rgb, int width, int height, const char* path) {
    if (!rgb || !path || width <= 0 || height <= 0) return OPH_EINVAL;
    const size_t stride = ((size_t)width * 3 + 3) & ~(size_t)3;
    std::vector<unsigned char> px(stride * height, 0);
    for (int y = 0; y < height; ++y) {
        const unsigned char* src = rgb + (size_t)(height - 1 - y) * width * 3;
        unsigned char* dst = px.data() + (size_t)y * stride;
        for (int x = 0; x < width; ++x) {
            dst[3 * x + 0] = src[3 * x + 2];
            dst[3 * x + 1] = src[3 * x + 1];
            dst[3 * x + 2] = src[3 * x + 0];
        }
    }
    std::vector<unsigned char> b;
    bmp_headers(b, (uint32_t)(54 + px.size()), 54, width, height, 24, (uint32_t)px.size(), 0);
    return write_all(path, b, px.data(), px.size()) ? OPH_OK : OPH_EIO;
}

extern "C" int oph_load_bmp_gray8(const char* path, int* width, int* height, unsigned char* dst, int64_t cap) {
    if (!width || !height) return OPH_EINVAL;
    std::string f;
    if (!read_file(path, f)) return OPH_EIO;
    const unsigned char* d = reinterpret_cast<const unsigned char*>(f.data());
    if (f.size() < 54 || d[0] != 'B' || d[1] != 'M') return OPH_EIO;
    const uint32_t off = get32(d + 10), dib = get32(d + 14);
    const int32_t w = (int32_t)get32(d + 18), hs = (int32_t)get32(d + 22);
    const unsigned bpp = get16(d + 28);
    const uint32_t comp = dib >= 40 ? get32(d + 30) : 0;
    if (w <= 0 || hs == 0 || (bpp != 8 && bpp != 24 && bpp != 32) || (comp != 0 && comp != 3)) return OPH_EIO;
    const bool top_down = hs < 0;
    const int h = top_down ? -hs : hs;
    const size_t stride = (((size_t)w * bpp + 31) / 32) * 4;
    if ((size_t)off + stride * (h - 1) + ((size_t)w * bpp + 7) / 8 > f.size()) return OPH_EIO;
    *width = w;
    *height = h;
    if (!dst) return OPH_OK;
    if (cap < (int64_t)w * h) return OPH_EINVAL;
    unsigned char pal[256];
    if (bpp == 8) {
        const uint32_t ncol0 = dib >= 40 ? get32(d + 46) : 0;
        const uint32_t ncol = ncol0 ? ncol0 : 256;
        const size_t pbase = 14 + dib;
        for (int i = 0; i < 256; ++i) {
            if ((uint32_t)i < ncol && pbase + 4 * i + 3 <= f.size()) {
                const unsigned b = d[pbase + 4 * i], g = d[pbase + 4 * i + 1], r = d[pbase + 4 * i + 2];
                pal[i] = (unsigned char)((r * 11 + g * 16 + b * 5) / 32);
            } else {
                pal[i] = 0;
            }
        }
    }
    for (int y = 0; y < h; ++y) {
        const unsigned char* row = d + off + stride * (top_down ? y : h - 1 - y);
        unsigned char* o = dst + (size_t)y * w;
        for (int x = 0; x < w; ++x) {
            if (bpp == 8) {
                o[x] = pal[row[x]];
            } else {
                const unsigned char* q = row + (size_t)x * (bpp / 8);
                o[x] = (unsigned char)((q[2] * 11 + q[1] * 16 + q[0] * 5) / 32);
            }
        }
    }
    return OPH_OK;
}

extern "C" int oph_load_xyz(const char* path, double* xyz, int64_t cap, int64_t* n) {
    if (!n) return OPH_EINVAL;
    std::string f;
    if (!read_file(path, f)) return OPH_EIO;
    std::vector<double> v;
    const char* p = f.c_str();
    const char* e = p + f.size();
    while (p < e) {
        while (p < e && isspace((unsigned char)*p)) ++p;
        if (p >= e) break;
        char* end = nullptr;
        const double x = strtod(p, &end);
        if (end == p) return OPH_EIO;  // a token that is not a number (std::stod throws)
        v.push_back(x);
        p = end;
        while (p < e && !isspace((unsigned char)*p)) ++p;  // rest of the token, as stod ignores it
    }
    const int64_t np = (int64_t)(v.size() / 3);
    *n = np;
    if (!xyz) return OPH_OK;
    const int64_t m = np < cap ? np : cap;
    memcpy(xyz, v.data(), sizeof(double) * 3 * (size_t)m);
    return OPH_OK;
}

extern "C" int oph_load_cwo_points(const char* path, double* xyz, int64_t cap, int64_t* n) {
    if (!n || !path) return OPH_EINVAL;
    const bool is_int = has_ext(path, "3d"), is_float = has_ext(path, "3df");
    if (!is_int && !is_float) return OPH_EINVAL;
    std::string f;
    if (!read_file(path, f)) return OPH_EIO;
    if (f.size() < 4) return OPH_EIO;
    const unsigned char* d = reinterpret_cast<const unsigned char*>(f.data());
    const int32_t num = (int32_t)get32(d);
    if (num < 0 || 4 + (size_t)num * 12 > f.size()) return OPH_EIO;
    *n = num;
    if (!xyz) return OPH_OK;
    const int64_t m = num < cap ? num : cap;
    for (int64_t i = 0; i < 3 * m; ++i) {
        const uint32_t w = get32(d + 4 + 4 * i);
        if (is_int) {
            xyz[i] = (double)(int32_t)w;
        } else {
            float fv;
            memcpy(&fv, &w, 4);
            xyz[i] = (double)fv;
        }
    }
    return OPH_OK;
}

extern "C" int oph_dm_config_load(const char* path, oph_dm_config* c) {
    if (!c) return OPH_EINVAL;
    std::string f;
    if (!read_file(path, f)) return OPH_EIO;
    memset(c, 0, sizeof *c);
    Reader r(f);
    bool ok = true;
    r.skip_lines(7);
    ok = ok && field_string(r, c->source_folder, sizeof c->source_folder);
    ok = ok && field_string(r, c->image_prefix, sizeof c->image_prefix);
    ok = ok && field_string(r, c->depth_prefix, sizeof c->depth_prefix);
    ok = ok && field_string(r, c->result_folder, sizeof c->result_folder);
    ok = ok && field_string(r, c->result_prefix, sizeof c->result_prefix);
    ok = ok && field_int(r, c->flag_static_image);
    ok = ok && field_int(r, c->start_of_frame_numbering);
    ok = ok && field_int(r, c->number_of_frame);
    ok = ok && field_int(r, c->number_of_digit);
    r.skip_lines(3);
    ok = ok && field_int(r, c->transform_method);
    ok = ok && field_int(r, c->propagation_method);
    ok = ok && field_int(r, c->encoding_method);
    r.skip_lines(3);
    ok = ok && field_double(r, c->field_lens);
    ok = ok && field_double(r, c->lambda);
    double v;
    ok = ok && field_expr(r, v);
    c->pn[0] = (int)v;
    ok = ok && field_expr(r, v);
    c->pn[1] = (int)v;
    ok = ok && field_expr(r, c->pp[0]);
    ok = ok && field_expr(r, c->pp[1]);
    if (!ok) return OPH_EIO;
    r.skip_lines(3);
    double near_m = 0, far_m = 0;
    ok = ok && field_double(r, near_m);
    ok = ok && field_double(r, far_m);
    c->near_depthmap = near_m < far_m ? near_m : far_m;
    c->far_depthmap = near_m < far_m ? far_m : near_m;
    ok = ok && field_int(r, c->flag_change_depth_quantization);
    ok = ok && field_int(r, c->default_depth_quantization);
    ok = ok && field_int(r, c->number_of_depth_quantization);
    if (!ok) return OPH_EIO;
    c->num_of_depth = c->flag_change_depth_quantization == 0 ? c->default_depth_quantization
                                                             : c->number_of_depth_quantization;
    if (!render_depth(r.line(), c)) return OPH_EIO;
    ok = ok && field_int(r, c->random_phase);
    r.skip_lines(3);
    ok = ok && field_string(r, c->sim_prefix, sizeof c->sim_prefix);
    ok = ok && field_double(r, c->test_pixel_number_scale);
    ok = ok && field_double(r, c->eye_length);
    ok = ok && field_double(r, c->eye_pupil_diameter);
    ok = ok && field_double(r, c->eye_center_xy[0]);
    ok = ok && field_double(r, c->eye_center_xy[1]);
    ok = ok && field_double(r, c->focus_distance);
    r.skip_lines(5);
    ok = ok && field_int(r, c->sim_type);
    ok = ok && field_double(r, c->sim_from);
    ok = ok && field_double(r, c->sim_to);
    ok = ok && field_int(r, c->sim_step_num);
    return ok ? OPH_OK : OPH_EIO;
}
