/*
 * film_io.cpp — the film writers of include/rt_film.h: --save-exr,
 * --save-img and --save-mat of the reference's command line
 * (src/main.cpp:216-258) and its output-name rule (src/main.cpp:70-81).
 * Host code; the tonemapper itself runs on the GPU (rt_film.hip).
 *
 * Films are height x width x 3 float RGB, row 0 at the top.
 */
#include <sys/stat.h>
#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rt_film.h"

namespace {

thread_local std::string g_film_error;

int fail(int status, const std::string& msg)
{
    g_film_error = msg;
    return status;
}

bool exists(const std::string& path)
{
    struct stat st;
    return ::stat(path.c_str(), &st) == 0;
}

struct Out {
    std::vector<unsigned char> b;
    void u8(unsigned v) { b.push_back((unsigned char)v); }
    void u16(unsigned v)
    {
        u8(v & 0xff);
        u8((v >> 8) & 0xff);
    }
    void u32(uint32_t v)
    {
        for (int k = 0; k < 4; ++k) u8((v >> (8 * k)) & 0xff);
    }
    void u32be(uint32_t v)
    {
        for (int k = 3; k >= 0; --k) u8((v >> (8 * k)) & 0xff);
    }
    void u64(uint64_t v)
    {
        for (int k = 0; k < 8; ++k) u8((unsigned)((v >> (8 * k)) & 0xff));
    }
    void f32(float f)
    {
        uint32_t u;
        std::memcpy(&u, &f, 4);
        u32(u);
    }
    void str0(const char* s)
    {
        while (*s) u8((unsigned char)*s++);
        u8(0);
    }
    void bytes(const void* p, size_t n)
    {
        const unsigned char* c = (const unsigned char*)p;
        b.insert(b.end(), c, c + n);
    }
};

int write_all(const char* path, const void* data, size_t n)
{
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return fail(RT_ERR_IO, std::string("cannot open ") + path + " for writing");
    size_t w = std::fwrite(data, 1, n, fp);
    int rc = std::fclose(fp);
    if (w != n || rc != 0) return fail(RT_ERR_IO, std::string("cannot write ") + path);
    return RT_OK;
}

bool ends_with(const std::string& s, const char* suf)
{
    size_t n = std::strlen(suf);
    if (s.size() < n) return false;
    for (size_t i = 0; i < n; ++i)
        if (std::tolower((unsigned char)s[s.size() - n + i]) != suf[i]) return false;
    return true;
}

// OpenEXR attribute: name, type, size, value
void exr_attr(Out& o, const char* name, const char* type, const Out& value)
{
    o.str0(name);
    o.str0(type);
    o.u32((uint32_t)value.b.size());
    o.bytes(value.b.data(), value.b.size());
}

}  // namespace

/* Radiance RGBE (.hdr) reader, the format OpenCV's imread decodes for the
 * reference's environment maps (src/environmentmap.cpp:20-33): a text header
 * ending in an empty line (FORMAT=32-bit_rle_rgbe), the resolution line
 * "-Y H +X W" (rows top to bottom), then H scanlines, each flat (4 bytes a
 * pixel) or run-length encoded per channel (marker 2 2 W/256 W%256; a count
 * above 128 repeats the next byte count-128 times, otherwise count literal
 * bytes follow).  A pixel (r, g, b, e) is (r, g, b) * 2^(e - 136), 0 when
 * e = 0. */
bool rt_read_rgbe(const std::string& path, int& w, int& h, std::vector<float>& rgb, std::string& err)
{
    FILE* fp = std::fopen(path.c_str(), "rb");
    if (!fp) {
        err = "cannot open " + path;
        return false;
    }
    std::vector<unsigned char> b;
    unsigned char buf[1 << 16];
    size_t r;
    while ((r = std::fread(buf, 1, sizeof buf, fp)) > 0) b.insert(b.end(), buf, buf + r);
    std::fclose(fp);
    size_t i = 0;
    auto line = [&](std::string& out) {
        out.clear();
        while (i < b.size() && b[i] != '\n') out += (char)b[i++];
        if (i < b.size()) ++i;
        return true;
    };
    std::string ln;
    line(ln);
    if (ln.rfind("#?", 0) != 0) {
        err = path + ": not a Radiance file";
        return false;
    }
    bool rgbe = false;
    for (;;) {
        if (i >= b.size()) {
            err = path + ": truncated header";
            return false;
        }
        line(ln);
        if (ln.empty()) break;
        if (ln == "FORMAT=32-bit_rle_rgbe") rgbe = true;
        if (ln.rfind("FORMAT=", 0) == 0 && !rgbe) {
            err = path + ": unsupported " + ln;
            return false;
        }
    }
    line(ln);
    char ya[3] = {0}, xa[3] = {0};
    if (std::sscanf(ln.c_str(), "%2s %d %2s %d", ya, &h, xa, &w) != 4 || std::string(ya) != "-Y" ||
        std::string(xa) != "+X" || w <= 0 || h <= 0 || (int64_t)w * h > (1ll << 28)) {
        err = path + ": unsupported resolution line \"" + ln + "\"";
        return false;
    }
    rgb.assign((size_t)w * h * 3, 0.f);
    std::vector<unsigned char> sl((size_t)w * 4);
    for (int y = 0; y < h; ++y) {
        if (i + 4 > b.size()) {
            err = path + ": truncated pixels";
            return false;
        }
        const bool rle = w >= 8 && w < 0x8000 && b[i] == 2 && b[i + 1] == 2 && !(b[i + 2] & 0x80);
        if (!rle) {
            if (i + sl.size() > b.size()) {
                err = path + ": truncated pixels";
                return false;
            }
            std::memcpy(sl.data(), &b[i], sl.size());
            i += sl.size();
        } else {
            if (((int)b[i + 2] << 8 | b[i + 3]) != w) {
                err = path + ": scanline width mismatch";
                return false;
            }
            i += 4;
            for (int ch = 0; ch < 4; ++ch) {
                int x = 0;
                while (x < w) {
                    if (i >= b.size()) {
                        err = path + ": truncated run";
                        return false;
                    }
                    int cnt = b[i++];
                    if (cnt > 128) {
                        cnt -= 128;
                        if (i >= b.size() || x + cnt > w) {
                            err = path + ": bad run";
                            return false;
                        }
                        unsigned char v = b[i++];
                        for (int k = 0; k < cnt; ++k) sl[4 * (size_t)(x++) + ch] = v;
                    } else {
                        if (cnt == 0 || i + cnt > b.size() || x + cnt > w) {
                            err = path + ": bad literal run";
                            return false;
                        }
                        for (int k = 0; k < cnt; ++k) sl[4 * (size_t)(x++) + ch] = b[i++];
                    }
                }
            }
        }
        for (int x = 0; x < w; ++x) {
            const unsigned char* px = &sl[4 * (size_t)x];
            if (px[3] == 0) continue;
            float f = (float)std::ldexp(1.0, (int)px[3] - (128 + 8));
            float* o = &rgb[3 * ((size_t)y * w + x)];
            o[0] = px[0] * f;
            o[1] = px[1] * f;
            o[2] = px[2] * f;
        }
    }
    return true;
}

extern "C" {

const char* rt_film_last_error(void) { return g_film_error.c_str(); }

int rt_image_read_hdr(const char* path, int32_t* width, int32_t* height, float* rgb, int64_t cap)
{
    if (!path || !width || !height) return fail(RT_ERR_INVALID, "invalid argument");
    int w = 0, h = 0;
    std::vector<float> img;
    std::string err;
    if (!rt_read_rgbe(path, w, h, img, err)) return fail(RT_ERR_IO, err);
    *width = w;
    *height = h;
    if (rgb) {
        if (cap < (int64_t)img.size()) return fail(RT_ERR_INVALID, "buffer too small");
        std::memcpy(rgb, img.data(), img.size() * sizeof(float));
    }
    return RT_OK;
}

/* OpenEXR 2 single-part scanline file: magic 20000630, version 2, the eight
 * required header attributes, a table of scanline offsets, then one block per
 * scanline (y, byte count, the B, G and R channels as 32-bit floats in
 * channel-name order).  NO_COMPRESSION. */
int rt_film_write_exr(const char* path, const float* film, int32_t w, int32_t h)
{
    if (!path || w <= 0 || h <= 0 || !film) return fail(RT_ERR_INVALID, "invalid argument");
    Out o;
    o.u32(20000630u);
    o.u32(2u);
    {
        Out v;
        for (const char* ch : {"B", "G", "R"}) {
            v.str0(ch);
            v.u32(2);  // FLOAT
            v.u8(0);   // pLinear
            v.u8(0), v.u8(0), v.u8(0);
            v.u32(1), v.u32(1);  // x / y sampling
        }
        v.u8(0);
        exr_attr(o, "channels", "chlist", v);
    }
    {
        Out v;
        v.u8(0);
        exr_attr(o, "compression", "compression", v);
    }
    for (const char* win : {"dataWindow", "displayWindow"}) {
        Out v;
        v.u32(0), v.u32(0), v.u32((uint32_t)(w - 1)), v.u32((uint32_t)(h - 1));
        exr_attr(o, win, "box2i", v);
    }
    {
        Out v;
        v.u8(0);  // INCREASING_Y
        exr_attr(o, "lineOrder", "lineOrder", v);
    }
    {
        Out v;
        v.f32(1.f);
        exr_attr(o, "pixelAspectRatio", "float", v);
    }
    {
        Out v;
        v.f32(0.f), v.f32(0.f);
        exr_attr(o, "screenWindowCenter", "v2f", v);
    }
    {
        Out v;
        v.f32(1.f);
        exr_attr(o, "screenWindowWidth", "float", v);
    }
    o.u8(0);  // end of header
    const uint64_t line_bytes = (uint64_t)w * 3 * 4;
    const uint64_t table = o.b.size();
    const uint64_t first = table + 8ull * (uint64_t)h;
    for (int y = 0; y < h; ++y) o.u64(first + (uint64_t)y * (8 + line_bytes));
    o.b.reserve(o.b.size() + (size_t)h * (8 + line_bytes));
    for (int y = 0; y < h; ++y) {
        o.u32((uint32_t)y);
        o.u32((uint32_t)line_bytes);
        for (int ch : {2, 1, 0})  // B, G, R
            for (int x = 0; x < w; ++x) o.f32(film[((size_t)y * w + x) * 3 + ch]);
    }
    return write_all(path, o.b.data(), o.b.size());
}

/* QImage::save of the tonemapped RGB32 image; the format follows the
 * extension.  PNG: 8-bit RGB, zlib-deflated scanlines with filter 0;
 * PPM: binary P6; BMP: 24-bit bottom-up rows padded to 4 bytes. */
int rt_film_write_image(const char* path, const uint8_t* rgb, int32_t w, int32_t h)
{
    if (!path || w <= 0 || h <= 0 || !rgb) return fail(RT_ERR_INVALID, "invalid argument");
    std::string p(path);
    Out o;
    if (ends_with(p, ".png")) {
        static const unsigned char sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
        o.bytes(sig, 8);
        auto chunk = [&](const char* type, const std::vector<unsigned char>& data) {
            o.u32be((uint32_t)data.size());
            size_t at = o.b.size();
            o.bytes(type, 4);
            o.bytes(data.data(), data.size());
            uLong crc = crc32(0L, Z_NULL, 0);
            crc = crc32(crc, o.b.data() + at, (uInt)(4 + data.size()));
            o.u32be((uint32_t)crc);
        };
        Out ihdr;
        ihdr.u32be((uint32_t)w), ihdr.u32be((uint32_t)h);
        ihdr.u8(8), ihdr.u8(2), ihdr.u8(0), ihdr.u8(0), ihdr.u8(0);
        chunk("IHDR", ihdr.b);
        std::vector<unsigned char> raw((size_t)h * (1 + (size_t)w * 3));
        for (int y = 0; y < h; ++y) {
            unsigned char* r = &raw[(size_t)y * (1 + (size_t)w * 3)];
            r[0] = 0;
            std::memcpy(r + 1, rgb + (size_t)y * w * 3, (size_t)w * 3);
        }
        uLongf zn = compressBound((uLong)raw.size());
        std::vector<unsigned char> z(zn);
        if (compress2(z.data(), &zn, raw.data(), (uLong)raw.size(), Z_DEFAULT_COMPRESSION) != Z_OK)
            return fail(RT_ERR_IO, "zlib compression failed");
        z.resize(zn);
        chunk("IDAT", z);
        chunk("IEND", {});
    } else if (ends_with(p, ".ppm")) {
        std::string head = "P6\n" + std::to_string(w) + " " + std::to_string(h) + "\n255\n";
        o.bytes(head.data(), head.size());
        o.bytes(rgb, (size_t)w * h * 3);
    } else if (ends_with(p, ".bmp")) {
        const uint32_t row = ((uint32_t)w * 3 + 3) & ~3u;
        const uint32_t img = row * (uint32_t)h;
        o.u8('B'), o.u8('M');
        o.u32(54 + img), o.u32(0), o.u32(54);
        o.u32(40), o.u32((uint32_t)w), o.u32((uint32_t)h), o.u16(1), o.u16(24), o.u32(0), o.u32(img);
        o.u32(2835), o.u32(2835), o.u32(0), o.u32(0);
        for (int y = h - 1; y >= 0; --y) {
            const uint8_t* s = rgb + (size_t)y * w * 3;
            for (int x = 0; x < w; ++x) o.u8(s[3 * x + 2]), o.u8(s[3 * x + 1]), o.u8(s[3 * x]);
            for (uint32_t k = (uint32_t)w * 3; k < row; ++k) o.u8(0);
        }
    } else {
        return fail(RT_ERR_UNSUPPORTED, "image format of " + p + " (use .png, .ppm or .bmp)");
    }
    return write_all(path, o.b.data(), o.b.size());
}

/* The MATLAB function the reference writes for --save-mat
 * (src/main.cpp:231-258): fixed notation with 6 decimals, the channels in
 * the order B, G, R of the OpenCV film, stored as planes 3, 2, 1. */
int rt_film_write_mat(const char* path, const float* film, int32_t w, int32_t h)
{
    if (!path || w <= 0 || h <= 0 || !film) return fail(RT_ERR_INVALID, "invalid argument");
    std::string p(path);
    size_t slash = p.find_last_of('/');
    std::string base = slash == std::string::npos ? p : p.substr(slash + 1);
    base = base.substr(0, base.find('.'));  // QFileInfo::baseName
    std::string t;
    t.reserve((size_t)w * h * 3 * 10 + 256);
    t += "function im = " + base + "()\n";
    t += "  im = zeros(" + std::to_string(h) + "," + std::to_string(w) + ",3);\n";
    char buf[64];
    for (int c = 0; c < 3; ++c) {
        const int ch = 2 - c;  // OpenCV channel c of BGR is RGB channel 2 - c
        t += "  im(:, :, " + std::to_string(3 - c) + ")=[";
        for (int i = 0; i < h; ++i) {
            for (int j = 0; j < w; ++j) {
                int n = std::snprintf(buf, sizeof buf, "%.6f", (double)film[((size_t)i * w + j) * 3 + ch]);
                t.append(buf, (size_t)n);
                t += ',';
            }
            t += ';';
        }
        t += "];\n";
    }
    t += "end\n";
    return write_all(path, t.data(), t.size());
}

/* getUniqueFilename: insert 0, 1, ... before the last '.' until the name is
 * free.  A name without '.' gets the number appended (the reference throws
 * std::out_of_range there). */
int64_t rt_unique_filename(const char* path, char* out, int64_t cap)
{
    if (!path || !out || cap <= 0) return fail(RT_ERR_INVALID, "invalid argument");
    std::string name(path), unique = name;
    size_t dot = name.find_last_of('.');
    for (long i = 0; exists(unique); ++i)
        unique = dot == std::string::npos ? name + std::to_string(i)
                                          : name.substr(0, dot) + std::to_string(i) + name.substr(dot);
    if ((int64_t)unique.size() + 1 > cap) return fail(RT_ERR_INVALID, "output buffer too small");
    std::memcpy(out, unique.c_str(), unique.size() + 1);
    return (int64_t)unique.size();
}

}  // extern "C"
