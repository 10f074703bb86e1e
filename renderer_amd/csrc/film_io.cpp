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
#include <cctype>
#include <climits>
#include <cstdint>
#include <exception>
#include <cstdlib>

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

namespace {

bool read_file(const std::string& path, std::vector<uint8_t>& buf)
{
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    uint8_t tmp[1 << 16];
    size_t n;
    while ((n = std::fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + n);
    std::fclose(f);
    return true;
}

uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }
uint32_t le32(const uint8_t* p) { return (uint32_t)p[3] << 24 | (uint32_t)p[2] << 16 | (uint32_t)p[1] << 8 | p[0]; }
uint16_t le16(const uint8_t* p) { return (uint16_t)(p[1] << 8 | p[0]); }

/* PNG (ISO/IEC 15948): IHDR, optional PLTE, IDAT stream inflated by zlib,
 * per-scanline filters None / Sub / Up / Average / Paeth; bit depth 8 or 16
 * (the high byte kept) and 1, 2, 4 for grey (scaled to 8 bits) and palette;
 * colour types grey, RGB, palette, grey + alpha, RGBA; alpha dropped, grey
 * replicated; no interlacing. */
bool decode_png(const std::vector<uint8_t>& b, int& w, int& h, std::vector<uint8_t>& rgb, std::string& err)
{
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (b.size() < 8 || std::memcmp(b.data(), sig, 8) != 0) return err = "not a PNG file", false;
    size_t pos = 8;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, pal;
    while (pos + 8 <= b.size()) {
        uint32_t len = be32(&b[pos]);
        std::string type((const char*)&b[pos + 4], 4);
        if (pos + 12 + (size_t)len > b.size()) return err = "truncated PNG chunk", false;
        const uint8_t* d = &b[pos + 8];
        if (type == "IHDR") {
            if (len < 13) return err = "bad IHDR", false;
            w = (int)be32(d);
            h = (int)be32(d + 4);
            depth = d[8];
            ctype = d[9];
            interlace = d[12];
        } else if (type == "PLTE") {
            pal.assign(d, d + len);
        } else if (type == "IDAT") {
            idat.insert(idat.end(), d, d + len);
        } else if (type == "IEND") {
            break;
        }
        pos += 12 + len;
    }
    if (w <= 0 || h <= 0 || (int64_t)w * h > (1 << 28)) return err = "bad PNG size", false;
    if (interlace) return err = "interlaced PNG is not supported", false;
    int ch = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : ctype == 6 ? 4 : 0;
    const bool sub = depth < 8;  // 1, 2, 4 bits: grey or palette only
    if (!ch || !(depth == 8 || (depth == 16 && ctype != 3) ||
                 (sub && (depth == 1 || depth == 2 || depth == 4) && (ctype == 0 || ctype == 3))))
        return err = "unsupported PNG pixel format", false;
    const size_t bpp = sub ? 1 : (size_t)ch * (depth / 8);
    const size_t stride = sub ? ((size_t)w * depth + 7) / 8 : bpp * w;
    std::vector<uint8_t> raw((stride + 1) * h);
    uLongf rn = (uLongf)raw.size();
    if (uncompress(raw.data(), &rn, idat.data(), (uLong)idat.size()) != Z_OK || rn != raw.size())
        return err = "corrupt PNG image data", false;
    std::vector<uint8_t> px(stride * h), prev(stride, 0);
    for (int y = 0; y < h; ++y) {
        const uint8_t f = raw[y * (stride + 1)];
        const uint8_t* in = &raw[y * (stride + 1) + 1];
        uint8_t* out = &px[y * stride];
        for (size_t i = 0; i < stride; ++i) {
            int a = i >= bpp ? out[i - bpp] : 0, up = prev[i], c = i >= bpp ? prev[i - bpp] : 0, v;
            switch (f) {
            case 0: v = in[i]; break;
            case 1: v = in[i] + a; break;
            case 2: v = in[i] + up; break;
            case 3: v = in[i] + ((a + up) >> 1); break;
            case 4: {
                int p = a + up - c, pa = std::abs(p - a), pb = std::abs(p - up), pc = std::abs(p - c);
                v = in[i] + (pa <= pb && pa <= pc ? a : pb <= pc ? up : c);
                break;
            }
            default: return err = "bad PNG filter", false;
            }
            out[i] = (uint8_t)v;
        }
        std::memcpy(prev.data(), out, stride);
    }
    rgb.resize((size_t)3 * w * h);
    if (sub) {  // unpack: grey scaled to 8 bits, palette indices kept
        std::vector<uint8_t> un((size_t)w * h);
        const int per = 8 / depth, mask = (1 << depth) - 1;
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                int v = (px[(size_t)y * stride + x / per] >> ((per - 1 - x % per) * depth)) & mask;
                un[(size_t)y * w + x] = (uint8_t)(ctype == 0 ? v * 255 / mask : v);
            }
        px.swap(un);
    }
    const size_t step = sub ? 1 : depth / 8, pbpp = sub ? 1 : bpp;
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        const uint8_t* q = &px[i * pbpp];
        auto sample = [&](int k) { return q[k * step]; };  // high byte of 16-bit samples
        uint8_t r, g, bl;
        if (ctype == 3) {
            size_t k = (size_t)q[0] * 3;
            if (k + 2 >= pal.size()) return err = "PNG palette index out of range", false;
            r = pal[k], g = pal[k + 1], bl = pal[k + 2];
        } else if (ctype == 0 || ctype == 4) {
            r = g = bl = sample(0);
        } else {
            r = sample(0), g = sample(1), bl = sample(2);
        }
        rgb[3 * i] = r, rgb[3 * i + 1] = g, rgb[3 * i + 2] = bl;
    }
    return true;
}

/* Binary PPM / PGM (P6, P5), maxval up to 65535 (high byte kept above 255). */
bool decode_pnm(const std::vector<uint8_t>& b, int& w, int& h, std::vector<uint8_t>& rgb, std::string& err)
{
    size_t pos = 2;
    auto field = [&](int& v) {
        while (pos < b.size()) {
            if (b[pos] == '#') {
                while (pos < b.size() && b[pos] != '\n') ++pos;
            } else if (std::isspace(b[pos])) {
                ++pos;
            } else {
                break;
            }
        }
        if (pos >= b.size() || !std::isdigit(b[pos])) return false;
        v = 0;
        while (pos < b.size() && std::isdigit(b[pos])) {
            const int dgt = b[pos++] - '0';
            if (v > (INT_MAX - dgt) / 10) return false;  /* the number does not fit an int */
            v = v * 10 + dgt;
        }
        return true;
    };
    const bool color = b[1] == '6';
    int maxv = 0;
    if (!field(w) || !field(h) || !field(maxv) || w <= 0 || h <= 0 || maxv <= 0 || maxv > 65535 ||
        (int64_t)w * h > (1ll << 28))
        return err = "bad PNM header", false;
    ++pos;  // one whitespace byte after maxval
    const int ch = color ? 3 : 1, bytes = maxv > 255 ? 2 : 1;
    /* w * h <= 2^28, so the byte count below stays far from overflow */
    if (pos > b.size() || (uint64_t)w * h * ch * bytes > b.size() - pos) return err = "truncated PNM data", false;
    rgb.resize((size_t)3 * w * h);
    for (size_t i = 0; i < (size_t)w * h; ++i)
        for (int k = 0; k < 3; ++k) {
            size_t at = pos + (i * ch + (color ? k : 0)) * bytes;
            rgb[3 * i + k] = b[at];
        }
    return true;
}

/* BMP: BITMAPINFOHEADER or later, uncompressed 24 or 32 bits, rows bottom-up
 * (or top-down for a negative height), each padded to 4 bytes, BGR(A). */
bool decode_bmp(const std::vector<uint8_t>& b, int& w, int& h, std::vector<uint8_t>& rgb, std::string& err)
{
    if (b.size() < 54) return err = "truncated BMP", false;
    const uint32_t off = le32(&b[10]);
    w = (int32_t)le32(&b[18]);
    int32_t hh = (int32_t)le32(&b[22]);
    const int bits = le16(&b[28]);
    const uint32_t comp = le32(&b[30]);
    if (w <= 0 || hh == 0 || hh == INT32_MIN || !(bits == 24 || bits == 32) || !(comp == 0 || (comp == 3 && bits == 32)))
        return err = "unsupported BMP format", false;
    h = hh < 0 ? -hh : hh;
    if ((int64_t)w * h > (1ll << 28)) return err = "bad BMP size", false;
    const size_t bpp = bits / 8, row = ((size_t)w * bpp + 3) & ~(size_t)3;
    /* w * h <= 2^28: row * h is below 2^31 * 4 and cannot wrap */
    if (off > b.size() || row * (size_t)h > b.size() - off) return err = "truncated BMP data", false;
    rgb.resize((size_t)3 * w * h);
    for (int y = 0; y < h; ++y) {
        const uint8_t* src = &b[off + row * (hh < 0 ? y : h - 1 - y)];
        for (int x = 0; x < w; ++x) {
            uint8_t* o = &rgb[3 * ((size_t)y * w + x)];
            o[0] = src[x * bpp + 2], o[1] = src[x * bpp + 1], o[2] = src[x * bpp];
        }
    }
    return true;
}

}  // namespace

/* An 8-bit image as the reference reads textures (cv::imread with the
 * default flags, then convertTo(CV_32F, 1./255.), src/texturematerial.cpp:46-60):
 * RGB floats byte / 255, row 0 at the top.  PNG, binary PPM / PGM and BMP. */
bool rt_read_image8(const std::string& path, int& w, int& h, std::vector<float>& rgb, std::string& err)
{
    std::vector<uint8_t> b;
    if (!read_file(path, b)) return err = "cannot open " + path, false;
    std::vector<uint8_t> px;
    bool ok;
    try {  /* no exception (bad_alloc of a large image) may leave through the C ABI */
        if (b.size() >= 8 && b[0] == 137 && b[1] == 'P' && b[2] == 'N' && b[3] == 'G')
            ok = decode_png(b, w, h, px, err);
        else if (b.size() >= 3 && b[0] == 'P' && (b[1] == '6' || b[1] == '5'))
            ok = decode_pnm(b, w, h, px, err);
        else if (b.size() >= 2 && b[0] == 'B' && b[1] == 'M')
            ok = decode_bmp(b, w, h, px, err);
        else
            return err = "unrecognised image format: " + path, false;
        if (!ok) return false;
        rgb.resize(px.size());
    } catch (const std::exception& e) {
        return err = std::string("cannot decode ") + path + ": " + e.what(), false;
    }
    for (size_t i = 0; i < px.size(); ++i) rgb[i] = (float)(px[i] * (1.0 / 255.0));
    return true;
}

extern "C" {

const char* rt_film_last_error(void) { return g_film_error.c_str(); }

int rt_image_read(const char* path, int32_t* width, int32_t* height, float* rgb, int64_t cap)
{
    if (!path || !width || !height) return fail(RT_ERR_INVALID, "invalid argument");
    int w = 0, h = 0;
    std::vector<float> img;
    std::string err;
    if (!rt_read_image8(path, w, h, img, err)) return fail(RT_ERR_IO, err);
    *width = w;
    *height = h;
    if (rgb) {
        if (cap < (int64_t)img.size()) return fail(RT_ERR_INVALID, "buffer too small");
        std::memcpy(rgb, img.data(), img.size() * sizeof(float));
    }
    return RT_OK;
}

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
