// png_io.cpp -- PNG input and result images of the image operators (C ABI
// ske_png_*).  The reference reads PNGs with lodepng and keeps, per pixel,
// the sum R + G + B of the 8-bit RGBA decode (src/helper.h:167-186); its
// showResults writes salida1.png (that grey level / 3) and salida2.png (the
// output mapped to grey) for the Radon and identity-image operators
// (src/main.h:124-141, src/kernelRADN.h:383-401).
//
// This is a PNG codec written from the PNG specification on top of zlib:
// non-interlaced images of every colour type (grey, RGB, palette,
// grey+alpha, RGBA) at every legal bit depth are decoded to 8-bit RGBA the
// way lodepng converts them (sub-byte grey scaled to 0..255, 16-bit samples
// keep their high byte, palette entries looked up); the encoder writes 8-bit
// grey or RGB with filter type 0.
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <zlib.h>

#include <string>
#include <vector>

#include "host_util.h"
#include "skewengine.h"

using namespace ske;

namespace {

const uint8_t kSig[8] = {137, 80, 78, 71, 13, 10, 26, 10};

uint32_t be32(const uint8_t* p) { return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]; }

void put32(std::vector<uint8_t>& v, uint32_t x) {
    v.push_back((uint8_t)(x >> 24));
    v.push_back((uint8_t)(x >> 16));
    v.push_back((uint8_t)(x >> 8));
    v.push_back((uint8_t)x);
}

int read_file(const char* path, std::vector<uint8_t>& data) {
    FILE* f = fopen(path, "rb");
    if (!f) return fail(SKE_EINVAL, std::string("cannot open ") + path);
    uint8_t buf[1 << 16];
    size_t n;
    while ((n = fread(buf, 1, sizeof buf, f)) > 0) data.insert(data.end(), buf, buf + n);
    fclose(f);
    return SKE_OK;
}

int paeth(int a, int b, int c) {
    const int p = a + b - c;
    const int pa = p > a ? p - a : a - p, pb = p > b ? p - b : b - p, pc = p > c ? p - c : c - p;
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

struct Png {
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = 0;
    std::vector<uint8_t> plte, trns, idat;
};

int channels_of(int ctype) {
    switch (ctype) {
        case 0: return 1;
        case 2: return 3;
        case 3: return 1;
        case 4: return 2;
        case 6: return 4;
        default: return 0;
    }
}

bool depth_ok(int ctype, int d) {
    switch (ctype) {
        case 0: return d == 1 || d == 2 || d == 4 || d == 8 || d == 16;
        case 3: return d == 1 || d == 2 || d == 4 || d == 8;
        case 2: case 4: case 6: return d == 8 || d == 16;
        default: return false;
    }
}

int parse(const std::vector<uint8_t>& f, Png& png) {
    if (f.size() < 8 || memcmp(f.data(), kSig, 8) != 0) return fail(SKE_EINVAL, "not a PNG file");
    size_t pos = 8;
    bool have_ihdr = false, have_iend = false;
    while (pos + 12 <= f.size()) {
        const uint32_t len = be32(&f[pos]);
        if (len > f.size() - pos - 12) return fail(SKE_EINVAL, "truncated PNG chunk");
        const uint8_t* type = &f[pos + 4];
        const uint8_t* d = &f[pos + 8];
        const uint32_t crc = be32(d + len);
        if ((uint32_t)crc32(0L, type, len + 4) != crc) return fail(SKE_EINVAL, "PNG chunk CRC mismatch");
        if (!memcmp(type, "IHDR", 4)) {
            if (len != 13) return fail(SKE_EINVAL, "bad IHDR");
            png.w = be32(d);
            png.h = be32(d + 4);
            png.depth = d[8];
            png.ctype = d[9];
            if (d[10] != 0 || d[11] != 0) return fail(SKE_EINVAL, "unknown PNG compression or filter method");
            if (d[12] != 0) return fail(SKE_EUNSUPPORTED, "interlaced (Adam7) PNG");
            if (!channels_of(png.ctype) || !depth_ok(png.ctype, png.depth))
                return fail(SKE_EINVAL, "invalid PNG colour type / bit depth");
            have_ihdr = true;
        } else if (!memcmp(type, "PLTE", 4)) {
            png.plte.assign(d, d + len);
        } else if (!memcmp(type, "tRNS", 4)) {
            png.trns.assign(d, d + len);
        } else if (!memcmp(type, "IDAT", 4)) {
            png.idat.insert(png.idat.end(), d, d + len);
        } else if (!memcmp(type, "IEND", 4)) {
            have_iend = true;
            break;
        } else if (!(type[0] & 0x20)) {
            return fail(SKE_EUNSUPPORTED, "unknown critical PNG chunk");
        }
        pos += 12 + (size_t)len;
    }
    if (!have_ihdr || !have_iend) return fail(SKE_EINVAL, "PNG without IHDR or IEND");
    if (png.w == 0 || png.h == 0 || png.w > 65535 || png.h > 65535) return fail(SKE_EINVAL, "PNG size out of range");
    if (png.ctype == 3 && (png.plte.empty() || png.plte.size() % 3)) return fail(SKE_EINVAL, "palette PNG without PLTE");
    return SKE_OK;
}

// Inflate and unfilter: one row of `stride` bytes per scanline.
int unfilter(const Png& png, std::vector<uint8_t>& raw, size_t& stride) {
    const int bits = channels_of(png.ctype) * png.depth;
    stride = ((size_t)png.w * bits + 7) / 8;
    const size_t bpp = std::max<size_t>(1, (size_t)bits / 8);
    std::vector<uint8_t> buf((stride + 1) * png.h);
    z_stream zs;
    memset(&zs, 0, sizeof zs);
    if (inflateInit(&zs) != Z_OK) return fail(SKE_ENOMEM, "inflateInit failed");
    zs.next_in = const_cast<Bytef*>(png.idat.data());
    zs.avail_in = (uInt)png.idat.size();
    zs.next_out = buf.data();
    zs.avail_out = (uInt)buf.size();
    const int zr = inflate(&zs, Z_FINISH);
    const size_t got = buf.size() - zs.avail_out;
    inflateEnd(&zs);
    if ((zr != Z_STREAM_END && zr != Z_BUF_ERROR) || got != buf.size())
        return fail(SKE_EINVAL, "PNG image data is corrupt or short");
    raw.assign(stride * png.h, 0);
    std::vector<uint8_t> prev(stride, 0);
    for (uint32_t y = 0; y < png.h; y++) {
        const uint8_t ft = buf[y * (stride + 1)];
        const uint8_t* src = &buf[y * (stride + 1) + 1];
        uint8_t* out = &raw[y * stride];
        for (size_t i = 0; i < stride; i++) {
            const int a = i >= bpp ? out[i - bpp] : 0, b = prev[i], c = i >= bpp ? prev[i - bpp] : 0;
            int v = src[i];
            switch (ft) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += b; break;
                case 3: v += (a + b) >> 1; break;
                case 4: v += paeth(a, b, c); break;
                default: return fail(SKE_EINVAL, "unknown PNG filter type");
            }
            out[i] = (uint8_t)v;
        }
        memcpy(prev.data(), out, stride);
    }
    return SKE_OK;
}

// Sample s (0-based, in channel units) of row `row` at `depth` bits.
unsigned sample(const uint8_t* row, size_t s, int depth) {
    if (depth == 8) return row[s];
    if (depth == 16) return ((unsigned)row[2 * s] << 8) | row[2 * s + 1];
    const size_t bit = s * depth;
    return (row[bit >> 3] >> (8 - depth - (bit & 7))) & ((1u << depth) - 1);
}

// 8-bit RGB of pixel x as lodepng's RGBA conversion gives it.
void rgb8(const Png& png, const uint8_t* row, uint32_t x, uint8_t rgb[3]) {
    const int d = png.depth;
    auto to8 = [d](unsigned v) -> uint8_t {
        if (d == 16) return (uint8_t)(v >> 8);
        if (d == 8) return (uint8_t)v;
        return (uint8_t)(v * 255 / ((1u << d) - 1));
    };
    switch (png.ctype) {
        case 0:
        case 4: {
            const uint8_t g = to8(sample(row, (size_t)x * channels_of(png.ctype), d));
            rgb[0] = rgb[1] = rgb[2] = g;
            break;
        }
        case 2:
        case 6: {
            const size_t c = channels_of(png.ctype);
            for (int k = 0; k < 3; k++) rgb[k] = to8(sample(row, (size_t)x * c + k, d));
            break;
        }
        case 3: {
            const unsigned idx = sample(row, x, d);
            if (3 * (size_t)idx + 2 < png.plte.size()) {
                for (int k = 0; k < 3; k++) rgb[k] = png.plte[3 * (size_t)idx + k];
            } else {
                rgb[0] = rgb[1] = rgb[2] = 0;  // out-of-range index: black
            }
            break;
        }
    }
}

int write_png(const char* path, int w, int h, int channels, const uint8_t* data) {
    const size_t stride = (size_t)w * channels;
    std::vector<uint8_t> raw((stride + 1) * h);
    for (int y = 0; y < h; y++) {
        raw[y * (stride + 1)] = 0;
        memcpy(&raw[y * (stride + 1) + 1], data + y * stride, stride);
    }
    uLongf zlen = compressBound(raw.size());
    std::vector<uint8_t> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), raw.size(), 6) != Z_OK) return fail(SKE_ENOMEM, "deflate failed");
    std::vector<uint8_t> out(kSig, kSig + 8);
    auto chunk = [&out](const char* type, const uint8_t* d, size_t len) {
        put32(out, (uint32_t)len);
        const size_t at = out.size();
        out.insert(out.end(), type, type + 4);
        out.insert(out.end(), d, d + len);
        put32(out, (uint32_t)crc32(0L, &out[at], (uInt)(len + 4)));
    };
    std::vector<uint8_t> ihdr;
    put32(ihdr, (uint32_t)w);
    put32(ihdr, (uint32_t)h);
    ihdr.push_back(8);
    ihdr.push_back(channels == 1 ? 0 : 2);
    ihdr.push_back(0);
    ihdr.push_back(0);
    ihdr.push_back(0);
    chunk("IHDR", ihdr.data(), ihdr.size());
    chunk("IDAT", z.data(), zlen);
    chunk("IEND", nullptr, 0);
    FILE* f = fopen(path, "wb");
    if (!f) return fail(SKE_EINVAL, std::string("cannot create ") + path);
    const size_t n = fwrite(out.data(), 1, out.size(), f);
    fclose(f);
    if (n != out.size()) return fail(SKE_EINVAL, std::string("short write to ") + path);
    return SKE_OK;
}

}  // namespace

extern "C" {

int ske_png_read(const char* path, int* width, int* height, float* pixels) {
    if (!path || !width || !height) return fail(SKE_EINVAL, "path, width or height is NULL");
    std::vector<uint8_t> file;
    SKE_TRY(read_file(path, file));
    Png png;
    SKE_TRY(parse(file, png));
    *width = (int)png.w;
    *height = (int)png.h;
    if (!pixels) return SKE_OK;
    std::vector<uint8_t> raw;
    size_t stride = 0;
    SKE_TRY(unfilter(png, raw, stride));
    for (uint32_t y = 0; y < png.h; y++) {
        const uint8_t* row = &raw[y * stride];
        for (uint32_t x = 0; x < png.w; x++) {
            uint8_t c[3];
            rgb8(png, row, x, c);
            // pixels[i] = R + G + B as a short, then copied to float
            // (src/helper.h:182, src/kernelRADN.h:89)
            pixels[(size_t)y * png.w + x] = (float)(short)(c[0] + c[1] + c[2]);
        }
    }
    return SKE_OK;
}

int ske_png_write(const char* path, int width, int height, int channels, const unsigned char* data) {
    if (!path || !data) return fail(SKE_EINVAL, "path or data is NULL");
    if (width < 1 || height < 1 || width > 65535 || height > 65535) return fail(SKE_EINVAL, "PNG size out of range");
    if (channels != 1 && channels != 3) return fail(SKE_EINVAL, "channels must be 1 (grey) or 3 (RGB)");
    return write_png(path, width, height, channels, data);
}

int ske_result_pngs(const char* dir, int kernel_id, int width, int height, const float* pixels, const float* output) {
    if (!dir || !pixels || !output) return fail(SKE_EINVAL, "dir, pixels or output is NULL");
    if (kernel_id != SKE_KERNEL_RADON && kernel_id != SKE_KERNEL_IDENTITYIMG)
        return fail(SKE_EINVAL, "result images exist for the Radon and identity-image operators only");
    if (width < 1 || height < 1) return fail(SKE_EINVAL, "bad size");
    const size_t n = (size_t)width * height;
    // src/main.h:125-131: min/max of the output (FLT_MAX / FLT_MIN starts),
    // escala = 255/(max-min) for Radon (shift = min), 1/540 for identity.
    float mn = 3.40282347e38f, mx = 1.17549435e-38f;
    for (size_t i = 0; i < n; i++) {
        mn = output[i] < mn ? output[i] : mn;
        mx = output[i] > mx ? output[i] : mx;
    }
    float escala, shift;
    if (kernel_id == SKE_KERNEL_RADON) {
        escala = 255 / (mx - mn);
        shift = mn;
    } else {
        escala = (float)(1.0 / 540);
        shift = 0;
    }
    std::vector<uint8_t> grey(n), rgb(3 * n);
    for (size_t i = 0; i < n; i++) {
        grey[i] = (uint8_t)((short)pixels[i] / 3);
        // (unsigned char)((int)(out - shift) * escala): the float is
        // truncated to an integer whose low byte is kept
        const float v = (float)(int)(output[i] - shift) * escala;
        const uint8_t c = (uint8_t)(int32_t)v;
        rgb[3 * i] = rgb[3 * i + 1] = rgb[3 * i + 2] = c;
    }
    const std::string d(dir);
    const std::string sep = (!d.empty() && d.back() != '/') ? "/" : "";
    SKE_TRY(write_png((d + sep + "salida1.png").c_str(), width, height, 1, grey.data()));
    return write_png((d + sep + "salida2.png").c_str(), width, height, 3, rgb.data());
}

}  // extern "C"
