// Image files of the scene description and the renderer's output, written
// from the published formats:
//   Radiance RGBE (.hdr): header, "-Y H +X W" resolution line, scanlines flat
//     or adaptive run-length encoded; value = mantissa * 2^(e - 136).
//   PFM: "PF\n W H\n scale\n", little-endian when scale < 0, rows bottom-up.
//   OpenEXR: scanline images with NO_COMPRESSION and HALF or FLOAT R, G, B
//     channels (read); written as FLOAT B, G, R, one scanline per chunk.
//   PNG (textures): non-interlaced, 8- or 16-bit grey, grey + alpha, RGB,
//     RGBA or 8-bit palette; zlib stream inflated by the system zlib, rows
//     unfiltered (None, Sub, Up, Average, Paeth), 16-bit samples reduced to
//     their high byte (as stb_image returns them).  Written (renders): 8-bit
//     sRGB RGB, unfiltered rows, deflated by zlib.
// The reference reads images through stb_image / OpenEXR
// (src/core/bitmap/bitmap.cpp:33-214), neither of which is vendored.
#include <zlib.h>

#include <cmath>
#include <cstring>
#include <fstream>

#include "drawlab/plugins.h"

namespace drawlab {
namespace {

std::vector<unsigned char> read_file(const std::string& fn) {
    std::ifstream in(fn, std::ios::binary);
    if (!in) throw Exception("Could not open image \"%s\"", fn);
    return std::vector<unsigned char>((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
}

Image load_hdr(const std::string& fn) {
    const std::vector<unsigned char> b = read_file(fn);
    size_t pos = 0;
    auto line = [&]() {
        std::string s;
        while (pos < b.size() && b[pos] != '\n') s += (char)b[pos++];
        if (pos < b.size()) ++pos;
        return s;
    };
    const std::string magic = line();
    if (magic.rfind("#?", 0) != 0) throw Exception("\"%s\" is not a Radiance HDR file", fn);
    for (;;) {
        if (pos >= b.size()) throw Exception("\"%s\": truncated header", fn);
        const std::string l = line();
        if (l.empty()) break;
        if (l.rfind("FORMAT=", 0) == 0 && l != "FORMAT=32-bit_rle_rgbe")
            throw Exception("\"%s\": unsupported format %s", fn, l);
    }
    const std::string res = line();
    Image img;
    char ya[3], xa[3];
    if (std::sscanf(res.c_str(), "%2s %d %2s %d", ya, &img.height, xa, &img.width) != 4 || std::string(ya) != "-Y" ||
        std::string(xa) != "+X")
        throw Exception("\"%s\": unsupported resolution line \"%s\"", fn, res);
    const int W = img.width, H = img.height;
    img.rgb.resize((size_t)W * H * 3);
    std::vector<unsigned char> scan((size_t)W * 4);
    auto need = [&](size_t n) {
        if (pos + n > b.size()) throw Exception("\"%s\": truncated pixel data", fn);
    };
    for (int y = 0; y < H; y++) {
        need(4);
        if (W >= 8 && W < 32768 && b[pos] == 2 && b[pos + 1] == 2 && ((b[pos + 2] << 8) | b[pos + 3]) == W) {
            pos += 4;
            for (int c = 0; c < 4; c++) {
                int x = 0;
                while (x < W) {
                    need(1);
                    int count = b[pos++];
                    if (count > 128) {
                        count -= 128;
                        need(1);
                        const unsigned char v = b[pos++];
                        if (x + count > W) throw Exception("\"%s\": bad run length", fn);
                        for (int i = 0; i < count; i++) scan[(size_t)4 * (x++) + c] = v;
                    } else {
                        if (count == 0 || x + count > W) throw Exception("\"%s\": bad run length", fn);
                        need(count);
                        for (int i = 0; i < count; i++) scan[(size_t)4 * (x++) + c] = b[pos++];
                    }
                }
            }
        } else {
            need((size_t)W * 4);
            std::memcpy(scan.data(), &b[pos], (size_t)W * 4);
            pos += (size_t)W * 4;
        }
        for (int x = 0; x < W; x++) {
            const unsigned char* p = &scan[(size_t)4 * x];
            float* o = &img.rgb[((size_t)y * W + x) * 3];
            if (p[3] == 0) {
                o[0] = o[1] = o[2] = 0.f;
            } else {
                const float f = std::ldexp(1.0f, (int)p[3] - (128 + 8));
                o[0] = p[0] * f;
                o[1] = p[1] * f;
                o[2] = p[2] * f;
            }
        }
    }
    return img;
}

Image load_pfm(const std::string& fn) {
    const std::vector<unsigned char> b = read_file(fn);
    size_t pos = 0;
    auto token = [&]() {
        while (pos < b.size() && std::isspace(b[pos])) ++pos;
        std::string s;
        while (pos < b.size() && !std::isspace(b[pos])) s += (char)b[pos++];
        return s;
    };
    const std::string magic = token();
    if (magic != "PF" && magic != "Pf") throw Exception("\"%s\" is not a PFM file", fn);
    const int channels = magic == "PF" ? 3 : 1;
    Image img;
    img.width = std::stoi(token());
    img.height = std::stoi(token());
    const float scale = std::stof(token());
    ++pos;  // the single whitespace byte before the raster
    const size_t n = (size_t)img.width * img.height * channels;
    if (pos + n * 4 > b.size()) throw Exception("\"%s\": truncated raster", fn);
    const bool little = scale < 0;
    img.rgb.resize((size_t)img.width * img.height * 3);
    for (int y = 0; y < img.height; y++)
        for (int x = 0; x < img.width; x++)
            for (int c = 0; c < 3; c++) {
                const size_t src = (((size_t)y * img.width + x) * channels + (channels == 3 ? c : 0)) * 4;
                unsigned char v[4] = {b[pos + src], b[pos + src + 1], b[pos + src + 2], b[pos + src + 3]};
                if (!little) std::swap(v[0], v[3]), std::swap(v[1], v[2]);
                float f;
                std::memcpy(&f, v, 4);
                // PFM rows run bottom-up
                img.rgb[(((size_t)(img.height - 1 - y)) * img.width + x) * 3 + c] = f;
            }
    return img;
}

float half_to_float(uint16_t h) {
    const int s = (h >> 15) & 1, e = (h >> 10) & 31, m = h & 1023;
    float v;
    if (e == 0)
        v = std::ldexp((float)m, -24);
    else if (e == 31)
        v = m ? NAN : INFINITY;
    else
        v = std::ldexp((float)(m + 1024), e - 25);
    return s ? -v : v;
}

template <typename T>
T rd(const std::vector<unsigned char>& b, size_t& pos) {
    if (pos + sizeof(T) > b.size()) throw Exception("truncated OpenEXR file");
    T v;
    std::memcpy(&v, &b[pos], sizeof(T));
    pos += sizeof(T);
    return v;
}

Image load_exr(const std::string& fn) {
    const std::vector<unsigned char> b = read_file(fn);
    size_t pos = 0;
    if (rd<uint32_t>(b, pos) != 20000630u) throw Exception("\"%s\" is not an OpenEXR file", fn);
    const uint32_t version = rd<uint32_t>(b, pos);
    if ((version & 0xff) != 2 || (version & 0x1e00) != 0) throw Exception("\"%s\": only single-part scanline OpenEXR files are supported", fn);
    struct Ch {
        std::string name;
        int type;
    };
    std::vector<Ch> chans;
    int compression = -1, xmin = 0, ymin = 0, xmax = -1, ymax = -1;
    for (;;) {
        std::string name, type;
        while (pos < b.size() && b[pos]) name += (char)b[pos++];
        ++pos;
        if (name.empty()) break;
        while (pos < b.size() && b[pos]) type += (char)b[pos++];
        ++pos;
        const int size = rd<int32_t>(b, pos);
        const size_t end = pos + (size_t)size;
        if (name == "channels") {
            while (pos < end && b[pos]) {
                Ch c;
                while (b[pos]) c.name += (char)b[pos++];
                ++pos;
                c.type = rd<int32_t>(b, pos);
                pos += 4;  // pLinear + reserved
                const int xs = rd<int32_t>(b, pos), ys = rd<int32_t>(b, pos);
                if (xs != 1 || ys != 1) c.type = -1;
                chans.push_back(c);
            }
        } else if (name == "compression") {
            compression = b[pos];
        } else if (name == "dataWindow") {
            xmin = rd<int32_t>(b, pos);
            ymin = rd<int32_t>(b, pos);
            xmax = rd<int32_t>(b, pos);
            ymax = rd<int32_t>(b, pos);
        }
        pos = end;
    }
    if (compression != 0) throw Exception("\"%s\": only uncompressed OpenEXR files are supported", fn);
    Image img;
    img.width = xmax - xmin + 1;
    img.height = ymax - ymin + 1;
    if (img.width <= 0 || img.height <= 0) throw Exception("\"%s\": empty data window", fn);
    int rgb_idx[3] = {-1, -1, -1};
    for (size_t i = 0; i < chans.size(); i++) {
        const std::string n = toLower(chans[i].name);
        const int which = (n == "r" || n == "red" || endsWith(n, ".r")) ? 0
                          : (n == "g" || n == "green" || endsWith(n, ".g")) ? 1
                          : (n == "b" || n == "blue" || endsWith(n, ".b")) ? 2 : -1;
        if (which >= 0 && rgb_idx[which] < 0 && chans[i].type >= 1) rgb_idx[which] = (int)i;
    }
    if (rgb_idx[0] < 0 || rgb_idx[1] < 0 || rgb_idx[2] < 0) throw Exception("This is not a standard RGB OpenEXR file!");
    size_t row_bytes = 0;
    for (const Ch& c : chans) row_bytes += (size_t)img.width * (c.type == 1 ? 2 : 4);
    img.rgb.resize((size_t)img.width * img.height * 3);
    pos += (size_t)img.height * 8;  // offset table
    for (int k = 0; k < img.height; k++) {
        const int y = rd<int32_t>(b, pos) - ymin;
        const int sz = rd<int32_t>(b, pos);
        if (y < 0 || y >= img.height || (size_t)sz != row_bytes || pos + row_bytes > b.size())
            throw Exception("\"%s\": malformed scanline chunk", fn);
        size_t off = pos;
        for (size_t ci = 0; ci < chans.size(); ci++) {
            const int bpp = chans[ci].type == 1 ? 2 : 4;
            for (int c = 0; c < 3; c++)
                if (rgb_idx[c] == (int)ci)
                    for (int x = 0; x < img.width; x++) {
                        float v;
                        if (bpp == 2) {
                            uint16_t h;
                            std::memcpy(&h, &b[off + (size_t)x * 2], 2);
                            v = half_to_float(h);
                        } else {
                            std::memcpy(&v, &b[off + (size_t)x * 4], 4);
                        }
                        img.rgb[((size_t)y * img.width + x) * 3 + c] = v;
                    }
            off += (size_t)img.width * bpp;
        }
        pos += row_bytes;
    }
    return img;
}

uint32_t be32(const unsigned char* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

// 8-bit samples of a PNG, `channels` per pixel, rows top first.
std::vector<unsigned char> load_png(const std::string& fn, int& w, int& h, int& channels) {
    const std::vector<unsigned char> b = read_file(fn);
    static const unsigned char sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (b.size() < 8 || std::memcmp(b.data(), sig, 8) != 0) throw Exception("\"%s\" is not a PNG file", fn);
    size_t pos = 8;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<unsigned char> idat, plte;
    w = h = 0;
    while (pos + 8 <= b.size()) {
        const uint32_t len = be32(&b[pos]);
        const std::string type(reinterpret_cast<const char*>(&b[pos + 4]), 4);
        pos += 8;
        if (pos + (size_t)len + 4 > b.size()) throw Exception("\"%s\": truncated PNG chunk", fn);
        const unsigned char* d = &b[pos];
        if (type == "IHDR") {
            if (len < 13) throw Exception("\"%s\": bad IHDR", fn);
            w = (int)be32(d);
            h = (int)be32(d + 4);
            depth = d[8];
            ctype = d[9];
            interlace = d[12];
        } else if (type == "PLTE") {
            plte.assign(d, d + len);
        } else if (type == "IDAT") {
            idat.insert(idat.end(), d, d + len);
        } else if (type == "IEND") {
            break;
        }
        pos += (size_t)len + 4;  // data + CRC
    }
    if (w <= 0 || h <= 0 || (uint64_t)w * (uint64_t)h > (1ull << 28)) throw Exception("\"%s\": bad PNG size", fn);
    if (interlace != 0) throw Exception("\"%s\": interlaced PNG is not supported", fn);
    int spp;  // samples per pixel in the file
    switch (ctype) {
        case 0: spp = 1; break;
        case 2: spp = 3; break;
        case 3: spp = 1; break;
        case 4: spp = 2; break;
        case 6: spp = 4; break;
        default: throw Exception("\"%s\": unknown PNG colour type %d", fn, ctype);
    }
    if (!((depth == 8) || (depth == 16 && ctype != 3)))
        throw Exception("\"%s\": PNG bit depth %d is not supported (8, or 16 without a palette)", fn, depth);
    const size_t bpp = (size_t)spp * (depth / 8);  // bytes per pixel
    const size_t stride = (size_t)w * bpp;
    std::vector<unsigned char> raw(((size_t)stride + 1) * (size_t)h);
    uLongf out_len = (uLongf)raw.size();
    if (uncompress(raw.data(), &out_len, idat.data(), (uLong)idat.size()) != Z_OK || out_len != raw.size())
        throw Exception("\"%s\": corrupt PNG image data", fn);
    std::vector<unsigned char> px(stride * (size_t)h), prev(stride, 0);
    for (int y = 0; y < h; y++) {
        const unsigned char f = raw[(size_t)y * (stride + 1)];
        const unsigned char* in = &raw[(size_t)y * (stride + 1) + 1];
        unsigned char* o = &px[(size_t)y * stride];
        for (size_t i = 0; i < stride; i++) {
            const int a = i >= bpp ? o[i - bpp] : 0, up = prev[i], c = i >= bpp ? prev[i - bpp] : 0;
            int v = in[i];
            switch (f) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += up; break;
                case 3: v += (a + up) / 2; break;
                case 4: {
                    const int p = a + up - c, pa = std::abs(p - a), pb = std::abs(p - up), pc = std::abs(p - c);
                    v += (pa <= pb && pa <= pc) ? a : (pb <= pc ? up : c);
                    break;
                }
                default: throw Exception("\"%s\": bad PNG filter %d", fn, (int)f);
            }
            o[i] = (unsigned char)v;
        }
        std::memcpy(prev.data(), o, stride);
    }
    // 8-bit samples; a palette expands to RGB
    channels = ctype == 3 ? 3 : spp;
    std::vector<unsigned char> out((size_t)w * h * channels);
    for (size_t i = 0; i < (size_t)w * h; i++) {
        if (ctype == 3) {
            const size_t e = px[i];
            if (3 * e + 2 >= plte.size()) throw Exception("\"%s\": palette index out of range", fn);
            for (int k = 0; k < 3; k++) out[3 * i + k] = plte[3 * e + k];
        } else {
            for (int k = 0; k < spp; k++) out[(size_t)spp * i + k] = px[bpp * i + (size_t)k * (depth / 8)];
        }
    }
    return out;
}

}  // namespace

Image loadTextureImage(const std::string& filename) {
    const std::string l = toLower(filename);
    Image img;
    if (endsWith(l, ".png")) {
        int w, h, c;
        std::vector<unsigned char> px = load_png(filename, w, h, c);
        // Bitmap::loadLDR (bitmap.cpp:131-157): stb loads the rows bottom up,
        // then the first 3 w h samples become the RGB triples (grey
        // replicated; with 4 channels the reference reads the interleaved
        // samples as if they were RGB, and so does this).  Grey + alpha has
        // only 2 w h samples, which the reference reads past: rejected.
        if (c == 2) throw Exception("\"%s\": grey + alpha PNG textures are not supported", filename);
        std::vector<unsigned char> flip(px.size());
        const size_t row = (size_t)w * c;
        for (int y = 0; y < h; y++) std::memcpy(&flip[(size_t)y * row], &px[(size_t)(h - 1 - y) * row], row);
        img.width = w;
        img.height = h;
        img.rgb.resize((size_t)w * h * 3);
        for (size_t i = 0; i < img.rgb.size(); i++) {
            const unsigned char v = c == 1 ? flip[i / 3] : flip[i];
            img.rgb[i] = (float)v / 255.0f;  // cudaReadModeNormalizedFloat of an 8-bit texel
        }
        return img;  // uploaded as loaded (bitmap.cpp:74-81)
    }
    if (endsWith(l, ".exr") || endsWith(l, ".pfm")) {
        // loaded top row first, flipped on upload (bitmap.cpp:86)
        img = loadImage(filename);
        const size_t row = (size_t)img.width * 3;
        for (int y = 0; y < img.height / 2; y++)
            for (size_t i = 0; i < row; i++) std::swap(img.rgb[(size_t)y * row + i], img.rgb[(size_t)(img.height - 1 - y) * row + i]);
        return img;
    }
    if (endsWith(l, ".hdr") || endsWith(l, ".pic")) {
        // flipped on load (bitmap.cpp:109) and again on upload: top row first
        return loadImage(filename);
    }
    throw Exception("Unsupported texture format \"%s\" (expected .png, .exr, .pfm or .hdr)", filename);
}

Image loadImage(const std::string& filename) {
    const std::string l = toLower(filename);
    if (endsWith(l, ".exr")) return load_exr(filename);
    if (endsWith(l, ".pfm")) return load_pfm(filename);
    if (endsWith(l, ".hdr") || endsWith(l, ".pic")) return load_hdr(filename);
    throw Exception("Unsupported image format \"%s\" (expected .hdr, .pfm or .exr)", filename);
}

// Bitmap::savePNG (src/core/bitmap/bitmap.cpp:190-214): each channel
// through Color3f::toSRGB (include/core/bitmap/color.h:76-86), times 255,
// clamped to [0, 255] and truncated to 8 bits; RGB rows top first.  Rows
// are stored unfiltered in one zlib stream (stb picks its own filters; the
// pixels are the same).
void savePNG(const std::string& filename, const Image& img) {
    const size_t stride = (size_t)img.width * 3 + 1;
    std::vector<unsigned char> raw(stride * (size_t)img.height);
    for (int y = 0; y < img.height; y++) {
        unsigned char* row = &raw[(size_t)y * stride];
        row[0] = 0;
        for (int i = 0; i < img.width * 3; i++) {
            const float v = img.rgb[(size_t)y * img.width * 3 + i];
            const float srgb = v <= 0.0031308f ? 12.92f * v : (1.0f + 0.055f) * std::pow(v, 1.0f / 2.4f) - 0.055f;
            const float c = std::min(std::max(255.f * srgb, 0.f), 255.f);
            row[1 + i] = (unsigned char)c;
        }
    }
    uLongf zlen = compressBound((uLong)raw.size());
    std::vector<unsigned char> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK)
        throw Exception("Could not compress \"%s\"", filename);
    z.resize(zlen);
    std::ofstream out(filename, std::ios::binary);
    if (!out) throw Exception("Could not write \"%s\"", filename);
    auto chunk = [&](const char* type, const unsigned char* data, size_t n) {
        const unsigned char len[4] = {(unsigned char)(n >> 24), (unsigned char)(n >> 16), (unsigned char)(n >> 8),
                                      (unsigned char)n};
        out.write(reinterpret_cast<const char*>(len), 4);
        out.write(type, 4);
        if (n) out.write(reinterpret_cast<const char*>(data), (std::streamsize)n);
        uLong crc = crc32(0L, reinterpret_cast<const Bytef*>(type), 4);
        if (n) crc = crc32(crc, data, (uInt)n);
        const unsigned char c[4] = {(unsigned char)(crc >> 24), (unsigned char)(crc >> 16), (unsigned char)(crc >> 8),
                                    (unsigned char)crc};
        out.write(reinterpret_cast<const char*>(c), 4);
    };
    static const unsigned char sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    out.write(reinterpret_cast<const char*>(sig), 8);
    const uint32_t w = (uint32_t)img.width, h = (uint32_t)img.height;
    const unsigned char ihdr[13] = {(unsigned char)(w >> 24), (unsigned char)(w >> 16), (unsigned char)(w >> 8),
                                    (unsigned char)w, (unsigned char)(h >> 24), (unsigned char)(h >> 16),
                                    (unsigned char)(h >> 8), (unsigned char)h, 8, 2, 0, 0, 0};
    chunk("IHDR", ihdr, 13);
    chunk("IDAT", z.data(), z.size());
    chunk("IEND", nullptr, 0);
    if (!out) throw Exception("Could not write \"%s\"", filename);
}

void savePFM(const std::string& filename, const Image& img) {
    std::ofstream out(filename, std::ios::binary);
    if (!out) throw Exception("Could not write \"%s\"", filename);
    out << "PF\n" << img.width << " " << img.height << "\n-1.0\n";
    for (int y = img.height - 1; y >= 0; y--)
        out.write(reinterpret_cast<const char*>(&img.rgb[(size_t)y * img.width * 3]), (std::streamsize)img.width * 12);
}

void saveEXR(const std::string& filename, const Image& img) {
    std::vector<unsigned char> h;
    auto put = [&](const void* p, size_t n) {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        h.insert(h.end(), c, c + n);
    };
    auto put_i32 = [&](int32_t v) { put(&v, 4); };
    auto put_str = [&](const char* s) { put(s, std::strlen(s) + 1); };
    const uint32_t magic = 20000630u, version = 2;
    put(&magic, 4);
    put(&version, 4);
    // channels: B, G, R (alphabetical), FLOAT
    put_str("channels");
    put_str("chlist");
    put_i32(3 * (2 + 16) + 1);
    for (const char* n : {"B", "G", "R"}) {
        put_str(n);
        put_i32(2);
        const unsigned char zero4[4] = {0, 0, 0, 0};
        put(zero4, 4);
        put_i32(1);
        put_i32(1);
    }
    h.push_back(0);
    put_str("compression");
    put_str("compression");
    put_i32(1);
    h.push_back(0);
    const int32_t box[4] = {0, 0, img.width - 1, img.height - 1};
    put_str("dataWindow");
    put_str("box2i");
    put_i32(16);
    put(box, 16);
    put_str("displayWindow");
    put_str("box2i");
    put_i32(16);
    put(box, 16);
    put_str("lineOrder");
    put_str("lineOrder");
    put_i32(1);
    h.push_back(0);
    const float one = 1.f, zero2[2] = {0.f, 0.f};
    put_str("pixelAspectRatio");
    put_str("float");
    put_i32(4);
    put(&one, 4);
    put_str("screenWindowCenter");
    put_str("v2f");
    put_i32(8);
    put(zero2, 8);
    put_str("screenWindowWidth");
    put_str("float");
    put_i32(4);
    put(&one, 4);
    h.push_back(0);
    const size_t row_bytes = (size_t)img.width * 12;
    uint64_t off = h.size() + (size_t)img.height * 8;
    for (int y = 0; y < img.height; y++) {
        put(&off, 8);
        off += 8 + row_bytes;
    }
    std::ofstream out(filename, std::ios::binary);
    if (!out) throw Exception("Could not write \"%s\"", filename);
    out.write(reinterpret_cast<const char*>(h.data()), (std::streamsize)h.size());
    std::vector<float> row((size_t)img.width * 3);
    for (int y = 0; y < img.height; y++) {
        const int32_t yy = y, sz = (int32_t)row_bytes;
        out.write(reinterpret_cast<const char*>(&yy), 4);
        out.write(reinterpret_cast<const char*>(&sz), 4);
        for (int c = 0; c < 3; c++)  // B, G, R planes
            for (int x = 0; x < img.width; x++) row[(size_t)c * img.width + x] = img.rgb[((size_t)y * img.width + x) * 3 + (2 - c)];
        out.write(reinterpret_cast<const char*>(row.data()), (std::streamsize)row_bytes);
    }
}

}  // namespace drawlab
