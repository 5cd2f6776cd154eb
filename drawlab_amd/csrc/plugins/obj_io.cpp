// Wavefront OBJ reading for the plugin layer's "obj" mesh
// (src/tracer/mesh/obj.cpp:17-87), and a binary cache of its result.
//
// Parsing: the file is read whole and scanned line by line with
// whitespace-separated tokens ("v x y z", "vt u v", "vn x y z",
// "f v[/t[/n]] ..."; other tags ignored).  Floats are rounded correctly
// (to nearest, ties to even), as the stream extraction of the reference's
// loader rounds them: a fast exact path for short decimals, std::from_chars
// otherwise; a missing coordinate reads as 0.  Face
// indices are 1-based or negative (relative to the end of the list read so
// far); corners are de-indexed per distinct (v, t, n) in order of first
// use, polygons fanned into triangles.  The distinct corners are found with
// an open-addressing hash table.
//
// Cache: when the environment variable DRAWLAB_MESH_CACHE names a
// directory, the parsed arrays (before any transform) are stored there as
// <16 hex digits>.dlmesh, keyed by the file's absolute path, size and
// modification time, and read back instead of parsing while those match.
// Layout (little-endian): "DLOBJC01", u64 size, i64 mtime_ns, u32 path
// length + path bytes, u64 corners, u64 triangles, u32 has_normals,
// u32 has_texcoords, then positions (3 floats per corner), normals (3, if
// any), texcoords (2, if any), indices (3 u32 per triangle).
#include <sys/stat.h>
#include <unistd.h>

#include <charconv>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <vector>

#include "drawlab/plugins.h"

namespace drawlab {

namespace {

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\f' || c == '\v'; }

// Decimal [b, e) -> nearest float.  Fast path (Clinger): up to 19 digits
// whose value m is below 2^53 and a decimal exponent within +-22 give the
// correctly rounded double m * 10^k or m / 10^-k (both operands exact);
// rounding that double to float is then correct unless the double lies
// exactly halfway between two floats (rounding to nearest cannot step over
// a midpoint, which is a double), and that case, like every other form,
// goes to std::from_chars.
float parse_float(const char* b, const char* e) {
    static const double kPow10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8,  1e9,  1e10, 1e11,
                                      1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
    const char* s = b;
    if (s < e && *s == '+') ++s;
    const char* q = s;
    bool neg = false;
    if (q < e && *q == '-') {
        neg = true;
        ++q;
    }
    uint64_t m = 0;
    int digits = 0, exp10 = 0;
    bool any = false;
    while (q < e && *q >= '0' && *q <= '9') {
        if (digits < 19) {
            m = m * 10 + (uint64_t)(*q - '0');
            if (m) ++digits;
        } else {
            ++exp10;
        }
        any = true;
        ++q;
    }
    if (q < e && *q == '.') {
        ++q;
        while (q < e && *q >= '0' && *q <= '9') {
            if (digits < 19) {
                m = m * 10 + (uint64_t)(*q - '0');
                if (m) ++digits;
                --exp10;
            }
            any = true;
            ++q;
        }
    }
    bool fast = any && digits < 19;
    if (fast && q < e && (*q == 'e' || *q == 'E')) {
        const char* r = q + 1;
        bool eneg = false;
        if (r < e && (*r == '+' || *r == '-')) eneg = *r++ == '-';
        int x = 0, nd = 0;
        while (r < e && *r >= '0' && *r <= '9' && nd < 6) {
            x = x * 10 + (*r++ - '0');
            ++nd;
        }
        if (nd == 0 || (r < e && *r >= '0' && *r <= '9')) fast = false;
        exp10 += eneg ? -x : x;
        q = r;
    }
    if (fast && m < (1ull << 53) && exp10 >= -22 && exp10 <= 22) {
        const double d = exp10 >= 0 ? (double)m * kPow10[exp10] : (double)m / kPow10[-exp10];
        const float f = (float)d;
        const double back = (double)f;
        bool tie = false;
        if (back != d) {
            const double other = (double)std::nextafter(f, back < d ? INFINITY : -INFINITY);
            tie = (d - back) == (other - d);
        }
        if (!tie && std::isfinite(f)) return neg ? -f : f;
    }
    float v = 0.0f;
    if (s < e) std::from_chars(s, e, v);
    return v;
}

struct Cursor {
    const char* p;
    const char* end;
    void skip() {
        while (p < end && is_space(*p)) ++p;
    }
    // next whitespace-separated token of the line; empty at its end
    std::pair<const char*, const char*> token() {
        skip();
        const char* b = p;
        while (p < end && !is_space(*p)) ++p;
        return {b, p};
    }
    float number() {
        auto t = token();
        return t.first < t.second ? parse_float(t.first, t.second) : 0.0f;
    }
};

// Leading integer of [b, e) (optional sign); false when there is none.
bool parse_int(const char* b, const char* e, long& out) {
    if (b < e && *b == '+') ++b;
    auto r = std::from_chars(b, e, out);
    return r.ec == std::errc() && r.ptr != b;
}

// Open-addressing map (v, t, n) -> corner index.
class CornerMap {
public:
    CornerMap() { resize(1u << 16); }
    // index of the corner (v, t, n), adding it as `next` when new
    uint32_t find_or_add(int v, int t, int n, uint32_t next, bool& added) {
        if ((size_t)(count_ + 1) * 2 > keys_.size() / 3) grow();
        const size_t mask = vals_.size() - 1;
        size_t i = hash(v, t, n) & mask;
        for (;;) {
            if (vals_[i] == kEmpty) {
                keys_[3 * i] = v;
                keys_[3 * i + 1] = t;
                keys_[3 * i + 2] = n;
                vals_[i] = next;
                ++count_;
                added = true;
                return next;
            }
            if (keys_[3 * i] == v && keys_[3 * i + 1] == t && keys_[3 * i + 2] == n) {
                added = false;
                return vals_[i];
            }
            i = (i + 1) & mask;
        }
    }

private:
    static constexpr uint32_t kEmpty = 0xffffffffu;
    static size_t hash(int v, int t, int n) {
        uint64_t h = (uint64_t)(uint32_t)v * 0x9E3779B97F4A7C15ull;
        h ^= ((uint64_t)(uint32_t)t + 0x632BE59BD9B4E019ull) * 0xC2B2AE3D27D4EB4Full;
        h ^= ((uint64_t)(uint32_t)n + 0x165667B19E3779F9ull) * 0x85EBCA77C2B2AE63ull;
        return (size_t)(h ^ (h >> 29));
    }
    void resize(size_t slots) {
        keys_.assign(3 * slots, 0);
        vals_.assign(slots, kEmpty);
        count_ = 0;
    }
    void grow() {
        std::vector<int> k = std::move(keys_);
        std::vector<uint32_t> v = std::move(vals_);
        resize(v.size() * 2);
        const size_t mask = vals_.size() - 1;
        for (size_t j = 0; j < v.size(); j++) {
            if (v[j] == kEmpty) continue;
            size_t i = hash(k[3 * j], k[3 * j + 1], k[3 * j + 2]) & mask;
            while (vals_[i] != kEmpty) i = (i + 1) & mask;
            keys_[3 * i] = k[3 * j];
            keys_[3 * i + 1] = k[3 * j + 1];
            keys_[3 * i + 2] = k[3 * j + 2];
            vals_[i] = v[j];
            ++count_;
        }
    }
    std::vector<int> keys_;
    std::vector<uint32_t> vals_;
    size_t count_ = 0;
};

ObjData parse_obj(const std::string& filename) {
    std::ifstream in(filename, std::ios::binary | std::ios::ate);
    if (!in) throw Exception("Loadobj fail: could not open \"%s\"", filename);
    std::vector<char> buf((size_t)in.tellg());
    in.seekg(0);
    if (!buf.empty() && !in.read(buf.data(), (std::streamsize)buf.size()))
        throw Exception("Loadobj fail: could not read \"%s\"", filename);
    std::vector<float> P, N, T;
    struct Corner {
        int v, t, n;
    };
    std::vector<Corner> corners;
    constexpr uint32_t kNone = 0xffffffffu;
    std::vector<uint32_t> first;  // position -> its first corner
    CornerMap map;
    ObjData out;
    std::vector<uint32_t> poly;
    const char* p = buf.data();
    const char* const end = p + buf.size();
    int lineno = 0;
    auto fix = [&](long idx, size_t count) -> int {
        if (idx > 0) return idx - 1 > INT_MAX ? INT_MAX : (int)(idx - 1);
        if (idx < 0) return (long)count + idx < INT_MIN ? INT_MIN : (int)((long)count + idx);
        throw Exception("Loadobj fail: index 0 at %s:%i", filename, lineno);
    };
    while (p < end) {
        const char* nl = static_cast<const char*>(std::memchr(p, '\n', (size_t)(end - p)));
        const char* le = nl ? nl : end;
        ++lineno;
        Cursor c{p, le};
        p = nl ? nl + 1 : end;
        auto tag = c.token();
        const size_t tl = (size_t)(tag.second - tag.first);
        if (tl == 1 && tag.first[0] == 'v') {
            const float x = c.number(), y = c.number(), z = c.number();
            P.insert(P.end(), {x, y, z});
        } else if (tl == 2 && tag.first[0] == 'v' && tag.first[1] == 'n') {
            const float x = c.number(), y = c.number(), z = c.number();
            N.insert(N.end(), {x, y, z});
        } else if (tl == 2 && tag.first[0] == 'v' && tag.first[1] == 't') {
            const float u = c.number(), v = c.number();
            T.insert(T.end(), {u, v});
        } else if (tl == 1 && tag.first[0] == 'f') {
            poly.clear();
            for (;;) {
                auto tok = c.token();
                if (tok.first == tok.second) break;
                const char* s1 = static_cast<const char*>(std::memchr(tok.first, '/', (size_t)(tok.second - tok.first)));
                long iv = 0;
                if (!parse_int(tok.first, s1 ? s1 : tok.second, iv))
                    throw Exception("Loadobj fail: bad face index at %s:%i", filename, lineno);
                int kv = fix(iv, P.size() / 3), kt = -1, kn = -1;
                if (s1) {
                    const char* s2 =
                        static_cast<const char*>(std::memchr(s1 + 1, '/', (size_t)(tok.second - (s1 + 1))));
                    const char* te = s2 ? s2 : tok.second;
                    long it = 0, in_ = 0;
                    if (te > s1 + 1) {
                        if (!parse_int(s1 + 1, te, it))
                            throw Exception("Loadobj fail: bad face index at %s:%i", filename, lineno);
                        kt = fix(it, T.size() / 2);
                    }
                    if (s2 && s2 + 1 < tok.second) {
                        if (!parse_int(s2 + 1, tok.second, in_))
                            throw Exception("Loadobj fail: bad face index at %s:%i", filename, lineno);
                        kn = fix(in_, N.size() / 3);
                    }
                }
                if (kv < 0 || kv >= (int)(P.size() / 3) || kt >= (int)(T.size() / 2) || kn >= (int)(N.size() / 3))
                    throw Exception("Loadobj fail: index out of range at %s:%i", filename, lineno);
                // the first corner of each position answers most lookups
                // (one corner per position is the common case); the hash
                // table holds the rest
                if (first.size() < P.size() / 3) first.resize(P.size() / 3, kNone);
                uint32_t id = first[(size_t)kv];
                if (id == kNone) {
                    id = (uint32_t)corners.size();
                    first[(size_t)kv] = id;
                    corners.push_back(Corner{kv, kt, kn});
                } else if (corners[id].t != kt || corners[id].n != kn) {
                    bool added;
                    id = map.find_or_add(kv, kt, kn, (uint32_t)corners.size(), added);
                    if (added) corners.push_back(Corner{kv, kt, kn});
                }
                poly.push_back(id);
            }
            for (size_t i = 1; i + 1 < poly.size(); ++i) out.F.insert(out.F.end(), {poly[0], poly[i], poly[i + 1]});
        }
    }
    bool has_n = !N.empty(), has_t = !T.empty();
    for (const Corner& k : corners) {
        if (k.n < 0) has_n = false;
        if (k.t < 0) has_t = false;
    }
    out.P.resize(3 * corners.size());
    if (has_n) out.N.resize(3 * corners.size());
    if (has_t) out.UV.resize(2 * corners.size());
    for (size_t i = 0; i < corners.size(); i++) {
        const Corner& k = corners[i];
        std::memcpy(&out.P[3 * i], &P[3 * (size_t)k.v], 3 * sizeof(float));
        if (has_n) std::memcpy(&out.N[3 * i], &N[3 * (size_t)k.n], 3 * sizeof(float));
        if (has_t) std::memcpy(&out.UV[2 * i], &T[2 * (size_t)k.t], 2 * sizeof(float));
    }
    return out;
}

// ---- binary cache ------------------------------------------------------------
constexpr char kMagic[8] = {'D', 'L', 'O', 'B', 'J', 'C', '0', '1'};

uint64_t fnv1a(const void* data, size_t n, uint64_t h = 1469598103934665603ull) {
    const unsigned char* b = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < n; i++) {
        h ^= b[i];
        h *= 1099511628211ull;
    }
    return h;
}

struct SourceId {
    std::string path;
    uint64_t size = 0;
    int64_t mtime_ns = 0;
};

bool source_id(const std::string& filename, SourceId& id) {
    char* real = realpath(filename.c_str(), nullptr);
    if (!real) return false;
    id.path = real;
    std::free(real);
    struct stat st;
    if (stat(id.path.c_str(), &st) != 0) return false;
    id.size = (uint64_t)st.st_size;
    id.mtime_ns = (int64_t)st.st_mtim.tv_sec * 1000000000ll + (int64_t)st.st_mtim.tv_nsec;
    return true;
}

std::string cache_file(const std::string& dir, const SourceId& id) {
    uint64_t h = fnv1a(id.path.data(), id.path.size());
    h = fnv1a(&id.size, sizeof(id.size), h);
    h = fnv1a(&id.mtime_ns, sizeof(id.mtime_ns), h);
    char name[40];
    std::snprintf(name, sizeof(name), "%016llx.dlmesh", (unsigned long long)h);
    return dir + "/" + name;
}

template <typename T>
bool read_vec(std::ifstream& f, std::vector<T>& v, size_t n) {
    v.resize(n);
    return n == 0 || (bool)f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(T)));
}

bool read_cache(const std::string& file, const SourceId& id, ObjData& out) {
    std::ifstream f(file, std::ios::binary);
    if (!f) return false;
    char magic[8];
    uint64_t size = 0, nc = 0, nt = 0;
    int64_t mt = 0;
    uint32_t plen = 0, has_n = 0, has_t = 0;
    if (!f.read(magic, 8) || std::memcmp(magic, kMagic, 8) != 0) return false;
    if (!f.read(reinterpret_cast<char*>(&size), 8) || !f.read(reinterpret_cast<char*>(&mt), 8)) return false;
    if (!f.read(reinterpret_cast<char*>(&plen), 4) || plen != id.path.size()) return false;
    std::string path(plen, '\0');
    if (!f.read(&path[0], plen) || path != id.path || size != id.size || mt != id.mtime_ns) return false;
    if (!f.read(reinterpret_cast<char*>(&nc), 8) || !f.read(reinterpret_cast<char*>(&nt), 8)) return false;
    if (!f.read(reinterpret_cast<char*>(&has_n), 4) || !f.read(reinterpret_cast<char*>(&has_t), 4)) return false;
    if (nc > (1ull << 32) || nt > (1ull << 32) || has_n > 1 || has_t > 1) return false;
    if (!read_vec(f, out.P, 3 * nc) || !read_vec(f, out.N, has_n ? 3 * nc : 0) ||
        !read_vec(f, out.UV, has_t ? 2 * nc : 0) || !read_vec(f, out.F, 3 * nt))
        return false;
    for (uint32_t i : out.F)
        if (i >= nc) return false;
    return true;
}

void write_cache(const std::string& file, const SourceId& id, const ObjData& d) {
    const std::string tmp = file + ".tmp." + std::to_string((long)getpid());
    {
        std::ofstream f(tmp, std::ios::binary);
        if (!f) return;
        const uint64_t nc = d.P.size() / 3, nt = d.F.size() / 3;
        const uint32_t plen = (uint32_t)id.path.size(), has_n = d.N.empty() ? 0 : 1, has_t = d.UV.empty() ? 0 : 1;
        f.write(kMagic, 8);
        f.write(reinterpret_cast<const char*>(&id.size), 8);
        f.write(reinterpret_cast<const char*>(&id.mtime_ns), 8);
        f.write(reinterpret_cast<const char*>(&plen), 4);
        f.write(id.path.data(), plen);
        f.write(reinterpret_cast<const char*>(&nc), 8);
        f.write(reinterpret_cast<const char*>(&nt), 8);
        f.write(reinterpret_cast<const char*>(&has_n), 4);
        f.write(reinterpret_cast<const char*>(&has_t), 4);
        f.write(reinterpret_cast<const char*>(d.P.data()), (std::streamsize)(d.P.size() * sizeof(float)));
        f.write(reinterpret_cast<const char*>(d.N.data()), (std::streamsize)(d.N.size() * sizeof(float)));
        f.write(reinterpret_cast<const char*>(d.UV.data()), (std::streamsize)(d.UV.size() * sizeof(float)));
        f.write(reinterpret_cast<const char*>(d.F.data()), (std::streamsize)(d.F.size() * sizeof(uint32_t)));
        if (!f) {
            f.close();
            std::remove(tmp.c_str());
            return;
        }
    }
    if (std::rename(tmp.c_str(), file.c_str()) != 0) std::remove(tmp.c_str());
}

}  // namespace

ObjData loadOBJ(const std::string& filename, bool* from_cache) {
    if (from_cache) *from_cache = false;
    const char* dir = std::getenv("DRAWLAB_MESH_CACHE");
    SourceId id;
    if (!dir || !*dir || !source_id(filename, id)) return parse_obj(filename);
    const std::string file = cache_file(dir, id);
    ObjData d;
    if (read_cache(file, id, d)) {
        if (from_cache) *from_cache = true;
        return d;
    }
    d = parse_obj(filename);
    write_cache(file, id, d);
    return d;
}

}  // namespace drawlab
