// Host codecs for lattice configurations:
//  * the reference's canonical SU(2) text format YM-SU2-CONFIG-V1 (one
//    header line, extents, link count, then one line per link with the four
//    components' IEEE-754 bit patterns as 16 lowercase hex digits), encoded
//    and decoded exactly, with the reference's validation order;
//  * a binary checkpoint of an SU(2) or SU(3) lattice in its storage
//    precision plus the chain key and the next sweep index.  The device RNG
//    is counter-based (key, sweep, site, direction, hit), so links + key +
//    next sweep are the whole chain state: a chain resumed from a checkpoint
//    continues bit for bit.
// Both work on host buffers in the reference AoS FP64 layout (what
// ym_lattice_download returns), so they need no device.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ym_amd.h"

namespace {

constexpr const char *kSu2Format = "YM-SU2-CONFIG-V1";  // lib.rs:23
constexpr double kSerializationTol = 2.0e-12;           // lib.rs:20
constexpr uint64_t kMaxReferenceLinks = 1ull << 20;     // lib.rs:21

constexpr char kCkptMagic[16] = {'Y', 'M', '-', 'C', 'K', 'P', 'T', '-', 'V', '1', '\n', 0, 0, 0, 0, 0};
constexpr int64_t kCkptHeader = 16 + 4 + 4 + 4 * 8 + 5 * 8;  // magic, group, precision, extents, 5 u64

bool is_fin(double v) { return std::isfinite(v); }

// Sum of squares in component order, as the reference's iterator sum.
double norm_sq(const double *c) {
    double s = 0.0;
    for (int k = 0; k < 4; ++k) s += c[k] * c[k];
    return s;
}

// Rust's usize::from_str: optional '+', then one or more ASCII digits, no overflow.
bool parse_usize(const std::string &s, uint64_t *out) {
    size_t i = 0;
    if (i < s.size() && s[i] == '+') ++i;
    if (i == s.size()) return false;
    uint64_t v = 0;
    for (; i < s.size(); ++i) {
        const char ch = s[i];
        if (ch < '0' || ch > '9') return false;
        const uint64_t d = (uint64_t)(ch - '0');
        if (v > (UINT64_MAX - d) / 10) return false;
        v = v * 10 + d;
    }
    *out = v;
    return true;
}

// str::lines: split on '\n', drop one '\r' before it, no empty line after a final '\n'.
std::vector<std::string> split_lines(const char *text, int64_t len) {
    std::vector<std::string> lines;
    int64_t start = 0;
    for (int64_t i = 0; i < len; ++i) {
        if (text[i] == '\n') {
            int64_t end = i;
            if (end > start && text[end - 1] == '\r') --end;
            lines.emplace_back(text + start, (size_t)(end - start));
            start = i + 1;
        }
    }
    if (start < len) {
        int64_t end = len;
        lines.emplace_back(text + start, (size_t)(end - start));
    }
    return lines;
}

std::vector<std::string> split(const std::string &s, char sep) {
    std::vector<std::string> parts;
    size_t start = 0;
    for (size_t i = 0; i <= s.size(); ++i)
        if (i == s.size() || s[i] == sep) {
            parts.emplace_back(s, start, i - start);
            start = i + 1;
        }
    return parts;
}

std::string encode_su2(const int64_t n[4], const double *aos, uint64_t links) {
    std::string out;
    out.reserve(64 + links * 68);
    char line[128];
    out += kSu2Format;
    out += '\n';
    std::snprintf(line, sizeof line, "extents=%lld,%lld,%lld,%lld\n", (long long)n[0], (long long)n[1],
                  (long long)n[2], (long long)n[3]);
    out += line;
    std::snprintf(line, sizeof line, "link_count=%llu\n", (unsigned long long)links);
    out += line;
    for (uint64_t l = 0; l < links; ++l) {
        uint64_t b[4];
        std::memcpy(b, aos + 4 * l, sizeof b);
        std::snprintf(line, sizeof line, "%016llx,%016llx,%016llx,%016llx\n", (unsigned long long)b[0],
                      (unsigned long long)b[1], (unsigned long long)b[2], (unsigned long long)b[3]);
        out += line;
    }
    return out;
}

// Lattice4::new_identity validation (lib.rs:242-256).
int32_t su2_link_count(const int64_t n[4], uint64_t *links) {
    for (int d = 0; d < 4; ++d)
        if (n[d] < 2) return YM_INVALID_EXTENT;
    unsigned __int128 s = 1;
    for (int d = 0; d < 4; ++d) {
        s *= (unsigned __int128)n[d];
        if (s > (unsigned __int128)UINT64_MAX) return YM_LATTICE_SIZE_OVERFLOW;
    }
    s *= 4;
    if (s > (unsigned __int128)UINT64_MAX) return YM_LATTICE_SIZE_OVERFLOW;
    if ((uint64_t)s > kMaxReferenceLinks) return YM_LATTICE_TOO_LARGE;
    *links = (uint64_t)s;
    return YM_OK;
}

// sites of the largest lattice a handle can be created with (ym_lattice_create)
constexpr uint64_t kCheckpointMaxSites = (uint64_t)1 << 30;

uint64_t fnv1a64(const unsigned char *p, int64_t n, uint64_t h = 0xcbf29ce484222325ull) {
    for (int64_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 0x100000001b3ull;
    }
    return h;
}

int32_t ckpt_sizes(int32_t group, int32_t precision, const int64_t n[4], uint64_t *links, int64_t *payload) {
    if ((group != 2 && group != 3) || (precision != 32 && precision != 64) || !n) return YM_INVALID_ARGUMENT;
    for (int d = 0; d < 4; ++d)
        if (n[d] < 2) return YM_INVALID_EXTENT;
    // checked after every factor (four extents of 2^32 would wrap a 128-bit
    // product), and the volume capped at what a lattice handle can hold
    unsigned __int128 v = 1;
    for (int d = 0; d < 4; ++d) {
        v *= (unsigned __int128)n[d];
        if (v > (unsigned __int128)kCheckpointMaxSites) return YM_LATTICE_TOO_LARGE;
    }
    const unsigned __int128 s = 4 * v;
    const unsigned __int128 bytes = s * (group == 3 ? 18 : 4) * (precision / 8);
    if (bytes > (unsigned __int128)INT64_MAX / 2) return YM_LATTICE_SIZE_OVERFLOW;
    *links = (uint64_t)s;
    *payload = (int64_t)bytes;
    return YM_OK;
}

template <typename T>
void put(unsigned char *&p, T v) {
    std::memcpy(p, &v, sizeof v);
    p += sizeof v;
}
template <typename T>
T get(const unsigned char *&p) {
    T v;
    std::memcpy(&v, p, sizeof v);
    p += sizeof v;
    return v;
}

}  // namespace

extern "C" {

// lib.rs:587-614 to_canonical_config on the FP64 AoS link array (4 doubles
// per link).  buf == NULL (or too small) only reports the length.
int32_t ym_su2_config_encode(const int64_t extents[4], const double *aos, char *buf, int64_t capacity,
                             int64_t *length) {
    if (!extents || !aos || !length) return YM_INVALID_ARGUMENT;
    uint64_t links;
    int32_t st = su2_link_count(extents, &links);
    if (st) return st;
    double worst = 0.0;
    for (uint64_t l = 0; l < links; ++l) {
        const double ns = norm_sq(aos + 4 * l);
        const double r = std::fabs(ns - 1.0);
        if (!is_fin(ns) || !is_fin(r)) {
            worst = INFINITY;
            break;
        }
        worst = std::fmax(worst, r);
    }
    if (worst > kSerializationTol) return YM_SU2_NON_UNIT;
    const std::string s = encode_su2(extents, aos, links);
    *length = (int64_t)s.size();
    if (buf && capacity >= (int64_t)s.size()) std::memcpy(buf, s.data(), s.size());
    else if (buf) return YM_INVALID_ARGUMENT;
    return YM_OK;
}

// lib.rs:616-694 from_canonical_config: extents first (so the caller can
// size `aos`), then the links; any deviation from the canonical encoding is
// YM_INVALID_CONFIGURATION_ENCODING.  aos == NULL only validates the header
// lines and returns the extents.
int32_t ym_su2_config_decode(const char *text, int64_t length, int64_t extents[4], double *aos, int64_t capacity) {
    if (!text || length < 0 || !extents) return YM_INVALID_ARGUMENT;
    const std::vector<std::string> lines = split_lines(text, length);
    size_t li = 0;
    if (lines.size() <= li || lines[li++] != kSu2Format) return YM_INVALID_CONFIGURATION_ENCODING;
    if (lines.size() <= li) return YM_INVALID_CONFIGURATION_ENCODING;
    const std::string &el = lines[li++];
    if (el.compare(0, 8, "extents=") != 0) return YM_INVALID_CONFIGURATION_ENCODING;
    const std::vector<std::string> ep = split(el.substr(8), ',');
    if (ep.size() < 4) return YM_INVALID_CONFIGURATION_ENCODING;
    int64_t n[4];
    for (int d = 0; d < 4; ++d) {
        uint64_t v;
        if (!parse_usize(ep[d], &v)) return YM_INVALID_CONFIGURATION_ENCODING;
        if (v > (uint64_t)INT64_MAX) return YM_LATTICE_SIZE_OVERFLOW;
        n[d] = (int64_t)v;
    }
    if (ep.size() > 4) return YM_INVALID_CONFIGURATION_ENCODING;
    if (lines.size() <= li) return YM_INVALID_CONFIGURATION_ENCODING;
    const std::string &cl = lines[li++];
    if (cl.compare(0, 11, "link_count=") != 0) return YM_INVALID_CONFIGURATION_ENCODING;
    uint64_t declared;
    if (!parse_usize(cl.substr(11), &declared)) return YM_INVALID_CONFIGURATION_ENCODING;
    if (declared > kMaxReferenceLinks) return YM_LATTICE_TOO_LARGE;
    uint64_t links;
    int32_t st = su2_link_count(n, &links);
    if (st) return st;
    if (declared != links) return YM_INVALID_CONFIGURATION_ENCODING;
    for (int d = 0; d < 4; ++d) extents[d] = n[d];
    if (!aos) return YM_OK;
    if (capacity < (int64_t)(4 * links)) return YM_INVALID_ARGUMENT;
    std::vector<double> tmp((size_t)(4 * links));
    for (uint64_t l = 0; l < links; ++l) {
        if (lines.size() <= li) return YM_INVALID_CONFIGURATION_ENCODING;
        const std::vector<std::string> parts = split(lines[li++], ',');
        double c[4];
        for (int k = 0; k < 4; ++k) {
            if ((int)parts.size() <= k) return YM_INVALID_CONFIGURATION_ENCODING;
            const std::string &tok = parts[k];
            if (tok.size() != 16) return YM_INVALID_CONFIGURATION_ENCODING;
            uint64_t bits = 0;
            for (char ch : tok) {
                int v;
                if (ch >= '0' && ch <= '9') v = ch - '0';
                else if (ch >= 'a' && ch <= 'f') v = ch - 'a' + 10;
                else return YM_INVALID_CONFIGURATION_ENCODING;
                bits = (bits << 4) | (uint64_t)v;
            }
            std::memcpy(&c[k], &bits, sizeof bits);
        }
        if (parts.size() > 4) return YM_INVALID_CONFIGURATION_ENCODING;
        // Su2::from_preserved_unit_components (lib.rs:163-177)
        for (int k = 0; k < 4; ++k)
            if (!is_fin(c[k])) return YM_SU2_NON_FINITE;
        for (int k = 0; k < 4; ++k)
            if (std::fabs(c[k]) > 1.0 + kSerializationTol) return YM_SU2_NON_UNIT;
        const double ns = norm_sq(c);
        if (!is_fin(ns) || std::fabs(ns - 1.0) > kSerializationTol) return YM_SU2_NON_UNIT;
        std::memcpy(tmp.data() + 4 * l, c, sizeof c);
    }
    if (lines.size() > li) return YM_INVALID_CONFIGURATION_ENCODING;
    // the text must be the canonical encoding of what it decodes to
    const std::string canon = encode_su2(n, tmp.data(), links);
    if ((int64_t)canon.size() != length || std::memcmp(canon.data(), text, (size_t)length) != 0)
        return YM_INVALID_CONFIGURATION_ENCODING;
    std::memcpy(aos, tmp.data(), tmp.size() * sizeof(double));
    return YM_OK;
}

int32_t ym_checkpoint_bytes(int32_t group, int32_t precision, const int64_t extents[4], int64_t *bytes) {
    if (!bytes) return YM_INVALID_ARGUMENT;
    uint64_t links;
    int64_t payload;
    int32_t st = ckpt_sizes(group, precision, extents, &links, &payload);
    if (st) return st;
    *bytes = kCkptHeader + payload;
    return YM_OK;
}

int32_t ym_checkpoint_encode(int32_t group, int32_t precision, const int64_t extents[4], const double *aos,
                             uint64_t key, uint64_t next_sweep, void *buf, int64_t capacity, int64_t *length) {
    if (!aos || !buf || !length) return YM_INVALID_ARGUMENT;
    uint64_t links;
    int64_t payload;
    int32_t st = ckpt_sizes(group, precision, extents, &links, &payload);
    if (st) return st;
    if (capacity < kCkptHeader + payload) return YM_INVALID_ARGUMENT;
    unsigned char *base = (unsigned char *)buf;
    unsigned char *body = base + kCkptHeader;
    const uint64_t reals = links * (group == 3 ? 18u : 4u);
    if (precision == 64) {
        std::memcpy(body, aos, (size_t)payload);
    } else {
        for (uint64_t i = 0; i < reals; ++i) {
            const float f = (float)aos[i];
            std::memcpy(body + 4 * i, &f, 4);
        }
    }
    unsigned char *p = base;
    std::memcpy(p, kCkptMagic, 16);
    p += 16;
    put<uint32_t>(p, (uint32_t)group);
    put<uint32_t>(p, (uint32_t)precision);
    for (int d = 0; d < 4; ++d) put<int64_t>(p, extents[d]);
    put<uint64_t>(p, key);
    put<uint64_t>(p, next_sweep);
    put<uint64_t>(p, links);
    put<uint64_t>(p, (uint64_t)payload);
    // checksum over the header fields above and the payload
    const uint64_t h = fnv1a64(body, payload, fnv1a64(base, (int64_t)(p - base)));
    put<uint64_t>(p, h);
    *length = kCkptHeader + payload;
    return YM_OK;
}

int32_t ym_checkpoint_header(const void *buf, int64_t length, int32_t *group, int32_t *precision, int64_t extents[4],
                             uint64_t *key, uint64_t *next_sweep) {
    if (!buf || !group || !precision || !extents || !key || !next_sweep) return YM_INVALID_ARGUMENT;
    if (length < kCkptHeader) return YM_INVALID_CONFIGURATION_ENCODING;
    const unsigned char *base = (const unsigned char *)buf;
    if (std::memcmp(base, kCkptMagic, 16) != 0) return YM_INVALID_CONFIGURATION_ENCODING;
    const unsigned char *p = base + 16;
    const int32_t g = (int32_t)get<uint32_t>(p), pr = (int32_t)get<uint32_t>(p);
    int64_t n[4];
    for (int d = 0; d < 4; ++d) n[d] = get<int64_t>(p);
    const uint64_t k = get<uint64_t>(p), s = get<uint64_t>(p), links = get<uint64_t>(p), pay = get<uint64_t>(p);
    const int64_t fields = (int64_t)(p - base);
    const uint64_t h = get<uint64_t>(p);
    uint64_t want_links;
    int64_t want_pay;
    if (ckpt_sizes(g, pr, n, &want_links, &want_pay) != YM_OK) return YM_INVALID_CONFIGURATION_ENCODING;
    if (links != want_links || pay != (uint64_t)want_pay || length != kCkptHeader + want_pay)
        return YM_INVALID_CONFIGURATION_ENCODING;
    if (fnv1a64(base + kCkptHeader, want_pay, fnv1a64(base, fields)) != h) return YM_INVALID_CONFIGURATION_ENCODING;
    *group = g;
    *precision = pr;
    for (int d = 0; d < 4; ++d) extents[d] = n[d];
    *key = k;
    *next_sweep = s;
    return YM_OK;
}

int32_t ym_checkpoint_decode(const void *buf, int64_t length, double *aos, int64_t count) {
    int32_t g, pr;
    int64_t n[4];
    uint64_t k, s;
    int32_t st = ym_checkpoint_header(buf, length, &g, &pr, n, &k, &s);
    if (st) return st;
    if (!aos) return YM_INVALID_ARGUMENT;
    uint64_t links;
    int64_t payload;
    ckpt_sizes(g, pr, n, &links, &payload);
    const uint64_t reals = links * (g == 3 ? 18u : 4u);
    if (count != (int64_t)reals) return YM_INVALID_ARGUMENT;
    const unsigned char *body = (const unsigned char *)buf + kCkptHeader;
    if (pr == 64) {
        std::memcpy(aos, body, (size_t)payload);
    } else {
        for (uint64_t i = 0; i < reals; ++i) {
            float f;
            std::memcpy(&f, body + 4 * i, 4);
            aos[i] = (double)f;
        }
    }
    return YM_OK;
}

}  // extern "C"
