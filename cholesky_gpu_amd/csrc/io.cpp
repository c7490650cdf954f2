// io.cpp -- host-side readers: the reference's settings.json options file and
// the big-endian idx files of the MNIST data set.
//
// settings.json (src/settings.json:1-14 of the reference) holds two objects,
// "matrix_settings" {n, condition_number, dist_type} and
// "factorization_settings" {block_size, eps_prime, floor, diag_pert, left},
// every value written as a JSON string.  The reference's reader
// (src/json_utils.hpp:2-32) looks the values up under other spellings
// ("matrix settings", "condition number") and strips the quotes of the dumped
// string; this reader accepts both spellings and either a string or a number
// for each value.  The JSON subset parsed here is the full grammar of objects,
// arrays, strings (with escapes), numbers, true/false/null; nothing else.
//
// idx (the MNIST container): magic 0x00 0x00 <type> <ndim>, then ndim
// big-endian uint32 dimensions, then the data in row-major order.  Only the
// unsigned-byte type (0x08) occurs in MNIST: labels are 1-D (magic 0x0801),
// images 3-D (0x0803, count x rows x cols).
#include "mpc_internal.hpp"
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>

namespace {

// ---------------------------------------------------------------- JSON
struct jvalue {
    enum kind_t { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<jvalue> arr;
    std::map<std::string, jvalue> obj;
};

struct jparser {
    const char* p;
    const char* end;
    std::string err;
    int depth = 0;

    void ws() {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    }
    bool fail(const char* what) {
        if (err.empty()) err = what;
        return false;
    }
    bool lit(const char* s) {
        const size_t n = std::strlen(s);
        if ((size_t)(end - p) < n || std::strncmp(p, s, n) != 0) return false;
        p += n;
        return true;
    }
    static void put_utf8(std::string& out, unsigned cp) {
        if (cp < 0x80) out += (char)cp;
        else if (cp < 0x800) { out += (char)(0xC0 | (cp >> 6)); out += (char)(0x80 | (cp & 63)); }
        else if (cp < 0x10000) {
            out += (char)(0xE0 | (cp >> 12));
            out += (char)(0x80 | ((cp >> 6) & 63));
            out += (char)(0x80 | (cp & 63));
        } else {
            out += (char)(0xF0 | (cp >> 18));
            out += (char)(0x80 | ((cp >> 12) & 63));
            out += (char)(0x80 | ((cp >> 6) & 63));
            out += (char)(0x80 | (cp & 63));
        }
    }
    bool hex4(unsigned& v) {
        if (end - p < 4) return fail("truncated \\u escape");
        v = 0;
        for (int i = 0; i < 4; ++i) {
            const char c = *p++;
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (unsigned)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (unsigned)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (unsigned)(c - 'A' + 10);
            else return fail("bad hex digit in \\u escape");
        }
        return true;
    }
    bool string(std::string& out) {
        if (p >= end || *p != '"') return fail("expected a string");
        ++p;
        while (p < end && *p != '"') {
            const unsigned char c = (unsigned char)*p;
            if (c < 0x20) return fail("control character in a string");
            if (c != '\\') { out += (char)c; ++p; continue; }
            if (++p >= end) return fail("truncated escape");
            const char e = *p++;
            switch (e) {
            case '"': out += '"'; break;
            case '\\': out += '\\'; break;
            case '/': out += '/'; break;
            case 'b': out += '\b'; break;
            case 'f': out += '\f'; break;
            case 'n': out += '\n'; break;
            case 'r': out += '\r'; break;
            case 't': out += '\t'; break;
            case 'u': {
                unsigned cp;
                if (!hex4(cp)) return false;
                if (cp >= 0xD800 && cp < 0xDC00) {          // surrogate pair
                    unsigned lo;
                    if (!lit("\\u") || !hex4(lo) || lo < 0xDC00 || lo >= 0xE000)
                        return fail("unpaired surrogate");
                    cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                }
                put_utf8(out, cp);
                break;
            }
            default: return fail("unknown escape");
            }
        }
        if (p >= end) return fail("unterminated string");
        ++p;
        return true;
    }
    bool number(double& v) {
        const char* s = p;
        if (p < end && *p == '-') ++p;
        if (p >= end || !std::isdigit((unsigned char)*p)) return fail("bad number");
        if (*p == '0') ++p;
        else while (p < end && std::isdigit((unsigned char)*p)) ++p;
        if (p < end && *p == '.') {
            ++p;
            if (p >= end || !std::isdigit((unsigned char)*p)) return fail("bad fraction");
            while (p < end && std::isdigit((unsigned char)*p)) ++p;
        }
        if (p < end && (*p == 'e' || *p == 'E')) {
            ++p;
            if (p < end && (*p == '+' || *p == '-')) ++p;
            if (p >= end || !std::isdigit((unsigned char)*p)) return fail("bad exponent");
            while (p < end && std::isdigit((unsigned char)*p)) ++p;
        }
        v = std::strtod(std::string(s, p).c_str(), nullptr);
        return true;
    }
    bool value(jvalue& v) {
        if (++depth > 64) return fail("nesting too deep");
        ws();
        if (p >= end) return fail("unexpected end of input");
        bool ok;
        if (*p == '{') {
            ++p;
            v.kind = jvalue::OBJ;
            ws();
            ok = true;
            if (p < end && *p == '}') ++p;
            else {
                for (;;) {
                    ws();
                    std::string key;
                    if (!string(key)) { ok = false; break; }
                    ws();
                    if (p >= end || *p != ':') { ok = fail("expected ':'"); break; }
                    ++p;
                    jvalue child;
                    if (!value(child)) { ok = false; break; }
                    v.obj[key] = std::move(child);
                    ws();
                    if (p < end && *p == ',') { ++p; continue; }
                    if (p < end && *p == '}') { ++p; break; }
                    ok = fail("expected ',' or '}'");
                    break;
                }
            }
        } else if (*p == '[') {
            ++p;
            v.kind = jvalue::ARR;
            ws();
            ok = true;
            if (p < end && *p == ']') ++p;
            else {
                for (;;) {
                    jvalue child;
                    if (!value(child)) { ok = false; break; }
                    v.arr.push_back(std::move(child));
                    ws();
                    if (p < end && *p == ',') { ++p; continue; }
                    if (p < end && *p == ']') { ++p; break; }
                    ok = fail("expected ',' or ']'");
                    break;
                }
            }
        } else if (*p == '"') {
            v.kind = jvalue::STR;
            ok = string(v.str);
        } else if (lit("true")) {
            v.kind = jvalue::BOOL; v.b = true; ok = true;
        } else if (lit("false")) {
            v.kind = jvalue::BOOL; v.b = false; ok = true;
        } else if (lit("null")) {
            v.kind = jvalue::NUL; ok = true;
        } else {
            v.kind = jvalue::NUM;
            ok = number(v.num);
        }
        --depth;
        return ok;
    }
};

// a member under any of the given spellings
const jvalue* member(const jvalue& o, std::initializer_list<const char*> names) {
    if (o.kind != jvalue::OBJ) return nullptr;
    for (const char* nm : names) {
        auto it = o.obj.find(nm);
        if (it != o.obj.end()) return &it->second;
    }
    return nullptr;
}

// a numeric setting written as a number or as a string holding one
int as_number(const jvalue* v, const char* name, double& out) {
    if (!v) return 1;                               // absent: keep the default
    if (v->kind == jvalue::NUM) { out = v->num; return 0; }
    if (v->kind == jvalue::BOOL) { out = v->b ? 1.0 : 0.0; return 0; }
    if (v->kind == jvalue::STR) {
        const char* s = v->str.c_str();
        char* e = nullptr;
        errno = 0;
        const double d = std::strtod(s, &e);
        while (e && *e && std::isspace((unsigned char)*e)) ++e;
        if (e != s && e && *e == '\0' && errno == 0 && std::isfinite(d)) { out = d; return 0; }
        if (v->str == "true") { out = 1.0; return 0; }
        if (v->str == "false") { out = 0.0; return 0; }
    }
    mpc_set_error("settings: %s is not a number", name);
    return MPC_ERR_ARG;
}

std::string lower(std::string s) {
    for (char& c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

}  // namespace

extern "C" void mpc_settings_default(mpc_settings* s) {
    if (!s) return;
    std::memset(s, 0, sizeof *s);
    s->n = 8192;
    s->condition_number = 100.0;
    s->dist_type = MPC_DIST_ARITHMETIC;
    s->block_size = 256;
    s->eps_prime = 0.25;
}

extern "C" int mpc_settings_parse(const char* text, size_t len, mpc_settings* out) {
    MPC_CHECK(text && out, "settings: NULL argument");
    jparser ps{text, text + len, {}};
    jvalue root;
    if (!ps.value(root)) {
        mpc_set_error("settings: JSON parse error at byte %lld: %s", (long long)(ps.p - text),
                      ps.err.c_str());
        return MPC_ERR_ARG;
    }
    ps.ws();
    MPC_CHECK(ps.p == ps.end, "settings: trailing characters after the JSON value");
    MPC_CHECK(root.kind == jvalue::OBJ, "settings: the top-level value is not an object");
    mpc_settings s;
    mpc_settings_default(&s);
    const jvalue* ms = member(root, {"matrix_settings", "matrix settings"});
    const jvalue* fs = member(root, {"factorization_settings", "factorization settings"});
    int rc;
    double v;
    if (ms) {
        if ((rc = as_number(member(*ms, {"n"}), "n", v)) < 0) return rc;
        if (rc == 0) {
            MPC_CHECK(v >= 1 && v <= 4.0e9 && v == std::floor(v), "settings: n = %g", v);
            s.n = (int64_t)v;
        }
        if ((rc = as_number(member(*ms, {"condition_number", "condition number"}),
                            "condition_number", v)) < 0)
            return rc;
        if (rc == 0) {
            MPC_CHECK(v >= 1.0, "settings: condition_number = %g (must be >= 1)", v);
            s.condition_number = v;
        }
        const jvalue* dt = member(*ms, {"dist_type", "dist type"});
        if (dt) {
            // DistType names (kernels_macros.cuh:24-29), or their ordinals
            int d = -1;
            if (dt->kind == jvalue::STR) {
                const std::string t = lower(dt->str);
                if (t == "randomloguniform" || t == "random_log_uniform" || t == "logrand" ||
                    t == "loguniform" || t == "log_uniform")
                    d = MPC_DIST_RANDOM_LOG_UNIFORM;
                else if (t == "clustered") d = MPC_DIST_CLUSTERED;
                else if (t == "arithmetic") d = MPC_DIST_ARITHMETIC;
                else if (t == "geometric") d = MPC_DIST_GEOMETRIC;
                else if (t == "logspaced" || t == "log_spaced") d = MPC_DIST_LOGSPACED;
                else if (as_number(dt, "dist_type", v) == 0 && v == std::floor(v)) d = (int)v;
            } else if (as_number(dt, "dist_type", v) == 0 && v == std::floor(v)) {
                d = (int)v;
            }
            MPC_CHECK(d >= 0 && d <= MPC_DIST_LOGSPACED, "settings: unknown dist_type");
            s.dist_type = d;
        }
    }
    if (fs) {
        if ((rc = as_number(member(*fs, {"block_size", "block size"}), "block_size", v)) < 0)
            return rc;
        if (rc == 0) {
            MPC_CHECK(v >= 1 && v <= 65536 && v == std::floor(v), "settings: block_size = %g", v);
            s.block_size = (int)v;
        }
        if ((rc = as_number(member(*fs, {"eps_prime", "eps prime"}), "eps_prime", v)) < 0) return rc;
        if (rc == 0) {
            MPC_CHECK(v > 0, "settings: eps_prime = %g (must be > 0)", v);
            s.eps_prime = v;
        }
        if ((rc = as_number(member(*fs, {"floor"}), "floor", v)) < 0) return rc;
        if (rc == 0) s.floor = v;
        if ((rc = as_number(member(*fs, {"diag_pert", "diag pert"}), "diag_pert", v)) < 0) return rc;
        if (rc == 0) s.diag_pert = v != 0.0;
        if ((rc = as_number(member(*fs, {"left"}), "left", v)) < 0) return rc;
        if (rc == 0) s.left = v != 0.0;
    }
    *out = s;
    return 0;
}

extern "C" int mpc_settings_read(const char* path, mpc_settings* out) {
    MPC_CHECK(path && out, "settings: NULL argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) { mpc_set_error("settings: cannot open %s", path); return MPC_ERR_ARG; }
    std::string text;
    char buf[4096];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) {
        text.append(buf, got);
        if (text.size() > (1u << 24)) {
            std::fclose(f);
            mpc_set_error("settings: %s is larger than 16 MiB", path);
            return MPC_ERR_ARG;
        }
    }
    std::fclose(f);
    return mpc_settings_parse(text.data(), text.size(), out);
}

// The reference's block size is the width of its trailing-update panels (the
// K of the update GEMM).  Here a block is 64, 128 or 256 columns and wider
// updates come from aggregating panels: block_size = a * 256 (a <= 4) maps to
// nb = 256 with agg_panels = a; 64, 128, 256 map to themselves.
extern "C" int mpc_settings_apply(const mpc_settings* s, mpc_matgen_opts* mo, mpc_factor_opts* fo) {
    MPC_CHECK(s, "settings: NULL argument");
    if (mo) {
        mo->dist = s->dist_type;
        mo->cond = s->condition_number;
    }
    if (fo) {
        const int b = s->block_size;
        if (b == 64 || b == 128 || b == 256) {
            fo->nb = b;
            fo->agg_panels = 1;
        } else if (b % 256 == 0 && b / 256 >= 2 && b / 256 <= MPC_AGG_MAX) {
            fo->nb = 256;
            fo->agg_panels = b / 256;
        } else {
            mpc_set_error("settings: block_size %d is not 64, 128 or a multiple of 256 up to %d", b,
                          256 * MPC_AGG_MAX);
            return MPC_ERR_UNSUPPORTED;
        }
        fo->eps_prime = s->eps_prime;
    }
    return 0;
}

// ---------------------------------------------------------------- idx
namespace {
uint32_t be32(const unsigned char* b) {
    return ((uint32_t)b[0] << 24) | ((uint32_t)b[1] << 16) | ((uint32_t)b[2] << 8) | b[3];
}
}  // namespace

extern "C" int mpc_idx_header(const char* path, int* type_code, int* ndim, int64_t dims[4],
                              int64_t* data_offset) {
    MPC_CHECK(path && type_code && ndim && dims, "idx: NULL argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) { mpc_set_error("idx: cannot open %s", path); return MPC_ERR_ARG; }
    unsigned char hdr[4 + 16];
    const size_t got = std::fread(hdr, 1, 4, f);
    if (got != 4 || hdr[0] != 0 || hdr[1] != 0) {
        std::fclose(f);
        mpc_set_error("idx: %s does not start with an idx magic (00 00 type ndim)", path);
        return MPC_ERR_ARG;
    }
    const int nd = hdr[3];
    if (nd < 1 || nd > 4) {
        std::fclose(f);
        mpc_set_error("idx: %s has %d dimensions (1..4 supported)", path, nd);
        return MPC_ERR_ARG;
    }
    if (std::fread(hdr + 4, 1, 4 * (size_t)nd, f) != 4 * (size_t)nd) {
        std::fclose(f);
        mpc_set_error("idx: %s: truncated header", path);
        return MPC_ERR_ARG;
    }
    std::fseek(f, 0, SEEK_END);
    const long fsize = std::ftell(f);
    std::fclose(f);
    int esz;
    switch (hdr[2]) {
    case 0x08: case 0x09: esz = 1; break;   // unsigned / signed byte
    case 0x0B: esz = 2; break;              // short
    case 0x0C: case 0x0D: esz = 4; break;   // int / float
    case 0x0E: esz = 8; break;              // double
    default:
        mpc_set_error("idx: %s: unknown element type 0x%02x", path, hdr[2]);
        return MPC_ERR_ARG;
    }
    int64_t count = 1;
    for (int i = 0; i < nd; ++i) {
        dims[i] = be32(hdr + 4 + 4 * i);
        count *= dims[i];
    }
    for (int i = nd; i < 4; ++i) dims[i] = 1;
    const int64_t off = 4 + 4 * (int64_t)nd;
    if (fsize < 0 || off + count * esz != (int64_t)fsize) {
        mpc_set_error("idx: %s: %lld bytes of data expected after the header, file has %ld bytes",
                      path, (long long)(count * esz), fsize);
        return MPC_ERR_ARG;
    }
    *type_code = hdr[2];
    *ndim = nd;
    if (data_offset) *data_offset = off;
    return 0;
}

// Unsigned-byte idx data (labels or images), the first `count` items (count
// <= dims[0]; all if count < 0) into out (row-major, the item's remaining
// dimensions contiguous).  Returns the number of items read in *items.
extern "C" int mpc_idx_read_u8(const char* path, int64_t count, uint8_t* out, int64_t out_bytes,
                               int64_t* items) {
    int type, nd;
    int64_t dims[4], off;
    int rc = mpc_idx_header(path, &type, &nd, dims, &off);
    if (rc) return rc;
    MPC_CHECK(type == 0x08, "idx: %s is not unsigned-byte data (type 0x%02x)", path, type);
    int64_t per = 1;
    for (int i = 1; i < nd; ++i) per *= dims[i];
    const int64_t nitem = count < 0 ? dims[0] : std::min<int64_t>(count, dims[0]);
    MPC_CHECK(out && out_bytes >= nitem * per, "idx: output buffer of %lld bytes, %lld needed",
              (long long)out_bytes, (long long)(nitem * per));
    FILE* f = std::fopen(path, "rb");
    if (!f) { mpc_set_error("idx: cannot open %s", path); return MPC_ERR_ARG; }
    std::fseek(f, (long)off, SEEK_SET);
    const size_t want = (size_t)(nitem * per);
    const size_t got = std::fread(out, 1, want, f);
    std::fclose(f);
    MPC_CHECK(got == want, "idx: %s: short read", path);
    if (items) *items = nitem;
    return 0;
}

// Images of an idx3 file as RBF points: n x (rows*cols) FP32, pixel / 255,
// written row-major (ld = rows*cols) to the device array dX.  This is the
// real-data form of the image-shaped RBF workload.
extern "C" int mpc_idx_points(mpc_handle* h, const char* path, int64_t n, float* dX, int64_t ldx,
                              int64_t* n_out, int* d_out) {
    MPC_CHECK(h && path && dX, "idx_points: NULL argument");
    int type, nd;
    int64_t dims[4], off;
    int rc = mpc_idx_header(path, &type, &nd, dims, &off);
    if (rc) return rc;
    MPC_CHECK(type == 0x08 && nd == 3, "idx_points: %s is not an unsigned-byte image file", path);
    const int64_t d = dims[1] * dims[2];
    const int64_t m = n < 0 ? dims[0] : std::min<int64_t>(n, dims[0]);
    MPC_CHECK(ldx >= d, "idx_points: ldx %lld < %lld pixels", (long long)ldx, (long long)d);
    std::vector<uint8_t> raw((size_t)(m * d));
    int64_t got = 0;
    if ((rc = mpc_idx_read_u8(path, m, raw.data(), (int64_t)raw.size(), &got))) return rc;
    std::vector<float> X((size_t)(m * ldx), 0.f);
    for (int64_t i = 0; i < m; ++i)
        for (int64_t c = 0; c < d; ++c) X[(size_t)(i * ldx + c)] = (float)raw[(size_t)(i * d + c)] / 255.f;
    MPC_HIP(hipSetDevice(h->device));
    MPC_HIP(hipMemcpy(dX, X.data(), sizeof(float) * X.size(), hipMemcpyHostToDevice));
    if (n_out) *n_out = m;
    if (d_out) *d_out = (int)d;
    return 0;
}
