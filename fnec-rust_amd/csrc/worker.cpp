// fnec_hip_worker — the nec_worker stdio protocol served from the GPU.
//
// Reads newline-delimited JSON task messages on stdin and writes one JSON
// result line per task on stdout, as run_worker_stdio does
// (crates/nec_worker/src/worker.rs:12-44): blank lines are skipped, and
// {"cmd":"shutdown"} or end of input ends the loop. Each task goes through
// the same steps and error mapping as process_task (worker.rs:46-128) and
// solve_deck_at_frequency_with_exec (crates/nec_worker/src/solve.rs:150-289):
// JSON -> base64 -> UTF-8 -> basis check -> parse -> geometry -> Hallén inputs
// with worker semantics (raw wire ranges, KCL junction rows, first type-0 EX)
// -> solve -> Z_in = V / I at the feed, VSWR against 50 ohm.
//
// Serving mode (the default, --serve): tasks are solved by a continuous-
// batching session (fnec_hip_session_*), so several are on the device at once
// and their replies are written as they finish, out of order, each carrying
// its task_id (TaskResult, crates/nec_worker/src/protocol.rs:178-229). A task
// the session cannot admit (its footprint exceeds the device budget, or the
// queue holds --max-queue tasks) is answered at once with error_code
// "resource_exhausted". On shutdown or end of input the worker stops reading,
// answers every task in flight, then exits. --serial answers each task before
// reading the next line, as the reference does.
//
// Protocol extensions (optional task fields, ignored by the reference):
//   "return_currents": true  -> the reply carries "currents": [[re, im], ...]
//                               (segment currents, then one constant per wire)
//   "rp": {"theta_deg": [...], "phi_deg": [...]}
//                            -> the reply carries "rp": {"total_radiated_w":
//                               P, "rows": [[gain_total_dbi, gain_theta_dbi,
//                               gain_phi_dbi, axial_ratio], ...]}
//   solver_config.precision  "fp64" (the reference's algorithm and rounding;
//                               the default when exec is "cpu"), "mixed" (the
//                               default when exec is "gpu") or "fp64_mfma".
//
// Result cache (--cache N, off by default): results of solved tasks are kept,
// up to N, and evicted oldest first (ResultCache, crates/nec_worker/src/
// result_cache.rs:27-108). The key is the reference's cache_key
// (result_cache.rs:11-22: SHA-256 of the deck text, basis, ground model and
// the frequency's bits) extended by what changes this worker's reply: the
// precision, return_currents and the rp angles. A task whose key is cached is
// answered at once with the stored reply under its own task_id, marked
// "cached": true. Only "ok" replies are stored.
// exec_used is always "gpu": the worker has no CPU solver. LD and TL cards
// are stamped into Z on the device (NT cards are ignored, as the reference's
// worker ignores them); decks the device path does not cover (plane-wave or
// current-source EX) are answered with error_code "unsupported_config".
#include <algorithm>
#include <atomic>
#include <charconv>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <iostream>
#include <map>
#include <unordered_map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fnec_host.h"

namespace {

constexpr double kPi = 3.14159265358979323846; // f64::consts::PI; to_degrees is x * (180 / PI)

// ---- JSON (RFC 8259 subset that serde_json accepts for these messages) ----

struct JVal {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;
    const JVal *get(const char *key) const
    {
        if (kind != Obj)
            return nullptr;
        for (auto &kv : obj)
            if (kv.first == key)
                return &kv.second;
        return nullptr;
    }
};

struct JParser {
    const std::string &s;
    size_t i = 0;
    std::string err;
    explicit JParser(const std::string &text) : s(text) {}

    bool fail(const char *what)
    {
        if (err.empty())
            err = std::string(what) + " at column " + std::to_string(i + 1);
        return false;
    }
    void ws()
    {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r'))
            ++i;
    }
    bool lit(const char *w)
    {
        size_t n = std::strlen(w);
        if (s.compare(i, n, w) != 0)
            return fail("expected value");
        i += n;
        return true;
    }
    static void put_utf8(std::string &o, uint32_t c)
    {
        if (c < 0x80) {
            o += (char)c;
        } else if (c < 0x800) {
            o += (char)(0xC0 | (c >> 6));
            o += (char)(0x80 | (c & 0x3F));
        } else if (c < 0x10000) {
            o += (char)(0xE0 | (c >> 12));
            o += (char)(0x80 | ((c >> 6) & 0x3F));
            o += (char)(0x80 | (c & 0x3F));
        } else {
            o += (char)(0xF0 | (c >> 18));
            o += (char)(0x80 | ((c >> 12) & 0x3F));
            o += (char)(0x80 | ((c >> 6) & 0x3F));
            o += (char)(0x80 | (c & 0x3F));
        }
    }
    bool hex4(uint32_t &v)
    {
        if (i + 4 > s.size())
            return fail("EOF while parsing a string");
        v = 0;
        for (int k = 0; k < 4; ++k) {
            char c = s[i++];
            v <<= 4;
            if (c >= '0' && c <= '9')
                v |= (uint32_t)(c - '0');
            else if (c >= 'a' && c <= 'f')
                v |= (uint32_t)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F')
                v |= (uint32_t)(c - 'A' + 10);
            else
                return fail("invalid escape");
        }
        return true;
    }
    bool string(std::string &o)
    {
        ++i; // opening quote
        while (true) {
            if (i >= s.size())
                return fail("EOF while parsing a string");
            unsigned char c = (unsigned char)s[i++];
            if (c == '"')
                return true;
            if (c < 0x20)
                return fail("control character (\\u0000-\\u001F) found while parsing a string");
            if (c != '\\') {
                o += (char)c;
                continue;
            }
            if (i >= s.size())
                return fail("EOF while parsing a string");
            char e = s[i++];
            switch (e) {
            case '"': o += '"'; break;
            case '\\': o += '\\'; break;
            case '/': o += '/'; break;
            case 'b': o += '\b'; break;
            case 'f': o += '\f'; break;
            case 'n': o += '\n'; break;
            case 'r': o += '\r'; break;
            case 't': o += '\t'; break;
            case 'u': {
                uint32_t v = 0;
                if (!hex4(v))
                    return false;
                if (v >= 0xD800 && v < 0xDC00) {
                    uint32_t lo = 0;
                    if (i + 2 > s.size() || s[i] != '\\' || s[i + 1] != 'u')
                        return fail("lone leading surrogate in hex escape");
                    i += 2;
                    if (!hex4(lo))
                        return false;
                    if (lo < 0xDC00 || lo >= 0xE000)
                        return fail("lone leading surrogate in hex escape");
                    v = 0x10000 + ((v - 0xD800) << 10) + (lo - 0xDC00);
                } else if (v >= 0xDC00 && v < 0xE000) {
                    return fail("lone trailing surrogate in hex escape");
                }
                put_utf8(o, v);
                break;
            }
            default:
                return fail("invalid escape");
            }
        }
    }
    bool number(double &v)
    {
        size_t b = i;
        if (s[i] == '-')
            ++i;
        if (i >= s.size())
            return fail("invalid number");
        if (s[i] == '0') {
            ++i;
        } else if (s[i] >= '1' && s[i] <= '9') {
            while (i < s.size() && std::isdigit((unsigned char)s[i]))
                ++i;
        } else {
            return fail("invalid number");
        }
        if (i < s.size() && s[i] == '.') {
            ++i;
            if (i >= s.size() || !std::isdigit((unsigned char)s[i]))
                return fail("invalid number");
            while (i < s.size() && std::isdigit((unsigned char)s[i]))
                ++i;
        }
        if (i < s.size() && (s[i] == 'e' || s[i] == 'E')) {
            ++i;
            if (i < s.size() && (s[i] == '+' || s[i] == '-'))
                ++i;
            if (i >= s.size() || !std::isdigit((unsigned char)s[i]))
                return fail("invalid number");
            while (i < s.size() && std::isdigit((unsigned char)s[i]))
                ++i;
        }
        const std::string t = s.substr(b, i - b);
        char *end = nullptr;
        v = std::strtod(t.c_str(), &end);
        if (!std::isfinite(v))
            return fail("number out of range");
        return true;
    }
    bool value(JVal &v, int depth)
    {
        if (depth > 128)
            return fail("recursion limit exceeded");
        ws();
        if (i >= s.size())
            return fail("EOF while parsing a value");
        char c = s[i];
        if (c == '{') {
            v.kind = JVal::Obj;
            ++i;
            ws();
            if (i < s.size() && s[i] == '}') {
                ++i;
                return true;
            }
            while (true) {
                ws();
                if (i >= s.size() || s[i] != '"')
                    return fail("key must be a string");
                std::string k;
                if (!string(k))
                    return false;
                ws();
                if (i >= s.size() || s[i] != ':')
                    return fail("expected `:`");
                ++i;
                JVal x;
                if (!value(x, depth + 1))
                    return false;
                v.obj.emplace_back(std::move(k), std::move(x));
                ws();
                if (i < s.size() && s[i] == ',') {
                    ++i;
                    continue;
                }
                if (i < s.size() && s[i] == '}') {
                    ++i;
                    return true;
                }
                return fail("expected `,` or `}`");
            }
        }
        if (c == '[') {
            v.kind = JVal::Arr;
            ++i;
            ws();
            if (i < s.size() && s[i] == ']') {
                ++i;
                return true;
            }
            while (true) {
                JVal x;
                if (!value(x, depth + 1))
                    return false;
                v.arr.push_back(std::move(x));
                ws();
                if (i < s.size() && s[i] == ',') {
                    ++i;
                    continue;
                }
                if (i < s.size() && s[i] == ']') {
                    ++i;
                    return true;
                }
                return fail("expected `,` or `]`");
            }
        }
        if (c == '"') {
            v.kind = JVal::Str;
            return string(v.str);
        }
        if (c == 't') {
            v.kind = JVal::Bool;
            v.b = true;
            return lit("true");
        }
        if (c == 'f') {
            v.kind = JVal::Bool;
            return lit("false");
        }
        if (c == 'n') {
            v.kind = JVal::Null;
            return lit("null");
        }
        if (c == '-' || (c >= '0' && c <= '9')) {
            v.kind = JVal::Num;
            return number(v.num);
        }
        return fail("expected value");
    }
    bool document(JVal &v)
    {
        if (!value(v, 0))
            return false;
        ws();
        if (i != s.size())
            return fail("trailing characters");
        return true;
    }
};

// shortest decimal that reads back as the same double; non-finite values are
// written as null, as serde_json writes them
std::string jnum(double x)
{
    if (!std::isfinite(x))
        return "null";
    char b[64];
    const double a = std::fabs(x);
    std::to_chars_result r;
    if (a == 0.0 || (a >= 1e-5 && a < 1e16))
        r = std::to_chars(b, b + sizeof b, x, std::chars_format::fixed);
    else
        r = std::to_chars(b, b + sizeof b, x, std::chars_format::scientific);
    std::string o(b, r.ptr);
    if (o.find_first_of(".e") == std::string::npos)
        o += ".0";
    return o;
}

std::string jstr(const std::string &s)
{
    std::string o = "\"";
    for (unsigned char c : s) {
        switch (c) {
        case '"': o += "\\\""; break;
        case '\\': o += "\\\\"; break;
        case '\n': o += "\\n"; break;
        case '\r': o += "\\r"; break;
        case '\t': o += "\\t"; break;
        case '\b': o += "\\b"; break;
        case '\f': o += "\\f"; break;
        default:
            if (c < 0x20) {
                char u[8];
                std::snprintf(u, sizeof u, "\\u%04x", c);
                o += u;
            } else {
                o += (char)c;
            }
        }
    }
    return o + "\"";
}

// ---- base64 (RFC 4648 standard alphabet, canonical padding, as the base64
// crate's STANDARD engine decodes) ----
bool b64_decode(const std::string &in, std::string &out, std::string &err)
{
    auto val = [](char c) -> int {
        if (c >= 'A' && c <= 'Z')
            return c - 'A';
        if (c >= 'a' && c <= 'z')
            return c - 'a' + 26;
        if (c >= '0' && c <= '9')
            return c - '0' + 52;
        if (c == '+')
            return 62;
        if (c == '/')
            return 63;
        return -1;
    };
    out.clear();
    const size_t n = in.size();
    for (size_t i = 0; i < n; ++i)
        if (val(in[i]) < 0 && in[i] != '=') {
            err = "Invalid symbol " + std::to_string((unsigned char)in[i]) + ", offset " + std::to_string(i) + ".";
            return false;
        }
    if (n % 4 != 0) {
        err = "Invalid padding";
        return false;
    }
    for (size_t q = 0; q < n; q += 4) {
        int v[4], pad = 0;
        for (int k = 0; k < 4; ++k) {
            const char c = in[q + k];
            if (c == '=') {
                if (q + 4 != n || k < 2) {
                    err = "Invalid symbol 61, offset " + std::to_string(q + k) + ".";
                    return false;
                }
                ++pad;
                v[k] = 0;
            } else {
                if (pad) {
                    err = "Invalid padding";
                    return false;
                }
                v[k] = val(c);
            }
        }
        const uint32_t w = (uint32_t)v[0] << 18 | (uint32_t)v[1] << 12 | (uint32_t)v[2] << 6 | (uint32_t)v[3];
        out += (char)(w >> 16);
        if (pad < 2)
            out += (char)((w >> 8) & 0xFF);
        if (pad < 1)
            out += (char)(w & 0xFF);
        // the unused bits of the last symbol must be zero
        if ((pad == 2 && (w & 0xFFFF)) || (pad == 1 && (w & 0xFF))) {
            err = "Invalid last symbol " + std::to_string((unsigned char)in[q + 3 - pad]) + ", offset " +
                  std::to_string(q + 3 - pad) + ".";
            return false;
        }
    }
    return true;
}

// UTF-8 validity (RFC 3629: no overlong forms, no surrogates, <= U+10FFFF)
bool utf8_valid(const std::string &s, size_t &bad)
{
    const unsigned char *p = (const unsigned char *)s.data();
    const size_t n = s.size();
    for (size_t i = 0; i < n;) {
        const unsigned char c = p[i];
        size_t len;
        uint32_t cp;
        if (c < 0x80) {
            ++i;
            continue;
        } else if ((c & 0xE0) == 0xC0) {
            len = 2, cp = c & 0x1F;
        } else if ((c & 0xF0) == 0xE0) {
            len = 3, cp = c & 0x0F;
        } else if ((c & 0xF8) == 0xF0) {
            len = 4, cp = c & 0x07;
        } else {
            bad = i;
            return false;
        }
        if (i + len > n) {
            bad = i;
            return false;
        }
        for (size_t k = 1; k < len; ++k) {
            if ((p[i + k] & 0xC0) != 0x80) {
                bad = i;
                return false;
            }
            cp = cp << 6 | (p[i + k] & 0x3F);
        }
        if ((len == 2 && cp < 0x80) || (len == 3 && cp < 0x800) || (len == 4 && cp < 0x10000) || cp > 0x10FFFF ||
            (cp >= 0xD800 && cp < 0xE000)) {
            bad = i;
            return false;
        }
        i += len;
    }
    return true;
}

// VSWR of z against z0, as worker.rs:131-143: infinite for a vanishing
// denominator or |Gamma| >= 1
double vswr(double re, double im, double z0)
{
    const double num_sq = (re - z0) * (re - z0) + im * im;
    const double den_sq = (re + z0) * (re + z0) + im * im;
    if (den_sq < 1e-100)
        return INFINITY;
    const double gamma = std::sqrt(num_sq) / std::sqrt(den_sq);
    if (gamma >= 1.0)
        return INFINITY;
    return (1.0 + gamma) / (1.0 - gamma);
}

// ---- the worker ----

// ---- SHA-256 (FIPS 180-4) and the result-cache key ----
std::string sha256_hex(const std::string &msg)
{
    static const uint32_t K[64] = {
        0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
        0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
        0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
        0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
        0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
        0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
        0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
        0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};
    uint32_t h[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a, 0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
    std::string m = msg;
    const uint64_t bits = (uint64_t)msg.size() * 8;
    m.push_back((char)0x80);
    while (m.size() % 64 != 56)
        m.push_back('\0');
    for (int i = 7; i >= 0; --i)
        m.push_back((char)((bits >> (8 * i)) & 0xff));
    auto rotr = [](uint32_t x, int n) { return (x >> n) | (x << (32 - n)); };
    for (size_t blk = 0; blk < m.size(); blk += 64) {
        uint32_t w[64];
        for (int t = 0; t < 16; ++t)
            w[t] = (uint32_t)(unsigned char)m[blk + 4 * t] << 24 | (uint32_t)(unsigned char)m[blk + 4 * t + 1] << 16 |
                   (uint32_t)(unsigned char)m[blk + 4 * t + 2] << 8 | (uint32_t)(unsigned char)m[blk + 4 * t + 3];
        for (int t = 16; t < 64; ++t) {
            const uint32_t s0 = rotr(w[t - 15], 7) ^ rotr(w[t - 15], 18) ^ (w[t - 15] >> 3);
            const uint32_t s1 = rotr(w[t - 2], 17) ^ rotr(w[t - 2], 19) ^ (w[t - 2] >> 10);
            w[t] = w[t - 16] + s0 + w[t - 7] + s1;
        }
        uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
        for (int t = 0; t < 64; ++t) {
            const uint32_t t1 = hh + (rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25)) + ((e & f) ^ (~e & g)) + K[t] + w[t];
            const uint32_t t2 = (rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22)) + ((a & b) ^ (a & c) ^ (b & c));
            hh = g;
            g = f;
            f = e;
            e = d + t1;
            d = c;
            c = b;
            b = a;
            a = t1 + t2;
        }
        h[0] += a, h[1] += b, h[2] += c, h[3] += d, h[4] += e, h[5] += f, h[6] += g, h[7] += hh;
    }
    char out[65];
    for (int i = 0; i < 8; ++i)
        std::snprintf(out + 8 * i, 9, "%08x", h[i]);
    return std::string(out, 64);
}

// cache_key (result_cache.rs:11-22): SHA-256 of deck \0 basis \0
// ground_model \0 the frequency's IEEE bits as 16 lower-case hex digits
std::string cache_key(const std::string &deck, const std::string &basis, const std::string &ground, double freq)
{
    uint64_t b;
    std::memcpy(&b, &freq, 8);
    char fb[17];
    std::snprintf(fb, sizeof fb, "%016llx", (unsigned long long)b);
    return sha256_hex(deck + std::string(1, '\0') + basis + std::string(1, '\0') + ground + std::string(1, '\0') +
                      fb);
}

std::mutex g_out_mu;
void emit(const std::string &line)
{
    std::lock_guard<std::mutex> lk(g_out_mu);
    std::fwrite(line.data(), 1, line.size(), stdout);
    std::fputc('\n', stdout);
    std::fflush(stdout);
}

std::string error_line(const std::string &task_id, double freq, const char *code, const std::string &msg)
{
    return "{\"status\":\"error\",\"task_id\":" + jstr(task_id) + ",\"frequency_hz\":" + jnum(freq) +
           ",\"error_code\":\"" + code + "\",\"error_message\":" + jstr(msg) + "}";
}

struct Task {
    std::string key; // result-cache key (empty: the cache is off)
    std::string task_id;
    double freq = 0.0;
    fnec_c64 v_feed{0.0, 0.0};
    int precision = FNEC_PREC_MIXED;
    bool want_currents = false;
    std::vector<fnec_c64> x;
    std::vector<double> theta, phi;
    std::vector<fnec_rp_row> rows;
};

const char *precision_name(int p)
{
    return p == FNEC_PREC_FP64 ? "fp64" : (p == FNEC_PREC_MIXED ? "mixed" : "fp64_mfma");
}

std::string ok_line(const Task &t, const fnec_result &r)
{
    // Z_in = V / I at the feed (solve.rs:243-268): V itself when |I| <= 1e-60
    const double ir = r.feed_current.re, ii = r.feed_current.im;
    const double mag = std::hypot(ir, ii);
    double zr = t.v_feed.re, zi = t.v_feed.im;
    if (mag > 1e-60) {
        zr = r.z_in.re;
        zi = r.z_in.im;
    }
    std::string o = "{\"status\":\"ok\",\"task_id\":" + jstr(t.task_id) + ",\"frequency_hz\":" + jnum(t.freq) +
                    ",\"impedance\":{\"re_ohm\":" + jnum(zr) + ",\"im_ohm\":" + jnum(zi) +
                    "},\"vswr_50\":" + jnum(vswr(zr, zi, 50.0)) + ",\"feedpoint_current_mag\":" + jnum(mag) +
                    ",\"feedpoint_current_phase_deg\":" + jnum(std::atan2(ii, ir) * (180.0 / kPi)) +
                    ",\"exec_used\":\"gpu\",\"precision\":\"" + precision_name(t.precision) + "\"";
    if (t.want_currents) {
        o += ",\"currents\":[";
        for (size_t k = 0; k < t.x.size(); ++k)
            o += (k ? ",[" : "[") + jnum(t.x[k].re) + "," + jnum(t.x[k].im) + "]";
        o += "]";
    }
    if (!t.rows.empty()) {
        o += ",\"rp\":{\"total_radiated_w\":" + jnum(r.rp_total) + ",\"rows\":[";
        for (size_t k = 0; k < t.rows.size(); ++k) {
            const fnec_rp_row &w = t.rows[k];
            o += (k ? ",[" : "[") + jnum(w.gain_total_dbi) + "," + jnum(w.gain_theta_dbi) + "," +
                 jnum(w.gain_phi_dbi) + "," + jnum(w.axial_ratio) + "]";
        }
        o += "]}";
    }
    return o + "}";
}

struct Options {
    bool serial = false;
    double budget_gib = 64.0;
    int max_batch = 256;
    int max_queue = 4096;
    size_t cache = 0; // result-cache capacity (0: off)
};

class Worker {
  public:
    explicit Worker(const Options &o) : opt_(o) {}
    ~Worker()
    {
        stop_poller();
        for (auto &s : sessions_)
            if (s)
                fnec_hip_session_destroy(s);
    }

    void handle_line(const std::string &line)
    {
        std::string msg;
        auto task = std::make_unique<Task>();
        fnec_deck *deck = nullptr;
        const char *code = prepare(line, *task, deck, msg);
        if (code) {
            emit(error_line(task->task_id.empty() ? "unknown" : task->task_id, task->freq, code, msg));
            if (deck)
                fnec_deck_free(deck);
            return;
        }
        if (!task->key.empty()) {
            std::string tail;
            {
                std::lock_guard<std::mutex> lk(cache_mu_);
                auto it = cache_.find(task->key);
                if (it != cache_.end())
                    tail = it->second;
            }
            if (!tail.empty()) {
                fnec_deck_free(deck);
                emit("{\"status\":\"ok\",\"task_id\":" + jstr(task->task_id) + tail.substr(0, tail.size() - 1) +
                     ",\"cached\":true}");
                return;
            }
        }
        submit(std::move(task), deck);
        if (opt_.serial)
            drain();
    }

    // answer every task in flight
    void drain()
    {
        std::unique_lock<std::mutex> lk(mu_);
        cv_done_.wait(lk, [&] { return tasks_.empty(); });
    }

  private:
    // validation and Hallén inputs; returns an error code name or nullptr
    const char *prepare(const std::string &line, Task &t, fnec_deck *&deck, std::string &msg)
    {
        JVal v;
        JParser p(line);
        if (!p.document(v)) {
            msg = "failed to deserialize task: " + p.err;
            return "parse_error";
        }
        // TaskMessage (protocol.rs:144-156): every field required
        const char *need[] = {"task_id", "deck_hash", "deck_b64", "solver_config", "frequency_hz"};
        if (v.kind != JVal::Obj) {
            msg = "failed to deserialize task: invalid type, expected struct TaskMessage";
            return "parse_error";
        }
        for (const char *k : need)
            if (!v.get(k)) {
                msg = std::string("failed to deserialize task: missing field `") + k + "`";
                return "parse_error";
            }
        const JVal *id = v.get("task_id"), *hash = v.get("deck_hash"), *b64 = v.get("deck_b64"),
                   *cfg = v.get("solver_config"), *fr = v.get("frequency_hz");
        if (id->kind != JVal::Str || hash->kind != JVal::Str || b64->kind != JVal::Str || cfg->kind != JVal::Obj ||
            fr->kind != JVal::Num) {
            msg = "failed to deserialize task: invalid type";
            return "parse_error";
        }
        std::string basis = "hallen", exec = "cpu", precision, ground = "none";
        for (auto &kv : cfg->obj) {
            if (kv.first == "basis" || kv.first == "ground_model" || kv.first == "exec" || kv.first == "precision") {
                if (kv.second.kind != JVal::Str) {
                    msg = "failed to deserialize task: invalid type for `" + kv.first + "`";
                    return "parse_error";
                }
                if (kv.first == "basis")
                    basis = kv.second.str;
                else if (kv.first == "ground_model")
                    ground = kv.second.str;
                else if (kv.first == "exec")
                    exec = kv.second.str;
                else if (kv.first == "precision")
                    precision = kv.second.str;
            }
        }
        t.task_id = id->str;
        t.freq = fr->num;
        std::string bytes, e;
        if (!b64_decode(b64->str, bytes, e)) {
            msg = "base64 decode failed: " + e;
            return "parse_error";
        }
        size_t bad = 0;
        if (!utf8_valid(bytes, bad)) {
            msg = "deck is not valid UTF-8: invalid utf-8 sequence at index " + std::to_string(bad);
            return "parse_error";
        }
        // solve.rs:156-160
        if (basis != "hallen") {
            msg = "basis '" + basis + "' not supported in worker; only 'hallen' is implemented";
            return "unsupported_config";
        }
        if (precision.empty())
            t.precision = exec == "gpu" ? FNEC_PREC_MIXED : FNEC_PREC_FP64;
        else if (precision == "fp64")
            t.precision = FNEC_PREC_FP64;
        else if (precision == "mixed")
            t.precision = FNEC_PREC_MIXED;
        else if (precision == "fp64_mfma")
            t.precision = FNEC_PREC_FP64_MFMA;
        else {
            msg = "precision '" + precision + "' not supported; use fp64, mixed or fp64_mfma";
            return "unsupported_config";
        }
        // extensions
        if (const JVal *rc = v.get("return_currents"))
            t.want_currents = rc->kind == JVal::Bool && rc->b;
        if (const JVal *rp = v.get("rp")) {
            const JVal *th = rp->get("theta_deg"), *ph = rp->get("phi_deg");
            if (!th || !ph || th->kind != JVal::Arr || ph->kind != JVal::Arr || th->arr.size() != ph->arr.size()) {
                msg = "rp: theta_deg and phi_deg must be arrays of equal length";
                return "parse_error";
            }
            for (size_t k = 0; k < th->arr.size(); ++k) {
                if (th->arr[k].kind != JVal::Num || ph->arr[k].kind != JVal::Num) {
                    msg = "rp: angles must be numbers";
                    return "parse_error";
                }
                t.theta.push_back(th->arr[k].num);
                t.phi.push_back(ph->arr[k].num);
            }
        }
        if (opt_.cache) {
            // the reference's key, then this worker's reply options
            std::string ext = cache_key(bytes, basis, ground, t.freq) + "|" + precision_name(t.precision) + "|" +
                              (t.want_currents ? "1" : "0");
            for (size_t k = 0; k < t.theta.size(); ++k) {
                char b[64];
                uint64_t u, w;
                std::memcpy(&u, &t.theta[k], 8);
                std::memcpy(&w, &t.phi[k], 8);
                std::snprintf(b, sizeof b, "|%016llx%016llx", (unsigned long long)u, (unsigned long long)w);
                ext += b;
            }
            t.key = sha256_hex(ext);
        }
        // parse, geometry, Hallén inputs (solve.rs:162-197)
        std::string text(bytes);
        char perr[512] = {0};
        deck = fnec_deck_parse(text.c_str(), perr, sizeof perr);
        if (!deck) {
            msg = std::string("parse error: ") + perr;
            return "parse_error";
        }
        if (fnec_deck_segment_count(deck) == 0) {
            msg = std::string("geometry error: ") + fnec_hip_last_error();
            return "parse_error";
        }
        fnec_deck_problem pr;
        const int st = fnec_deck_build(deck, FNEC_SEM_WORKER, t.freq, &pr);
        if (st == FNEC_UNAVAILABLE) {
            msg = std::string("unsupported config: ") + fnec_hip_last_error();
            return "unsupported_config";
        }
        if (st != FNEC_OK) {
            msg = std::string("parse error: ") + fnec_hip_last_error();
            return "parse_error";
        }
        if (pr.n_feeds == 0) {
            msg = "no EX type-0 card found in deck";
            return "parse_error";
        }
        t.v_feed = pr.req.v_feed;
        req_ = pr.req;
        return nullptr;
    }

    void submit(std::unique_ptr<Task> t, fnec_deck *deck)
    {
        fnec_request q = req_;
        q.rp_theta = t->theta.empty() ? nullptr : t->theta.data();
        q.rp_phi = t->phi.empty() ? nullptr : t->phi.data();
        q.rp_npts = t->theta.size();
        t->x.resize(q.n + std::max<size_t>(q.n_wires, 1));
        q.x_out = t->x.data();
        t->rows.resize(q.rp_npts);
        q.rp_out = t->rows.empty() ? nullptr : t->rows.data();
        const int prec = t->precision;
        std::string err;
        int st = FNEC_OK;
        {
            std::lock_guard<std::mutex> lk(mu_);
            if ((int)tasks_.size() >= opt_.max_queue) {
                st = FNEC_RESOURCE_EXHAUSTED;
                err = "worker queue is full (" + std::to_string(opt_.max_queue) + " tasks in flight)";
            } else if (!session(prec)) {
                st = FNEC_UNAVAILABLE;
                err = fnec_hip_device_count() <= 0 ? std::string("no usable gfx950 device")
                                                   : std::string("session: ") + fnec_hip_last_error();
            } else {
                q.tag = next_tag_++;
                st = fnec_hip_session_submit(sessions_[prec], &q);
                if (st == FNEC_OK) {
                    tasks_[q.tag] = std::move(t);
                    start_poller();
                } else {
                    err = fnec_hip_last_error();
                }
            }
        }
        fnec_deck_free(deck); // submit copied the inputs
        if (st == FNEC_OK)
            return;
        const char *code = st == FNEC_RESOURCE_EXHAUSTED ? "resource_exhausted"
                           : st == FNEC_INVALID_ARG || st == FNEC_DIM_MISMATCH ? "parse_error"
                                                                               : "internal";
        emit(error_line(t->task_id, t->freq, code, err));
    }

    // lazily created, one per precision; called with mu_ held
    fnec_session *session(int prec)
    {
        if (!sessions_[prec]) {
            if (fnec_hip_device_count() <= 0) {
                return nullptr;
            }
            sessions_[prec] =
                fnec_hip_session_create((size_t)(opt_.budget_gib * (double)(1ull << 30)), prec, opt_.max_batch);
        }
        return sessions_[prec];
    }

    void start_poller()
    {
        if (!poller_.joinable())
            poller_ = std::thread([this] { poll_loop(); });
    }

    void stop_poller()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            quit_ = true;
        }
        if (poller_.joinable())
            poller_.join();
    }

    void poll_loop()
    {
        std::vector<fnec_result> res(256);
        while (true) {
            bool idle;
            {
                std::lock_guard<std::mutex> lk(mu_);
                idle = tasks_.empty();
                if (idle && quit_)
                    return;
            }
            if (idle) {
                std::this_thread::sleep_for(std::chrono::microseconds(200));
                continue;
            }
            for (int p = 0; p < 3; ++p) {
                fnec_session *s;
                {
                    std::lock_guard<std::mutex> lk(mu_);
                    s = sessions_[p];
                }
                if (!s)
                    continue;
                const int k = fnec_hip_session_poll(s, res.data(), (int)res.size(), 1000);
                for (int j = 0; j < k; ++j) {
                    const fnec_result &r = res[j];
                    std::unique_ptr<Task> t;
                    {
                        std::lock_guard<std::mutex> lk(mu_);
                        auto it = tasks_.find(r.tag);
                        if (it == tasks_.end())
                            continue;
                        t = std::move(it->second);
                        tasks_.erase(it);
                    }
                    if (r.status == FNEC_OK) {
                        const std::string line = ok_line(*t, r);
                        if (!t->key.empty())
                            cache_put(t->key, line.substr(line.find(",\"frequency_hz\"")));
                        emit(line);
                    } else if (r.status == FNEC_SINGULAR)
                        emit(error_line(t->task_id, t->freq, "singular_matrix", "impedance matrix is singular"));
                    else if (r.status == FNEC_RESOURCE_EXHAUSTED)
                        emit(error_line(t->task_id, t->freq, "resource_exhausted", "device memory exhausted"));
                    else
                        emit(error_line(t->task_id, t->freq, "internal",
                                        std::string("solve failed (status ") + std::to_string(r.status) + ")"));
                    std::lock_guard<std::mutex> lk(mu_);
                    if (tasks_.empty())
                        cv_done_.notify_all();
                }
            }
        }
    }

    // FIFO result cache (ResultCache::insert, result_cache.rs:76-96): a key
    // already present is replaced in place; at capacity the oldest goes first
    void cache_put(const std::string &key, const std::string &tail)
    {
        std::lock_guard<std::mutex> lk(cache_mu_);
        auto it = cache_.find(key);
        if (it != cache_.end()) {
            it->second = tail;
            return;
        }
        while (cache_.size() >= opt_.cache && !cache_order_.empty()) {
            cache_.erase(cache_order_.front());
            cache_order_.pop_front();
        }
        cache_order_.push_back(key);
        cache_[key] = tail;
    }

    Options opt_;
    std::mutex cache_mu_;
    std::unordered_map<std::string, std::string> cache_; // key -> reply after its task_id
    std::deque<std::string> cache_order_;
    fnec_request req_{};
    std::mutex mu_;
    std::condition_variable cv_done_;
    std::map<uint64_t, std::unique_ptr<Task>> tasks_;
    fnec_session *sessions_[3] = {nullptr, nullptr, nullptr};
    uint64_t next_tag_ = 1;
    bool quit_ = false;
    std::thread poller_;
};

// worker.rs:38-44: {"cmd":"shutdown"} as any JSON object
bool is_shutdown(const std::string &line)
{
    JVal v;
    JParser p(line);
    if (!p.document(v))
        return false;
    const JVal *c = v.get("cmd");
    return c && c->kind == JVal::Str && c->str == "shutdown";
}

std::string trim(const std::string &s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a]))
        ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1]))
        --b;
    return s.substr(a, b - a);
}

int usage()
{
    std::fprintf(stderr,
                 "usage: fnec_hip_worker [--serve | --serial] [--budget-gib G] [--max-batch B] [--max-queue Q]\n"
                 "                       [--cache N]\n"
                 "       fnec_hip_worker --vswr RE IM [Z0]\n"
                 "       fnec_hip_worker --cache-key BASIS GROUND_MODEL FREQ_HZ < deck\n");
    return 2;
}

} // namespace

int main(int argc, char **argv)
{
    Options opt;
    for (int a = 1; a < argc; ++a) {
        const std::string s = argv[a];
        if (s == "--serve") {
            opt.serial = false;
        } else if (s == "--serial") {
            opt.serial = true;
        } else if (s == "--budget-gib" && a + 1 < argc) {
            opt.budget_gib = std::atof(argv[++a]);
        } else if (s == "--max-batch" && a + 1 < argc) {
            opt.max_batch = std::atoi(argv[++a]);
        } else if (s == "--max-queue" && a + 1 < argc) {
            opt.max_queue = std::max(1, std::atoi(argv[++a]));
        } else if (s == "--cache" && a + 1 < argc) {
            opt.cache = (size_t)std::max(0, std::atoi(argv[++a]));
        } else if (s == "--cache-key" && a + 3 < argc) {
            // the reference's cache_key of the deck text on stdin
            std::string deck((std::istreambuf_iterator<char>(std::cin)), std::istreambuf_iterator<char>());
            std::printf("%s\n", cache_key(deck, argv[a + 1], argv[a + 2], std::atof(argv[a + 3])).c_str());
            return 0;
        } else if (s == "--vswr" && a + 2 < argc) {
            const double z0 = a + 3 < argc ? std::atof(argv[a + 3]) : 50.0;
            std::printf("%s\n", jnum(vswr(std::atof(argv[a + 1]), std::atof(argv[a + 2]), z0)).c_str());
            return 0;
        } else {
            return usage();
        }
    }
    Worker w(opt);
    std::string line;
    while (std::getline(std::cin, line)) {
        const std::string t = trim(line);
        if (t.empty())
            continue;
        if (is_shutdown(t))
            break;
        w.handle_line(t);
    }
    w.drain();
    return 0;
}
