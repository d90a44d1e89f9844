// Host side above the fnec_hip ABI (include/fnec_host.h): NEC deck parsing,
// geometry, wire topology, Hallén excitation vectors, the CLI pipeline and
// report contract v1. Plain C++; the numerics of each builder follow the
// reference's operation order so that segment data and right-hand sides are
// the same doubles the reference's callers pass to nec_accel.
#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fnec_hip.h"
#include "../../include/fnec_host.h"
#include "engine.h"

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kC0 = 299792458.0;
constexpr double kMu0 = 4.0 * kPi * 1e-7;
constexpr double kEta0 = kMu0 * kC0;

inline double deg2rad(double d) { return d * (kPi / 180.0); }
inline double rad2deg(double r) { return r * (180.0 / kPi); }

struct Card {
    std::string mn;
    int nf = 0;
    double f[16] = {0};
    std::vector<std::string> raw; // NT: the fields as written (card.rs NtCard::raw_fields)
};

uint32_t as_u32(double v)
{
    if (std::isnan(v) || v <= 0.0)
        return 0;
    if (v >= 4294967295.0)
        return 4294967295u;
    return (uint32_t)v;
}

int32_t as_i32(double v)
{
    if (std::isnan(v))
        return 0;
    if (v <= -2147483648.0)
        return INT32_MIN;
    if (v >= 2147483647.0)
        return INT32_MAX;
    return (int32_t)v;
}

// Rust str::parse::<f64> accepts [+-]digits[.digits][e[+-]digits] and inf/nan.
bool parse_f64(const std::string &s, double &out)
{
    size_t p = 0;
    if (p < s.size() && (s[p] == '+' || s[p] == '-'))
        ++p;
    std::string rest = s.substr(p);
    std::string low = rest;
    for (auto &c : low)
        c = (char)std::tolower((unsigned char)c);
    if (low == "inf" || low == "infinity" || low == "nan") {
        out = std::strtod(s.c_str(), nullptr);
        return true;
    }
    size_t q = p, digits = 0;
    while (q < s.size() && std::isdigit((unsigned char)s[q])) {
        ++q;
        ++digits;
    }
    if (q < s.size() && s[q] == '.') {
        ++q;
        while (q < s.size() && std::isdigit((unsigned char)s[q])) {
            ++q;
            ++digits;
        }
    }
    if (!digits)
        return false;
    if (q < s.size() && (s[q] == 'e' || s[q] == 'E')) {
        ++q;
        if (q < s.size() && (s[q] == '+' || s[q] == '-'))
            ++q;
        if (q >= s.size() || !std::isdigit((unsigned char)s[q]))
            return false;
        while (q < s.size() && std::isdigit((unsigned char)s[q]))
            ++q;
    }
    if (q != s.size())
        return false;
    out = std::strtod(s.c_str(), nullptr);
    return true;
}

struct Seg {
    uint32_t tag, idx;
    double start[3], end[3], mid[3], dir[3];
    double length, radius;
};

struct Ground {
    int kind = 0; // 0 free, 1 PEC, 2 finite, 3 deferred (treated as free)
    double eps_r = 0.0, sigma = 0.0;
};

} // namespace

struct fnec_deck {
    std::vector<Card> cards;
    // built geometry (lazily)
    bool geo_built = false;
    int geo_status = FNEC_OK;
    std::string geo_err;
    std::vector<Seg> segs;
    // storage for the last fnec_deck_problem
    std::vector<fnec_seg> fsegs;
    std::vector<fnec_c64> rhs, feed_v;
    std::vector<double> cosv, js, theta, phi;
    std::vector<size_t> wf, wl, ja, jb, path_of, free_ends, feed_idx;
    std::vector<uint32_t> feed_tag, feed_seg, seg_tag, seg_index;
    std::vector<fnec_stamp> stamps;
    // plane-wave / current-source storage
    std::vector<double> sinv, recv_theta, recv_phi;
    std::vector<fnec_c64> rhs_all;
    std::vector<unsigned char> feed_port;
};

namespace {

bool card_is(const Card &c, const char *m) { return c.mn == m; }
int ex_type(const Card &c) { return (int)as_u32(c.f[0]); }
uint32_t ex_tag(const Card &c) { return as_u32(c.f[1]); }
uint32_t ex_seg(const Card &c) { return as_u32(c.f[2]); }
double ex_vr(const Card &c) { return c.nf > 4 ? c.f[4] : 0.0; }
double ex_vi(const Card &c) { return c.nf > 5 ? c.f[5] : 0.0; }
bool ex_voltage(const Card &c) { int t = ex_type(c); return t == 0 || t == 5; }
bool ex_plane(const Card &c) { int t = ex_type(c); return t >= 1 && t <= 3; }
bool ex_current(const Card &c) { return ex_type(c) == 4; }

int find_seg(const std::vector<Seg> &s, uint32_t tag, uint32_t idx)
{
    for (size_t i = 0; i < s.size(); ++i)
        if (s[i].tag == tag && s[i].idx == idx)
            return (int)i;
    return -1;
}

// ---- geometry (geometry.rs:676-879) ----

void rot_x(double p[3], double c, double s) { double y = c * p[1] - s * p[2], z = s * p[1] + c * p[2]; p[1] = y; p[2] = z; }
void rot_y(double p[3], double c, double s) { double x = c * p[0] + s * p[2], z = -s * p[0] + c * p[2]; p[0] = x; p[2] = z; }
void rot_z(double p[3], double c, double s) { double x = c * p[0] - s * p[1], y = s * p[0] + c * p[1]; p[0] = x; p[1] = y; }

void recompute_dir(Seg &g)
{
    double dx = g.end[0] - g.start[0], dy = g.end[1] - g.start[1], dz = g.end[2] - g.start[2];
    double len = std::sqrt(dx * dx + dy * dy + dz * dz);
    if (len == 0.0) {
        g.dir[0] = 0.0; g.dir[1] = 0.0; g.dir[2] = 1.0;
    } else {
        g.dir[0] = dx / len; g.dir[1] = dy / len; g.dir[2] = dz / len;
    }
}

void gm_point(double p[3], const Card &c)
{
    rot_x(p, std::cos(deg2rad(c.f[2])), std::sin(deg2rad(c.f[2])));
    rot_y(p, std::cos(deg2rad(c.f[3])), std::sin(deg2rad(c.f[3])));
    rot_z(p, std::cos(deg2rad(c.f[4])), std::sin(deg2rad(c.f[4])));
    p[0] = p[0] + c.f[5];
    p[1] = p[1] + c.f[6];
    p[2] = p[2] + c.f[7];
}

bool tag_in_range(uint32_t t, uint32_t first, uint32_t last)
{
    return (first == 0 || t >= first) && (last == 0 || t <= last);
}

int build_geometry(const std::vector<Card> &cards, std::vector<Seg> &out, std::string &err)
{
    out.clear();
    for (const Card &c : cards) {
        if (card_is(c, "GW")) {
            uint32_t tag = as_u32(c.f[0]), ns = as_u32(c.f[1]);
            double x1 = c.f[2], y1 = c.f[3], z1 = c.f[4];
            double wdx = c.f[5] - x1, wdy = c.f[6] - y1, wdz = c.f[7] - z1;
            if (ns == 0) {
                err = "GW tag " + std::to_string(tag) + ": segment count must be >= 1";
                return FNEC_INVALID_ARG;
            }
            double wl = std::sqrt(wdx * wdx + wdy * wdy + wdz * wdz);
            if (wl == 0.0) {
                err = "GW tag " + std::to_string(tag) + ": wire has zero length (start == end)";
                return FNEC_INVALID_ARG;
            }
            if (out.size() + ns > ((size_t)1 << 20)) {
                err = "too many segments";
                return FNEC_INVALID_ARG;
            }
            double dir[3] = {wdx / wl, wdy / wl, wdz / wl};
            double nn = (double)ns, seglen = wl / nn;
            for (uint32_t i = 0; i < ns; ++i) {
                Seg g;
                double t0 = (double)i / nn, t1 = ((double)i + 1.0) / nn, tm = (t0 + t1) * 0.5;
                g.tag = tag;
                g.idx = i + 1;
                g.start[0] = x1 + wdx * t0; g.start[1] = y1 + wdy * t0; g.start[2] = z1 + wdz * t0;
                g.end[0] = x1 + wdx * t1;   g.end[1] = y1 + wdy * t1;   g.end[2] = z1 + wdz * t1;
                g.mid[0] = x1 + wdx * tm;   g.mid[1] = y1 + wdy * tm;   g.mid[2] = z1 + wdz * tm;
                std::memcpy(g.dir, dir, sizeof(dir));
                g.length = seglen;
                g.radius = c.f[8];
                out.push_back(g);
            }
        } else if (card_is(c, "GM")) {
            uint32_t inc = as_u32(c.f[0]), last = as_u32(c.f[1]), first = as_u32(c.f[8]);
            size_t n0 = out.size();
            if (inc == 0) {
                for (size_t i = 0; i < n0; ++i) {
                    Seg &g = out[i];
                    if (!tag_in_range(g.tag, first, last))
                        continue;
                    gm_point(g.start, c);
                    gm_point(g.end, c);
                    gm_point(g.mid, c);
                    recompute_dir(g);
                }
            } else {
                for (size_t i = 0; i < n0; ++i) {
                    Seg g = out[i];
                    if (!tag_in_range(g.tag, first, last))
                        continue;
                    g.tag += inc;
                    gm_point(g.start, c);
                    gm_point(g.end, c);
                    gm_point(g.mid, c);
                    recompute_dir(g);
                    out.push_back(g);
                }
            }
        } else if (card_is(c, "GR")) {
            uint32_t inc = as_u32(c.f[0]), count = as_u32(c.f[1]);
            std::vector<Seg> orig = out;
            for (uint32_t k = 1; k <= count; ++k) {
                double ang = c.f[2] * (double)k;
                double ca = std::cos(deg2rad(ang)), sa = std::sin(deg2rad(ang));
                for (const Seg &o : orig) {
                    Seg g = o;
                    g.tag += inc * k;
                    rot_z(g.start, ca, sa);
                    rot_z(g.end, ca, sa);
                    rot_z(g.mid, ca, sa);
                    recompute_dir(g);
                    out.push_back(g);
                }
            }
        }
    }
    if (out.empty()) {
        err = "deck contains no GW (wire) cards";
        return FNEC_INVALID_ARG;
    }
    return FNEC_OK;
}

Ground ground_model(const std::vector<Card> &cards)
{
    Ground g;
    for (const Card &c : cards) {
        if (!card_is(c, "GN"))
            continue;
        int t = as_i32(c.f[0]);
        if (t == -1)
            g.kind = 0;
        else if (t == 1)
            g.kind = 1;
        else if (t == 0 || t == 2) {
            g.kind = 2;
            g.eps_r = c.nf > 4 ? c.f[4] : 13.0;
            g.sigma = c.nf > 5 ? c.f[5] : 0.005;
        } else
            g.kind = 3;
        return g;
    }
    for (const Card &c : cards)
        if (card_is(c, "GE") && c.nf > 0 && as_i32(c.f[0]) == 1) {
            g.kind = 1;
            return g;
        }
    return g;
}

// ---- wire topology (geometry.rs:168-339) ----

std::vector<std::pair<size_t, size_t>> wire_endpoints(const std::vector<Seg> &s)
{
    std::vector<std::pair<size_t, size_t>> out;
    uint32_t cur = 0xFFFFFFFFu;
    size_t first = 0;
    for (size_t i = 0; i < s.size(); ++i) {
        if (s[i].tag != cur) {
            if (cur != 0xFFFFFFFFu)
                out.push_back({first, i - 1});
            cur = s[i].tag;
            first = i;
        }
    }
    if (cur != 0xFFFFFFFFu)
        out.push_back({first, s.size() - 1});
    return out;
}

double dist2(const double *a, const double *b)
{
    double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
    return d0 * d0 + d1 * d1 + d2 * d2;
}

bool dir_aligned(const double *a, const double *b, double tol)
{
    double cr[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    double z[3] = {0, 0, 0};
    double dot = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    return dist2(cr, z) < tol * tol && dot > 0.0;
}

std::vector<std::pair<size_t, size_t>> merge_collinear(const std::vector<Seg> &s)
{
    auto base = wire_endpoints(s);
    if (base.size() < 2)
        return base;
    const double tol = 1e-6;
    std::vector<std::pair<size_t, size_t>> merged;
    auto cur = base[0];
    for (size_t i = 1; i < base.size(); ++i) {
        const Seg &a = s[cur.second], &b = s[base[i].first];
        bool contiguous = cur.second + 1 == base[i].first;
        bool connects = dist2(a.end, b.start) < tol * tol;
        bool collinear = dir_aligned(a.dir, b.dir, tol);
        bool same_r = std::fabs(a.radius - b.radius) < tol;
        if (contiguous && connects && collinear && same_r)
            cur.second = base[i].second;
        else {
            merged.push_back(cur);
            cur = base[i];
        }
    }
    merged.push_back(cur);
    return merged;
}

struct Junction {
    size_t a, b;
    double sign;
};

std::vector<Junction> detect_junctions(const std::vector<Seg> &s,
                                       const std::vector<std::pair<size_t, size_t>> &wires, double tol)
{
    struct Node {
        size_t seg;
        const double *p;
        bool start;
    };
    std::vector<Node> nodes;
    for (auto &w : wires) {
        nodes.push_back({w.first, s[w.first].start, true});
        nodes.push_back({w.second, s[w.second].end, false});
    }
    std::vector<Junction> out;
    for (size_t i = 0; i < nodes.size(); ++i)
        for (size_t j = i + 1; j < nodes.size(); ++j) {
            if (nodes[i].seg == nodes[j].seg)
                continue;
            double dx = nodes[i].p[0] - nodes[j].p[0], dy = nodes[i].p[1] - nodes[j].p[1],
                   dz = nodes[i].p[2] - nodes[j].p[2];
            if (std::sqrt(dx * dx + dy * dy + dz * dz) > tol)
                continue;
            out.push_back({nodes[i].seg, nodes[j].seg, nodes[i].start == nodes[j].start ? 1.0 : -1.0});
        }
    return out;
}

struct Paths {
    bool ok = false;
    bool nontrivial = false;
    size_t count = 0;
    std::vector<size_t> path_of;
    std::vector<double> sign_of, s_of;
    std::vector<size_t> free_ends; // two per path
};

size_t uf_find(std::vector<size_t> &parent, size_t x)
{
    size_t r = x;
    while (parent[r] != r)
        r = parent[r];
    while (parent[x] != x) {
        size_t nx = parent[x];
        parent[x] = r;
        x = nx;
    }
    return r;
}

Paths conductor_paths(const std::vector<Seg> &s)
{
    Paths P;
    auto wires = wire_endpoints(s);
    const size_t w = wires.size(), n = s.size();
    P.path_of.assign(n, 0);
    P.sign_of.assign(n, 1.0);
    P.s_of.assign(n, 0.0);
    if (w == 0) {
        P.ok = true;
        return P;
    }
    const double tol = 1e-6;
    auto ep = [&](size_t wi, int end) -> const double * {
        return end == 0 ? s[wires[wi].first].start : s[wires[wi].second].end;
    };
    std::vector<long> node_of(2 * w, -1);
    std::vector<std::vector<size_t>> members;
    for (size_t id = 0; id < 2 * w; ++id) {
        if (node_of[id] >= 0)
            continue;
        const double *p = ep(id / 2, (int)(id & 1));
        long node = (long)members.size();
        members.push_back({id});
        node_of[id] = node;
        for (size_t j = 0; j < 2 * w; ++j) {
            if (node_of[j] >= 0)
                continue;
            if (dist2(p, ep(j / 2, (int)(j & 1))) < tol * tol) {
                node_of[j] = node;
                members.back().push_back(j);
            }
        }
    }
    for (auto &m : members)
        if (m.size() >= 3)
            return P;
    std::vector<size_t> parent(w);
    for (size_t i = 0; i < w; ++i)
        parent[i] = i;
    for (auto &m : members)
        if (m.size() == 2) {
            size_t a = m[0] / 2, b = m[1] / 2;
            if (a == b)
                return P;
            size_t ra = uf_find(parent, a), rb = uf_find(parent, b);
            parent[ra] = rb;
        }
    std::map<size_t, std::vector<size_t>> comp;
    for (size_t wi = 0; wi < w; ++wi)
        comp[uf_find(parent, wi)].push_back(wi);
    for (auto &kv : comp) {
        std::vector<size_t> fe;
        for (size_t wi : kv.second)
            for (int end = 0; end < 2; ++end)
                if (members[node_of[wi * 2 + end]].size() == 1)
                    fe.push_back(wi * 2 + end);
        if (fe.size() != 2)
            return P;
        size_t cur = fe[0] / 2;
        int enter = (int)(fe[0] & 1);
        std::vector<char> visited(w, 0);
        std::vector<size_t> order;
        double arc = 0.0;
        for (;;) {
            visited[cur] = 1;
            size_t first = wires[cur].first, last = wires[cur].second;
            bool fwd = enter == 0;
            double sg = fwd ? 1.0 : -1.0;
            for (size_t k = 0; k <= last - first; ++k) {
                size_t m = fwd ? first + k : last - k;
                double half = s[m].length * 0.5;
                P.s_of[m] = arc + half;
                arc += s[m].length;
                order.push_back(m);
                P.path_of[m] = P.count;
                P.sign_of[m] = sg;
            }
            int exit_end = 1 - enter;
            long exit_node = node_of[cur * 2 + exit_end];
            if (members[exit_node].size() == 1)
                break;
            long nw = -1;
            int ne = 0;
            for (size_t mem : members[exit_node])
                if (mem / 2 != cur && !visited[mem / 2]) {
                    nw = (long)(mem / 2);
                    ne = (int)(mem & 1);
                }
            if (nw < 0)
                break;
            cur = (size_t)nw;
            enter = ne;
        }
        for (size_t m : order)
            P.s_of[m] = P.s_of[m] - arc * 0.5;
        for (size_t i = 0; i < order.size(); ++i) {
            if (P.sign_of[order[i]] != 1.0)
                P.nontrivial = true;
            if (i > 0 && order[i] != order[i - 1] + 1)
                P.nontrivial = true;
        }
        P.free_ends.push_back(order.front());
        P.free_ends.push_back(order.back());
        P.count++;
    }
    P.ok = true;
    return P;
}

// ---- excitation (excitation.rs:78-341) ----

int build_excitation(const std::vector<Card> &cards, const std::vector<Seg> &s, std::vector<fnec_c64> &v)
{
    v.assign(s.size(), fnec_c64{0.0, 0.0});
    for (const Card &c : cards) {
        if (!card_is(c, "EX") || ex_plane(c) || ex_current(c))
            continue;
        if (!ex_voltage(c))
            return FNEC_INVALID_ARG;
        int idx = find_seg(s, ex_tag(c), ex_seg(c));
        if (idx < 0)
            return FNEC_INVALID_ARG;
        double dl = s[idx].length;
        v[idx].re = v[idx].re + ex_vr(c) / dl;
        v[idx].im = v[idx].im + ex_vi(c) / dl;
    }
    return FNEC_OK;
}

// (0, b) * (vr, vi) with num-complex multiplication order
inline void add_src(fnec_c64 &acc, double b, double vr, double vi, double scale_sign)
{
    double re = 0.0 * vr - b * vi, im = 0.0 * vi + b * vr;
    if (scale_sign != 1.0) {
        re *= scale_sign;
        im *= scale_sign;
    }
    acc.re = acc.re + re;
    acc.im = acc.im + im;
}

int hallen_rhs(const std::vector<Card> &cards, const std::vector<Seg> &s, double freq,
               std::vector<fnec_c64> &rhs, std::vector<double> &cosv)
{
    const size_t n = s.size();
    const Card *first = nullptr;
    for (const Card &c : cards) {
        if (!card_is(c, "EX") || ex_plane(c) || ex_current(c))
            continue;
        if (!ex_voltage(c))
            return FNEC_INVALID_ARG;
        if (!first)
            first = &c;
    }
    rhs.assign(n, fnec_c64{0.0, 0.0});
    cosv.assign(n, 0.0);
    if (!first)
        return FNEC_OK;
    int feed = find_seg(s, ex_tag(*first), ex_seg(*first));
    if (feed < 0)
        return FNEC_INVALID_ARG;
    const double k = 2.0 * kPi * freq / kC0;
    const double scale = 2.0 * kPi / kEta0;
    const uint32_t driven = s[feed].tag;
    std::map<uint32_t, std::pair<size_t, fnec_c64>> src;
    src[driven] = {(size_t)feed, fnec_c64{ex_vr(*first), ex_vi(*first)}};
    for (const Card &c : cards) {
        if (!card_is(c, "EX") || !ex_voltage(c) || ex_tag(c) == driven)
            continue;
        int idx = find_seg(s, ex_tag(c), ex_seg(c));
        if (idx >= 0)
            src[ex_tag(c)] = {(size_t)idx, fnec_c64{ex_vr(c), ex_vi(c)}};
    }
    auto comps = merge_collinear(s);
    std::vector<size_t> comp_of(n, 0);
    for (size_t ci = 0; ci < comps.size(); ++ci)
        for (size_t m = comps[ci].first; m <= comps[ci].second && m < n; ++m)
            comp_of[m] = ci;
    for (size_t m = 0; m < n; ++m) {
        const Seg &a = s[comps[comp_of[m]].first], &b = s[comps[comp_of[m]].second];
        double wm[3] = {(a.mid[0] + b.mid[0]) / 2.0, (a.mid[1] + b.mid[1]) / 2.0, (a.mid[2] + b.mid[2]) / 2.0};
        double dl[3] = {s[m].mid[0] - wm[0], s[m].mid[1] - wm[1], s[m].mid[2] - wm[2]};
        double sl = dl[0] * a.dir[0] + dl[1] * a.dir[1] + dl[2] * a.dir[2];
        cosv[m] = std::cos(k * sl);
    }
    for (auto &kv : src) {
        size_t fi = kv.second.first;
        fnec_c64 v = kv.second.second;
        auto cr = comps[comp_of[fi]];
        for (size_t m = cr.first; m <= cr.second; ++m) {
            double d[3] = {s[m].mid[0] - s[fi].mid[0], s[m].mid[1] - s[fi].mid[1], s[m].mid[2] - s[fi].mid[2]};
            double sv = d[0] * s[fi].dir[0] + d[1] * s[fi].dir[1] + d[2] * s[fi].dir[2];
            add_src(rhs[m], -scale * std::sin(k * std::fabs(sv)), v.re, v.im, 1.0);
        }
    }
    return FNEC_OK;
}

int hallen_rhs_paths(const std::vector<Card> &cards, const std::vector<Seg> &s, double freq,
                     const Paths &P, std::vector<fnec_c64> &rhs, std::vector<double> &cosv)
{
    const size_t n = s.size();
    const double k = 2.0 * kPi * freq / kC0;
    const double scale = 2.0 * kPi / kEta0;
    cosv.assign(n, 0.0);
    rhs.assign(n, fnec_c64{0.0, 0.0});
    for (size_t m = 0; m < n; ++m)
        cosv[m] = P.sign_of[m] * std::cos(k * P.s_of[m]);
    for (const Card &c : cards) {
        if (!card_is(c, "EX") || ex_plane(c) || ex_current(c))
            continue;
        if (!ex_voltage(c))
            return FNEC_INVALID_ARG;
        int fi = find_seg(s, ex_tag(c), ex_seg(c));
        if (fi < 0)
            return FNEC_INVALID_ARG;
        const double vr = ex_vr(c), vi = ex_vi(c);
        const size_t sp = P.path_of[fi];
        const double ssrc = P.s_of[fi], fs = P.sign_of[fi];
        for (size_t m = 0; m < n; ++m) {
            if (P.path_of[m] != sp)
                continue;
            double ds = std::fabs(P.s_of[m] - ssrc);
            add_src(rhs[m], -scale * std::sin(k * ds), vr, vi, P.sign_of[m] * fs);
        }
    }
    return FNEC_OK;
}

// ---- plane-wave and current-source inputs (planewave.rs:54-347,
// excitation.rs:350-447) ----
//
// Where the reference takes the sine and the cosine of one angle, its
// optimised code makes one sincos call; these builders call sincos there, and
// sin alone where the reference needs only the sine, so every value is the
// same libm result the reference's build computes.

double card_f(const Card &c, int i) { return c.nf > i ? c.f[i] : 0.0; }

fnec_c64 cmul(fnec_c64 a, fnec_c64 b) { return fnec_c64{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }

// IncidentPlaneWave::from_ex_card (planewave.rs:59-76) at incidence (theta,
// phi) in degrees: the propagation unit vector r_hat and the complex
// polarisation vector major + j sense AR minor (planewave.rs:79-118).
struct Wave {
    double r_hat[3];
    fnec_c64 pol[3];
};

Wave plane_wave(const Card &ex, double theta_deg, double phi_deg)
{
    const int type = ex_type(ex);
    const double ar = type == 1 ? 0.0 : card_f(ex, 9);
    const double sense = type == 2 ? 1.0 : (type == 3 ? -1.0 : 0.0);
    double st, ct, sp, cp, se, ce;
    sincos(deg2rad(theta_deg), &st, &ct);
    sincos(deg2rad(phi_deg), &sp, &cp);
    sincos(deg2rad(card_f(ex, 6)), &se, &ce);
    Wave w;
    w.r_hat[0] = st * cp;
    w.r_hat[1] = st * sp;
    w.r_hat[2] = ct;
    const double th[3] = {ct * cp, ct * sp, -st}, ph[3] = {-sp, cp, 0.0};
    const double jm = sense * ar; // jminor = (0, sense AR)
    for (int i = 0; i < 3; ++i) {
        const double major = ce * th[i] + se * ph[i];
        const double minor = -se * th[i] + ce * ph[i];
        // (major, 0) + jminor * minor
        w.pol[i] = fnec_c64{major + 0.0 * minor, 0.0 + jm * minor};
    }
    return w;
}

// (pol . d) E0, E0 = 1: the sum of pol[c] d[c] from zero, then scaled by E0.
fnec_c64 coupling(const Wave &w, const double d[3])
{
    fnec_c64 acc{0.0, 0.0};
    for (int c = 0; c < 3; ++c) {
        acc.re = acc.re + w.pol[c].re * d[c];
        acc.im = acc.im + w.pol[c].im * d[c];
    }
    return fnec_c64{acc.re * 1.0, acc.im * 1.0};
}

// build_planewave_hallen (planewave.rs:153-243) on the raw wire ranges, or
// build_planewave_hallen_paths (planewave.rs:271-347) when P is given: for
// each incident wave q, the forcing rhs_q(s_m) = -j (2 pi / eta0) sum_{p on
// m's wire or path} E_t,q(s_p) sin(k |s_m - s_p|) dl_p (into rhs + q n), and
// the cos / sin columns of k s_m (the same for every wave). The kernel
// sin(k |s_m - s_p|) does not depend on the wave: each row's values are
// computed once and serve every wave (the receive sweep's incidences), in the
// reference's operation order for each.
// f(begin, end) over contiguous ranges of [0, count) on up to 16 host
// threads when the loop is large (work: its elementary operations); every item
// is computed by one thread with the operations of the serial loop.
template <class F> void parallel_ranges(size_t count, double work, F f)
{
    unsigned nt = std::thread::hardware_concurrency();
    nt = std::min(nt ? nt : 1u, 16u);
    if (work < 4e6 || nt < 2 || count < 2) {
        f((size_t)0, count);
        return;
    }
    const size_t per = (count + nt - 1) / nt;
    std::vector<std::thread> th;
    for (size_t b = per; b < count; b += per)
        th.emplace_back(f, b, std::min(count, b + per));
    f((size_t)0, std::min(count, per));
    for (auto &t : th)
        t.join();
}

void planewave_rhs(const std::vector<Wave> &waves, const std::vector<Seg> &s, double freq,
                   const std::vector<std::pair<size_t, size_t>> &wires, const Paths *P, fnec_c64 *rhs,
                   double *cosv, double *sinv)
{
    const size_t n = s.size(), nq = waves.size();
    const double k = 2.0 * kPi * freq / kC0;
    const double scale = 2.0 * kPi / kEta0;
    std::vector<size_t> grp(n, 0);
    std::vector<double> sc(n, 0.0), sg(n, 1.0);
    std::vector<fnec_c64> et(n * nq);
    std::vector<const double *> wdir;
    std::vector<std::array<double, 3>> wmid;
    if (P) {
        grp = P->path_of;
        sc = P->s_of;
        sg = P->sign_of;
    } else {
        for (size_t wi = 0; wi < wires.size(); ++wi) {
            for (size_t m = wires[wi].first; m <= wires[wi].second && m < n; ++m)
                grp[m] = wi;
            const Seg &a = s[wires[wi].first], &b = s[wires[wi].second];
            wdir.push_back(a.dir);
            wmid.push_back({(a.mid[0] + b.mid[0]) / 2.0, (a.mid[1] + b.mid[1]) / 2.0,
                            (a.mid[2] + b.mid[2]) / 2.0});
        }
    }
    size_t ngrp = 0;
    for (size_t i = 0; i < n; ++i)
        ngrp = std::max(ngrp, grp[i] + 1);
    // the incident field along each segment, for every wave
    parallel_ranges(n, 60.0 * n * nq, [&](size_t i0, size_t i1) {
        for (size_t i = i0; i < i1; ++i) {
            const double *dir = P ? s[i].dir : wdir[grp[i]];
            if (!P) {
                const double d[3] = {s[i].mid[0] - wmid[grp[i]][0], s[i].mid[1] - wmid[grp[i]][1],
                                     s[i].mid[2] - wmid[grp[i]][2]};
                sc[i] = d[0] * dir[0] + d[1] * dir[1] + d[2] * dir[2];
            }
            for (size_t q = 0; q < nq; ++q) {
                const Wave &wv = waves[q];
                const double phase = k * (wv.r_hat[0] * s[i].mid[0] + wv.r_hat[1] * s[i].mid[1] +
                                          wv.r_hat[2] * s[i].mid[2]);
                double sp, cp;
                sincos(phase, &sp, &cp);
                fnec_c64 e = cmul(coupling(wv, dir), fnec_c64{1.0 * cp, 1.0 * sp});
                if (P)
                    e = fnec_c64{e.re * sg[i], e.im * sg[i]};
                et[q * n + i] = e;
            }
        }
    });
    // members of each wire or path in ascending order (the reference's p loop)
    std::vector<std::vector<size_t>> members(ngrp);
    size_t summed = 0;
    for (size_t p = 0; p < n; ++p)
        members[grp[p]].push_back(p);
    for (size_t m = 0; m < n; ++m)
        summed += members[grp[m]].size();
    // each row's sums over its wire or path, for every wave
    parallel_ranges(n, 6.0 * (double)summed * nq, [&](size_t m0, size_t m1) {
        std::vector<double> kern;
        for (size_t m = m0; m < m1; ++m) {
            const std::vector<size_t> &mem = members[grp[m]];
            kern.resize(mem.size());
            for (size_t j = 0; j < mem.size(); ++j)
                kern[j] = std::sin(k * std::fabs(sc[m] - sc[mem[j]]));
            for (size_t q = 0; q < nq; ++q) {
                const fnec_c64 *e = et.data() + q * n;
                fnec_c64 acc{0.0, 0.0};
                for (size_t j = 0; j < mem.size(); ++j) {
                    const size_t p = mem[j];
                    const double tr = e[p].re * kern[j] * s[p].length, ti = e[p].im * kern[j] * s[p].length;
                    acc.re = acc.re + tr;
                    acc.im = acc.im + ti;
                }
                fnec_c64 r = cmul(fnec_c64{0.0, -scale}, acc);
                if (P)
                    r = fnec_c64{r.re * sg[m], r.im * sg[m]};
                rhs[q * n + m] = r;
            }
            double sv, cv;
            sincos(k * sc[m], &sv, &cv);
            cosv[m] = P ? sg[m] * cv : cv;
            sinv[m] = P ? sg[m] * sv : sv;
        }
    });
}

// build_current_source_shape(_paths) (excitation.rs:350-447): the delta-gap
// forcing and cos column of the deck with its first EX card replaced by a
// unit voltage source at (tag, seg) and every other EX card dropped
// (excitation.rs:365-392).
int current_source_shape(const std::vector<Card> &cards, const std::vector<Seg> &s, double freq,
                         uint32_t tag, uint32_t seg, const Paths *P, std::vector<fnec_c64> &shape,
                         std::vector<double> &cosv, int &src)
{
    src = find_seg(s, tag, seg);
    if (src < 0)
        return FNEC_INVALID_ARG;
    std::vector<Card> u;
    bool placed = false;
    for (const Card &c : cards) {
        if (card_is(c, "EX")) {
            if (!placed) {
                Card e;
                e.mn = "EX";
                e.nf = 6;
                e.f[0] = 0.0;
                e.f[1] = (double)tag;
                e.f[2] = (double)seg;
                e.f[3] = 0.0;
                e.f[4] = 1.0;
                e.f[5] = 0.0;
                u.push_back(e);
                placed = true;
            }
            continue;
        }
        u.push_back(c);
    }
    return P ? hallen_rhs_paths(u, s, freq, *P, shape, cosv) : hallen_rhs(u, s, freq, shape, cosv);
}

// ---- Rust number formatting (nec_report/src/lib.rs:437-482) ----

std::string fmt_fixed(double v, int prec)
{
    if (std::isnan(v))
        return "NaN";
    if (std::isinf(v))
        return v < 0 ? "-inf" : "inf";
    char b[64];
    std::snprintf(b, sizeof(b), "%.*f", prec, v);
    return b;
}

std::string fmt_exp(double v, int prec)
{
    if (std::isnan(v) || std::isinf(v))
        return fmt_fixed(v, prec);
    char b[64];
    std::snprintf(b, sizeof(b), "%.*e", prec, v);
    char *e = std::strchr(b, 'e');
    int ex = std::atoi(e + 1);
    *e = 0;
    return std::string(b) + "e" + std::to_string(ex);
}

// ---- LD / TL / NT stamps on Z (loads.rs:34-223, tl.rs:47-244,
// network.rs:55-140): complex arithmetic in num-complex 0.4's definitions
// (naive division, norm = hypot), as the reference computes them.
struct Cz {
    double re, im;
};
Cz cz(double r, double i) { return Cz{r, i}; }
Cz operator+(Cz a, Cz b) { return Cz{a.re + b.re, a.im + b.im}; }
Cz operator-(Cz a, Cz b) { return Cz{a.re - b.re, a.im - b.im}; }
Cz operator*(Cz a, Cz b) { return Cz{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
Cz operator*(Cz a, double s) { return Cz{a.re * s, a.im * s}; }
Cz operator/(Cz a, double s) { return Cz{a.re / s, a.im / s}; }
Cz operator/(Cz a, Cz b)
{
    const double ns = b.re * b.re + b.im * b.im;
    return Cz{(a.re * b.re + a.im * b.im) / ns, (a.im * b.re - a.re * b.im) / ns};
}
double cnorm(Cz a) { return std::hypot(a.re, a.im); }

constexpr double kTwoPiH = 2.0 * 3.14159265358979323846;
constexpr double kMu0H = 4.0 * 3.14159265358979323846 * 1e-7;
constexpr double kC0H = 299792458.0;

// I0(z) / I1(z) by the power series (|z| <= 15) or the asymptotic expansion
// (loads.rs:34-59)
Cz bessel_i0_over_i1(Cz z)
{
    if (cnorm(z) > 15.0) {
        const Cz inv = cz(1.0, 0.0) / z;
        return cz(1.0, 0.0) + inv * 0.5 + inv * inv * 0.375;
    }
    const Cz half = z * 0.5, half_sq = half * half;
    Cz i0 = cz(1.0, 0.0), t0 = cz(1.0, 0.0), i1 = half, t1 = half;
    for (int k = 1; k < 300; ++k) {
        const double kf = (double)k;
        t0 = t0 * (half_sq / (kf * kf));
        i0 = i0 + t0;
        t1 = t1 * (half_sq / (kf * (kf + 1.0)));
        i1 = i1 + t1;
        if (cnorm(t0) <= 1e-17 * cnorm(i0) && cnorm(t1) <= 1e-17 * cnorm(i1))
            break;
    }
    return i0 / i1;
}

// the round wire's internal (skin-effect) impedance (loads.rs:73-80)
Cz wire_skin_impedance(double sigma, double a, double dl, double omega)
{
    if (omega <= 0.0)
        return cz(dl / (sigma * 3.14159265358979323846 * a * a), 0.0);
    const double delta = std::sqrt(2.0 / (omega * kMu0H * sigma));
    const Cz gamma = cz(1.0 / delta, 1.0 / delta);
    return gamma / (kTwoPiH * a * sigma) * bessel_i0_over_i1(gamma * a) * dl;
}

// (tag, segment) -> segment index; segment 0 is the tag's centre, the lower
// one of an even count (tl.rs:186-244)
int stamp_seg(const std::vector<Seg> &s, uint32_t tag, uint32_t seg)
{
    if (seg != 0) {
        for (size_t i = 0; i < s.size(); ++i)
            if (s[i].tag == tag && s[i].idx == seg)
                return (int)i;
        return -1;
    }
    std::vector<size_t> of_tag;
    for (size_t i = 0; i < s.size(); ++i)
        if (s[i].tag == tag)
            of_tag.push_back(i);
    if (of_tag.empty())
        return -1;
    size_t pick;
    if (of_tag.size() % 2 == 1) {
        pick = of_tag.size() / 2;
    } else {
        const uint32_t lower = (uint32_t)(of_tag.size() / 2);
        pick = of_tag.size() / 2 >= 1 ? of_tag.size() / 2 - 1 : 0;
        for (size_t k = 0; k < of_tag.size(); ++k)
            if (s[of_tag[k]].idx == lower) {
                pick = k;
                break;
            }
    }
    return (int)of_tag[pick];
}

// str::parse::<u32>: optional '+', decimal digits, no overflow
bool parse_u32_strict(const std::string &t, uint32_t &out)
{
    size_t p = t.size() && t[0] == '+' ? 1 : 0;
    if (p >= t.size())
        return false;
    unsigned long long v = 0;
    for (; p < t.size(); ++p) {
        if (t[p] < '0' || t[p] > '9')
            return false;
        v = v * 10 + (unsigned)(t[p] - '0');
        if (v > 4294967295ull)
            return false;
    }
    out = (uint32_t)v;
    return true;
}

// The stamps in the order the reference adds them to Z: the LD loads (one
// diagonal stamp per loaded segment, its loads summed in card order), then
// the TL stamps, then (with_nt: the CLI route) the NT stamps. Cards the
// reference warns about and skips are skipped.
void build_stamps(const std::vector<Card> &cards, const std::vector<Seg> &s, double freq, bool with_nt,
                  std::vector<fnec_stamp> &out)
{
    out.clear();
    const size_t n = s.size();
    const double omega = kTwoPiH * freq;
    std::vector<Cz> loads(n, cz(0.0, 0.0));
    for (const Card &c : cards) {
        if (!card_is(c, "LD"))
            continue;
        const int32_t type = as_i32(c.f[0]);
        const uint32_t tag = as_u32(c.f[1]), first = as_u32(c.f[2]);
        uint32_t last = as_u32(c.f[3]);
        const double f1 = c.nf > 4 ? c.f[4] : 0.0, f2 = c.nf > 5 ? c.f[5] : 0.0, f3 = c.nf > 6 ? c.f[6] : 0.0;
        if (last == 0)
            last = first;
        for (size_t i = 0; i < n; ++i) {
            if (tag != 0 && s[i].tag != tag)
                continue;
            if (first != 0 && (s[i].idx < first || s[i].idx > last))
                continue;
            Cz z;
            if (type == 0) {
                z = cz(f1, omega * f2 - (f3 > 0.0 ? 1.0 / (omega * f3) : 0.0));
            } else if (type == 1) {
                Cz y = cz(0.0, 0.0);
                if (f1 > 0.0)
                    y = y + cz(1.0 / f1, 0.0);
                if (f2 > 0.0)
                    y = y + cz(0.0, -1.0 / (omega * f2));
                if (f3 > 0.0)
                    y = y + cz(0.0, omega * f3);
                z = cnorm(y) < 1e-30 ? cz(0.0, 0.0) : cz(1.0, 0.0) / y;
            } else if (type == 2) {
                z = cz(f1, omega * f2);
            } else if (type == 3) {
                z = cz(f1, -(f3 > 0.0 ? 1.0 / (omega * f3) : 0.0));
            } else if (type == 4) {
                z = cz(f1, f2);
            } else if (type == 5 && f1 > 0.0) {
                z = wire_skin_impedance(f1, s[i].radius, s[i].length, omega);
            } else {
                continue; // sigma <= 0 or an unsupported type: ignored with a warning
            }
            loads[i] = loads[i] + z;
        }
    }
    for (size_t i = 0; i < n; ++i)
        if (loads[i].re != 0.0 || loads[i].im != 0.0)
            out.push_back(fnec_stamp{i, i, fnec_c64{loads[i].re, loads[i].im}});
    auto two_port = [&](int a, int b, Cz zd, Cz zo) {
        out.push_back(fnec_stamp{(size_t)a, (size_t)a, fnec_c64{zd.re, zd.im}});
        out.push_back(fnec_stamp{(size_t)b, (size_t)b, fnec_c64{zd.re, zd.im}});
        out.push_back(fnec_stamp{(size_t)a, (size_t)b, fnec_c64{zo.re, zo.im}});
        out.push_back(fnec_stamp{(size_t)b, (size_t)a, fnec_c64{zo.re, zo.im}});
    };
    const double k = kTwoPiH * freq / kC0H;
    for (const Card &c : cards) {
        if (!card_is(c, "TL"))
            continue;
        const int a = stamp_seg(s, as_u32(c.f[0]), as_u32(c.f[1])), b = stamp_seg(s, as_u32(c.f[2]), as_u32(c.f[3]));
        const uint32_t type = as_u32(c.f[5]);
        const double z0 = c.f[6], len = c.f[7], f3 = c.nf > 8 ? c.f[8] : 1.0;
        if (a < 0 || b < 0 || a == b || z0 <= 0.0 || len <= 0.0)
            continue;
        if (type != 0) {
            // lossy: gamma l = F3 ln10 / 20 + j k len; Z0 coth, Z0 csch with
            // num-complex sinh = (sinh a cos b, cosh a sin b), cosh = (cosh a cos b, sinh a sin b)
            const double bl = k * len, al = f3 * 2.302585092994046 / 20.0;
            const Cz sh = cz(std::sinh(al) * std::cos(bl), std::cosh(al) * std::sin(bl));
            const Cz ch = cz(std::cosh(al) * std::cos(bl), std::sinh(al) * std::sin(bl));
            if (cnorm(sh) < 1e-9)
                continue;
            const Cz z0c = cz(z0, 0.0);
            two_port(a, b, z0c * ch / sh, z0c / sh);
            continue;
        }
        const double vf = (f3 >= 0.0 && f3 <= 1.0 && f3 > 0.0) ? f3 : 1.0;
        const double theta = k * len / vf, st = std::sin(theta);
        if (std::fabs(st) < 1e-9)
            continue;
        const double cot = std::cos(theta) / st, csc = 1.0 / st;
        two_port(a, b, cz(0.0, -z0 * cot), cz(0.0, -z0 * csc));
    }
    if (!with_nt)
        return;
    for (const Card &c : cards) {
        if (!card_is(c, "NT") || c.raw.size() < 10)
            continue;
        uint32_t t1, s1, t2, s2;
        if (!parse_u32_strict(c.raw[0], t1) || !parse_u32_strict(c.raw[1], s1) || !parse_u32_strict(c.raw[2], t2) ||
            !parse_u32_strict(c.raw[3], s2))
            continue;
        double y[6];
        bool ok = true;
        for (int q = 0; q < 6; ++q)
            ok = parse_f64(c.raw[4 + q], y[q]) && ok;
        if (!ok)
            continue;
        const int a = stamp_seg(s, t1, s1), b = stamp_seg(s, t2, s2);
        if (a < 0 || b < 0 || a == b)
            continue;
        const Cz y11 = cz(y[0], y[1]), y12 = cz(y[2], y[3]), y21 = y12, y22 = cz(y[4], y[5]);
        const Cz det = y11 * y22 - y12 * y21;
        if (cnorm(det) < 1e-30)
            continue;
        out.push_back(fnec_stamp{(size_t)a, (size_t)a, [&] { Cz z = y22 / det; return fnec_c64{z.re, z.im}; }()});
        out.push_back(fnec_stamp{(size_t)b, (size_t)b, [&] { Cz z = y11 / det; return fnec_c64{z.re, z.im}; }()});
        out.push_back(fnec_stamp{(size_t)a, (size_t)b,
                                 [&] { Cz z = cz(-y12.re, -y12.im) / det; return fnec_c64{z.re, z.im}; }()});
        out.push_back(fnec_stamp{(size_t)b, (size_t)a,
                                 [&] { Cz z = cz(-y21.re, -y21.im) / det; return fnec_c64{z.re, z.im}; }()});
    }
}

} // namespace

extern "C" {

fnec_deck *fnec_deck_parse(const char *text, char *err, size_t errlen)
{
    if (err && errlen)
        err[0] = 0;
    if (!text) {
        if (err && errlen)
            std::snprintf(err, errlen, "null deck text");
        return nullptr;
    }
    fnec_deck *d = new fnec_deck;
    const char *p = text;
    int lineno = 0;
    static const char *known[] = {"GW", "GE", "GM", "GR", "GN", "EX", "NE", "NH", "FR", "RP", "LD", "TL", "NT", "PT"};
    while (*p) {
        const char *eol = std::strchr(p, '\n');
        std::string line = eol ? std::string(p, eol) : std::string(p);
        p = eol ? eol + 1 : p + line.size();
        ++lineno;
        std::vector<std::string> toks;
        size_t i = 0;
        while (i < line.size()) {
            while (i < line.size() && std::isspace((unsigned char)line[i]))
                ++i;
            if (i >= line.size())
                break;
            size_t j = i;
            while (j < line.size() && !std::isspace((unsigned char)line[j]))
                ++j;
            toks.push_back(line.substr(i, j - i));
            i = j;
        }
        if (toks.empty())
            continue;
        std::string mn = toks[0];
        for (auto &c : mn)
            c = (char)std::toupper((unsigned char)c);
        if (mn == "CM" || mn == "CE")
            continue;
        if (mn == "EN")
            break;
        bool isknown = false;
        for (const char *k : known)
            if (mn == k)
                isknown = true;
        if (!isknown)
            continue; // unknown card: warning only (nec_parser lib.rs:403-408)
        int nf = (int)toks.size() - 1, need = 0, nparse = 0;
        if (mn == "GW") { need = 9; nparse = 9; }
        else if (mn == "GE") { nparse = nf > 0 ? 1 : 0; }
        else if (mn == "GM") { need = 9; nparse = 9; }
        else if (mn == "GR") { need = 3; nparse = 3; }
        else if (mn == "GN") { need = 1; nparse = nf > 5 ? 6 : (nf > 4 ? 5 : 1); }
        else if (mn == "EX" || mn == "NE" || mn == "NH") { need = 4; nparse = std::min(nf, 10); }
        else if (mn == "FR") { need = 4; nparse = nf >= 6 ? 6 : 4; }
        else if (mn == "RP") { need = 7; nparse = nf >= 8 ? 8 : 7; }
        else if (mn == "LD") { need = 4; nparse = std::min(nf, 7); }
        else if (mn == "TL") { need = 8; nparse = nf > 8 ? 9 : 8; }
        if (nf < need) {
            if (err && errlen)
                std::snprintf(err, errlen, "line %d: %s needs %d fields, got %d", lineno, mn.c_str(), need, nf);
            delete d;
            return nullptr;
        }
        Card c;
        c.mn = mn;
        c.nf = nf;
        if (mn == "NT")
            c.raw.assign(toks.begin() + 1, toks.end());
        for (int q = 0; q < nparse && q < 16; ++q)
            if (!parse_f64(toks[1 + q], c.f[q])) {
                if (err && errlen)
                    std::snprintf(err, errlen, "line %d: %s field %d: cannot parse '%s'", lineno,
                                  mn.c_str(), q + 1, toks[1 + q].c_str());
                delete d;
                return nullptr;
            }
        d->cards.push_back(c);
    }
    return d;
}

void fnec_deck_free(fnec_deck *d) { delete d; }

size_t fnec_deck_frequencies(const fnec_deck *d, int semantics, double *out, size_t max)
{
    if (!d)
        return 0;
    size_t cnt = 0;
    auto put = [&](double v) {
        if (out && cnt < max)
            out[cnt] = v;
        ++cnt;
    };
    for (const Card &c : d->cards) {
        if (!card_is(c, "FR"))
            continue;
        int fi = c.nf >= 6 ? 4 : 2;
        uint32_t type = as_u32(c.f[0]), steps = std::max<uint32_t>(as_u32(c.f[1]), 1);
        double f = c.f[fi], st = c.f[fi + 1];
        auto powi = [](double a, int b) {
            double r = 1.0;
            for (;;) {
                if (b & 1)
                    r *= a;
                b /= 2;
                if (b == 0)
                    break;
                a *= a;
            }
            return r;
        };
        if (semantics == FNEC_SEM_PY) {
            for (uint32_t i = 0; i < steps; ++i)
                put((type == 1 ? f * powi(st, (int)i) : f + st * (double)i) * 1e6);
            continue;
        }
        if (type == 0)
            for (uint32_t i = 0; i < steps; ++i)
                put((f + (double)i * st) * 1e6);
        else if (type == 1)
            for (uint32_t i = 0; i < steps; ++i)
                put(f * powi(st, (int)i) * 1e6);
        else
            put(f * 1e6);
        break;
    }
    return cnt;
}

size_t fnec_deck_segment_count(fnec_deck *d)
{
    if (!d)
        return 0;
    if (!d->geo_built) {
        d->geo_built = true;
        d->geo_status = build_geometry(d->cards, d->segs, d->geo_err);
    }
    if (d->geo_status) {
        fnec::set_error(d->geo_err);
        return 0;
    }
    return d->segs.size();
}

int fnec_deck_build(fnec_deck *d, int semantics, double freq_hz, fnec_deck_problem *out)
{
    if (!d || !out) {
        fnec::set_error("null deck or output");
        return FNEC_INVALID_ARG;
    }
    if (fnec_deck_segment_count(d) == 0)
        return FNEC_INVALID_ARG;
    // the first plane-wave card wins over current sources (solve_session.rs:1327-1337)
    int variant = 0;
    for (const Card &c : d->cards)
        if (card_is(c, "EX") && ex_plane(c))
            variant = 1;
        else if (card_is(c, "EX") && ex_current(c) && !variant)
            variant = 2;
    if (variant && semantics != FNEC_SEM_CLI) {
        fnec::set_error("plane-wave and current-source excitations stay on the CPU path of the worker and "
                        "the Python binding");
        return FNEC_UNAVAILABLE;
    }
    const std::vector<Seg> &s = d->segs;
    const size_t n = s.size();
    std::memset(out, 0, sizeof(*out));
    d->fsegs.resize(n);
    d->seg_tag.resize(n);
    d->seg_index.resize(n);
    for (size_t i = 0; i < n; ++i) {
        for (int c = 0; c < 3; ++c) {
            d->fsegs[i].mid[c] = s[i].mid[c];
            d->fsegs[i].dir[c] = s[i].dir[c];
        }
        d->fsegs[i].length = s[i].length;
        d->fsegs[i].radius = s[i].radius;
        d->seg_tag[i] = s[i].tag;
        d->seg_index[i] = s[i].idx;
    }
    Ground g = ground_model(d->cards);
    fnec_ground fg{g.kind == 1 ? FNEC_GROUND_PEC : (g.kind == 2 ? FNEC_GROUND_FINITE : FNEC_GROUND_FREE),
                   g.eps_r, g.sigma};
    std::vector<fnec_c64> vvec;
    if (semantics != FNEC_SEM_WORKER && build_excitation(d->cards, s, vvec)) {
        fnec::set_error("EX: bad excitation");
        return FNEC_INVALID_ARG;
    }
    d->wf.clear(); d->wl.clear(); d->ja.clear(); d->jb.clear(); d->js.clear();
    d->path_of.clear(); d->free_ends.clear();
    d->sinv.clear(); d->rhs_all.clear(); d->recv_theta.clear(); d->recv_phi.clear();
    int route = 0, src = -1;
    fnec_c64 i0{0.0, 0.0};
    if (variant) {
        // solve_plane_wave_hallen / solve_current_source_hallen
        // (solve_session.rs:262-301, 392-448): the conductor-path route when
        // some path is not trivial, else the raw wire ranges
        Paths P = conductor_paths(s);
        const bool paths = P.ok && P.count > 0 && P.nontrivial;
        auto wires = wire_endpoints(s);
        if (paths) {
            route = 1;
            d->path_of = P.path_of;
            d->free_ends = P.free_ends;
        } else
            for (auto &w : wires) {
                d->wf.push_back(w.first);
                d->wl.push_back(w.second);
            }
        if (variant == 1) {
            const Card *ex = nullptr;
            for (const Card &c : d->cards)
                if (!ex && card_is(c, "EX") && ex_plane(c))
                    ex = &c;
            if (!paths && !detect_junctions(s, wires, 1e-6).empty()) {
                fnec::set_error("EX: incident plane wave is supported on straight, non-junctioned wires; "
                                "junctioned geometry is not yet supported");
                return FNEC_INVALID_ARG;
            }
            // the card's own incidence, then the receive sweep's (NTHETA NPHI > 1)
            std::vector<std::pair<double, double>> inc{{card_f(*ex, 4), card_f(*ex, 5)}};
            const uint64_t nth = std::max<uint32_t>(ex_tag(*ex), 1), nph = std::max<uint32_t>(ex_seg(*ex), 1);
            if (nth * nph > 1) {
                if (nth * nph > 65536) {
                    fnec::set_error("receive sweep: more than 65536 incidences");
                    return FNEC_INVALID_ARG;
                }
                for (uint64_t it = 0; it < nth; ++it)
                    for (uint64_t ip = 0; ip < nph; ++ip) {
                        const double th = card_f(*ex, 4) + (double)it * card_f(*ex, 7);
                        const double ph = card_f(*ex, 5) + (double)ip * card_f(*ex, 8);
                        inc.push_back({th, ph});
                        d->recv_theta.push_back(th);
                        d->recv_phi.push_back(ph);
                    }
            }
            d->rhs_all.resize(n * inc.size());
            d->cosv.resize(n);
            d->sinv.resize(n);
            std::vector<Wave> waves;
            for (auto &a : inc)
                waves.push_back(plane_wave(*ex, a.first, a.second));
            planewave_rhs(waves, s, freq_hz, wires, paths ? &P : nullptr, d->rhs_all.data(), d->cosv.data(),
                          d->sinv.data());
            d->rhs.assign(d->rhs_all.begin(), d->rhs_all.begin() + n);
        } else {
            const Card *cs = nullptr;
            for (const Card &c : d->cards)
                if (!cs && card_is(c, "EX") && ex_current(c))
                    cs = &c;
            if (!paths && !P.ok && !detect_junctions(s, wires, 1e-6).empty()) {
                fnec::set_error("EX: current source is supported on straight or degree-2 junctioned wires; "
                                "degree-3+ (T/Y) junctions and closed loops are not yet supported");
                return FNEC_INVALID_ARG;
            }
            if (current_source_shape(d->cards, s, freq_hz, ex_tag(*cs), ex_seg(*cs), paths ? &P : nullptr, d->rhs,
                                     d->cosv, src)) {
                fnec::set_error("EX: current-source segment not found");
                return FNEC_INVALID_ARG;
            }
            i0 = fnec_c64{ex_vr(*cs), ex_vi(*cs)};
        }
    } else if (semantics == FNEC_SEM_CLI) {
        Paths P = conductor_paths(s);
        if (P.ok && P.count > 0 && P.nontrivial) {
            route = 1;
            if (hallen_rhs_paths(d->cards, s, freq_hz, P, d->rhs, d->cosv)) {
                fnec::set_error("EX: bad excitation");
                return FNEC_INVALID_ARG;
            }
            d->path_of = P.path_of;
            d->free_ends = P.free_ends;
        }
    }
    if (route == 0 && !variant) {
        if (hallen_rhs(d->cards, s, freq_hz, d->rhs, d->cosv)) {
            fnec::set_error("EX: bad excitation");
            return FNEC_INVALID_ARG;
        }
        auto wires = semantics == FNEC_SEM_CLI ? merge_collinear(s) : wire_endpoints(s);
        for (auto &w : wires) {
            d->wf.push_back(w.first);
            d->wl.push_back(w.second);
        }
        auto js = detect_junctions(s, wires, 1e-6);
        std::vector<size_t> comp(n, 0);
        for (size_t ci = 0; ci < wires.size(); ++ci)
            for (size_t m = wires[ci].first; m <= wires[ci].second; ++m)
                comp[m] = ci;
        for (auto &j : js) {
            if (semantics == FNEC_SEM_CLI && comp[j.a] == comp[j.b])
                continue;
            d->ja.push_back(j.a);
            d->jb.push_back(j.b);
            d->js.push_back(j.sign);
        }
    }
    // feedpoints (solve_session.rs:877-944; nec_worker solve.rs:243-268; fnec_py lib.rs:84-92)
    d->feed_idx.clear(); d->feed_v.clear(); d->feed_tag.clear(); d->feed_seg.clear(); d->feed_port.clear();
    for (const Card &c : d->cards) {
        if (!card_is(c, "EX") || ex_plane(c))
            continue;
        if (semantics == FNEC_SEM_WORKER && ex_type(c) != 0)
            continue;
        int idx = find_seg(s, ex_tag(c), ex_seg(c));
        if (idx < 0)
            continue;
        fnec_c64 v;
        if (semantics == FNEC_SEM_WORKER)
            v = fnec_c64{ex_vr(c), ex_vi(c)};
        else
            v = fnec_c64{vvec[idx].re * s[idx].length, vvec[idx].im * s[idx].length};
        d->feed_idx.push_back((size_t)idx);
        d->feed_v.push_back(v);
        d->feed_tag.push_back(s[idx].tag);
        d->feed_seg.push_back(s[idx].idx);
        d->feed_port.push_back(variant == 2 && ex_current(c) ? 1 : 0);
        if (semantics != FNEC_SEM_CLI)
            break;
    }
    // RP points: all RP cards, phi outer / theta inner (farfield.rs:677-699)
    d->theta.clear();
    d->phi.clear();
    if (semantics == FNEC_SEM_CLI)
        for (const Card &c : d->cards) {
            if (!card_is(c, "RP"))
                continue;
            int a = c.nf >= 8 ? 4 : 3;
            uint32_t nt = std::max<uint32_t>(as_u32(c.f[1]), 1), np = std::max<uint32_t>(as_u32(c.f[2]), 1);
            for (uint32_t ip = 0; ip < np; ++ip) {
                double ph = c.f[a + 1] + (double)ip * c.f[a + 3];
                for (uint32_t it = 0; it < nt; ++it) {
                    d->theta.push_back(c.f[a] + (double)it * c.f[a + 2]);
                    d->phi.push_back(ph);
                }
            }
        }
    fnec_request &r = out->req;
    r.segs = d->fsegs.data();
    r.n = n;
    r.rhs = d->rhs.data();
    r.cos_vec = d->cosv.data();
    r.wire_first = d->wf.empty() ? nullptr : d->wf.data();
    r.wire_last = d->wl.empty() ? nullptr : d->wl.data();
    r.n_wires = d->wf.size();
    r.j_a = d->ja.empty() ? nullptr : d->ja.data();
    r.j_b = d->jb.empty() ? nullptr : d->jb.data();
    r.j_sign = d->js.empty() ? nullptr : d->js.data();
    r.n_junc = d->ja.size();
    r.freq_hz = freq_hz;
    r.ground = fg;
    r.feed_index = d->feed_idx.empty() ? 0 : d->feed_idx[0];
    r.v_feed = d->feed_v.empty() ? fnec_c64{0.0, 0.0} : d->feed_v[0];
    r.rp_theta = d->theta.empty() ? nullptr : d->theta.data();
    r.rp_phi = d->phi.empty() ? nullptr : d->phi.data();
    r.rp_npts = d->theta.size();
    // LD and TL stamps for every caller, NT for the CLI (solve_session.rs:1285-1309;
    // nec_worker solve.rs:179-185; fnec_py lib.rs:58-63)
    build_stamps(d->cards, s, freq_hz, semantics == FNEC_SEM_CLI, d->stamps);
    r.stamps = d->stamps.empty() ? nullptr : d->stamps.data();
    r.n_stamps = d->stamps.size();
    out->route = route;
    out->path_of = d->path_of.empty() ? nullptr : d->path_of.data();
    out->free_ends = d->free_ends.empty() ? nullptr : d->free_ends.data();
    out->n_free = d->free_ends.size();
    out->n_feeds = d->feed_idx.size();
    out->feed_idx = d->feed_idx.empty() ? nullptr : d->feed_idx.data();
    out->feed_v = d->feed_v.empty() ? nullptr : d->feed_v.data();
    out->feed_tag = d->feed_tag.empty() ? nullptr : d->feed_tag.data();
    out->feed_seg = d->feed_seg.empty() ? nullptr : d->feed_seg.data();
    out->seg_tag = d->seg_tag.data();
    out->seg_index = d->seg_index.data();
    out->variant = variant;
    out->sin_vec = d->sinv.empty() ? nullptr : d->sinv.data();
    out->n_rhs = variant == 1 ? d->rhs_all.size() / n : 1;
    out->rhs_all = variant == 1 ? d->rhs_all.data() : r.rhs;
    out->n_recv = d->recv_theta.size();
    out->recv_theta = d->recv_theta.empty() ? nullptr : d->recv_theta.data();
    out->recv_phi = d->recv_phi.empty() ? nullptr : d->recv_phi.data();
    out->src = src >= 0 ? (size_t)src : 0;
    out->i0 = i0;
    out->feed_port = d->feed_port.empty() ? nullptr : d->feed_port.data();
    return FNEC_OK;
}

int fnec_deck_report(fnec_deck *d, size_t freq_index, int precision, char *outbuf, size_t cap,
                     size_t *needed)
{
    double freqs[4096];
    size_t nf = fnec_deck_frequencies(d, FNEC_SEM_CLI, freqs, 4096);
    if (freq_index >= nf || freq_index >= 4096) {
        fnec::set_error("no such frequency");
        return FNEC_INVALID_ARG;
    }
    const double f = freqs[freq_index];
    fnec_deck_problem pr;
    int st = fnec_deck_build(d, FNEC_SEM_CLI, f, &pr);
    if (st)
        return st;
    const fnec_request &r = pr.req;
    const size_t n = r.n;
    // route 1 writes one constant per conductor path (max path_of + 1)
    size_t ncons = r.n_wires;
    if (pr.route == 1) {
        ncons = 0;
        for (size_t i = 0; i < n; ++i)
            ncons = std::max(ncons, pr.path_of[i] + 1);
    }
    ncons = std::max<size_t>(ncons, 1);
    // a plane wave carries cos and sin constants per range or path, a current
    // source the port voltage after its constants
    const size_t s_len = n + (pr.variant == 1 ? 2 * ncons : ncons + (pr.variant == 2 ? 1 : 0));
    std::vector<fnec_c64> x(s_len * (pr.variant == 1 ? pr.n_rhs : 1));
    fnec_solve_stats stats;
    st = fnec_deck_solve(&pr, precision, x.data(), &stats);
    if (st)
        return st;
    const fnec_c64 port_v = pr.variant == 2 ? x[s_len - 1] : fnec_c64{0.0, 0.0};
    std::string o;
    o += "FNEC FEEDPOINT REPORT\nFORMAT_VERSION 1\n";
    o += "FREQ_MHZ " + fmt_fixed(f / 1e6, 6) + "\n";
    o += std::string("SOLVER_MODE ") +
         (pr.variant == 1 ? "hallen-planewave" : pr.variant == 2 ? "hallen-current-source" : "hallen") +
         "\nPULSE_RHS Nec2\n\nFEEDPOINTS\nTAG SEG V_RE V_IM I_RE I_IM Z_RE Z_IM\n";
    double p_in = 0.0;
    for (size_t i = 0; i < pr.n_feeds; ++i) {
        fnec_c64 v = pr.feed_port && pr.feed_port[i] ? port_v : pr.feed_v[i], c = x[pr.feed_idx[i]], z;
        if (std::hypot(c.re, c.im) > 1e-60) {
            double ns = c.re * c.re + c.im * c.im;
            z.re = (v.re * c.re + v.im * c.im) / ns;
            z.im = (v.im * c.re - v.re * c.im) / ns;
        } else
            z = v;
        p_in += 0.5 * (v.re * c.re - v.im * (-c.im));
        o += std::to_string(pr.feed_tag[i]) + " " + std::to_string(pr.feed_seg[i]) + " " +
             fmt_fixed(v.re, 6) + " " + fmt_fixed(v.im, 6) + " " + fmt_fixed(c.re, 6) + " " +
             fmt_fixed(c.im, 6) + " " + fmt_fixed(z.re, 6) + " " + fmt_fixed(z.im, 6) + "\n";
    }
    int nsrc = 0;
    for (const Card &c : d->cards)
        if (card_is(c, "EX"))
            ++nsrc;
    if (nsrc) {
        o += "\nSOURCES\nN_SOURCES " + std::to_string(nsrc) + "\nTYPE TAG SEG I4 V_RE V_IM\n";
        for (const Card &c : d->cards)
            if (card_is(c, "EX"))
                o += std::to_string(as_u32(c.f[0])) + " " + std::to_string(ex_tag(c)) + " " +
                     std::to_string(ex_seg(c)) + " " + std::to_string(as_u32(c.f[3])) + " " +
                     fmt_fixed(ex_vr(c), 6) + " " + fmt_fixed(ex_vi(c), 6) + "\n";
    }
    // LD cards as parsed (solve_session.rs:966-982, nec_report lib.rs:335-344)
    int nld = 0;
    for (const Card &c : d->cards)
        if (card_is(c, "LD"))
            ++nld;
    if (nld) {
        o += "\nLOADS\nN_LOADS " + std::to_string(nld) + "\nTYPE TAG SEG_FIRST SEG_LAST F1 F2 F3\n";
        for (const Card &c : d->cards)
            if (card_is(c, "LD"))
                o += std::to_string(as_i32(c.f[0])) + " " + std::to_string(as_u32(c.f[1])) + " " +
                     std::to_string(as_u32(c.f[2])) + " " + std::to_string(as_u32(c.f[3])) + " " +
                     fmt_fixed(c.nf > 4 ? c.f[4] : 0.0, 6) + " " + fmt_fixed(c.nf > 5 ? c.f[5] : 0.0, 6) +
                     " " + fmt_fixed(c.nf > 6 ? c.f[6] : 0.0, 6) + "\n";
    }
    o += "\nCURRENTS\nTAG SEG I_RE I_IM I_MAG I_PHASE\n";
    for (size_t i = 0; i < n; ++i) {
        fnec_c64 c = x[i];
        o += std::to_string(pr.seg_tag[i]) + " " + std::to_string(pr.seg_index[i]) + " " +
             fmt_exp(c.re, 6) + " " + fmt_exp(c.im, 6) + " " + fmt_exp(std::hypot(c.re, c.im), 6) +
             " " + fmt_fixed(rad2deg(std::atan2(c.im, c.re)), 4) + "\n";
    }
    // receive pattern (solve_session.rs:313-371): each sweep incidence's peak
    // |I| over the segments, in dB below the largest (nec_report lib.rs)
    std::string recv;
    if (pr.variant == 1 && pr.n_recv) {
        std::vector<double> peak(pr.n_recv, 0.0);
        double max_peak = 0.0;
        for (size_t q = 0; q < pr.n_recv; ++q) {
            const fnec_c64 *xq = x.data() + (q + 1) * s_len;
            for (size_t m = 0; m < n; ++m)
                peak[q] = std::fmax(peak[q], std::hypot(xq[m].re, xq[m].im));
            max_peak = std::fmax(max_peak, peak[q]);
        }
        recv = "\nRECEIVE_PATTERN\nN_POINTS " + std::to_string(pr.n_recv) + "\nTHETA PHI RESPONSE_DB\n";
        for (size_t q = 0; q < pr.n_recv; ++q) {
            const double db = peak[q] > 0.0 && max_peak > 0.0 ? 20.0 * std::log10(peak[q] / max_peak) : -999.99;
            recv += fmt_fixed(pr.recv_theta[q], 4) + " " + fmt_fixed(pr.recv_phi[q], 4) + " " + fmt_fixed(db, 4) +
                    "\n";
        }
    }
    const size_t np = r.rp_npts;
    bool normalize = false, avg = false;
    for (const Card &c : d->cards)
        if (card_is(c, "RP") && c.nf >= 8) {
            uint32_t xn = as_u32(c.f[3]);
            if ((xn / 1000) % 10 != 0)
                normalize = true;
            if (xn % 10 != 0)
                avg = true;
        }
    std::vector<fnec_rp_row> rows(np);
    if (np) {
        double total = 0.0;
        st = fnec_hip_rp(r.segs, x.data(), n, f, &r.ground, r.rp_theta, r.rp_phi, np, rows.data(), &total);
        if (st)
            return st;
        if (r.ground.kind == FNEC_GROUND_FINITE && p_in > 0.0) {
            // radiation efficiency (farfield.rs:308-322, solve_session.rs:1664-1684):
            // the integral at k = 2 pi f / c0
            double integral = 0.0;
            st = fnec_hip_rp(r.segs, x.data(), n, f, &r.ground, nullptr, nullptr, 0, nullptr, &integral);
            if (st)
                return st;
            const double k = 2.0 * kPi * f / kC0;
            double p_rad = k * k * kEta0 / (32.0 * kPi * kPi) * integral;
            double eta = p_rad / p_in;
            eta = eta < 0.0 ? 0.0 : (eta > 1.0 ? 1.0 : eta);
            double ddb = 10.0 * std::log10(eta);
            for (auto &row : rows)
                for (double *g : {&row.gain_total_dbi, &row.gain_theta_dbi, &row.gain_phi_dbi})
                    if (*g > -900.0)
                        *g += ddb;
        }
        o += "\nRADIATION_PATTERN\nN_POINTS " + std::to_string(np) +
             "\nTHETA PHI GAIN_DB GAIN_V_DB GAIN_H_DB AXIAL_RATIO\n";
        for (size_t i = 0; i < np; ++i)
            o += fmt_fixed(r.rp_theta[i], 4) + " " + fmt_fixed(r.rp_phi[i], 4) + " " +
                 fmt_fixed(rows[i].gain_total_dbi, 4) + " " + fmt_fixed(rows[i].gain_theta_dbi, 4) +
                 " " + fmt_fixed(rows[i].gain_phi_dbi, 4) + " " + fmt_fixed(rows[i].axial_ratio, 4) + "\n";
        o += recv;
        if (normalize) {
            double peak = -INFINITY;
            for (auto &row : rows)
                peak = std::fmax(peak, row.gain_total_dbi);
            o += "\nNORMALIZED_PATTERN\nN_POINTS " + std::to_string(np) + "\nTHETA PHI GAIN_NORM_DB\n";
            for (size_t i = 0; i < np; ++i) {
                double g = rows[i].gain_total_dbi;
                o += fmt_fixed(r.rp_theta[i], 4) + " " + fmt_fixed(r.rp_phi[i], 4) + " " +
                     fmt_fixed(g > -900.0 ? g - peak : g, 4) + "\n";
            }
        }
        if (avg) {
            double num = 0.0, den = 0.0;
            for (size_t i = 0; i < np; ++i) {
                double g = rows[i].gain_total_dbi;
                if (g <= -900.0)
                    continue;
                double w = std::sin(deg2rad(r.rp_theta[i]));
                num += std::pow(10.0, g / 10.0) * w;
                den += w;
            }
            if (den > 0.0)
                o += "\nAVERAGE_POWER_GAIN " + fmt_fixed(num / den, 4) + "\n";
        }
    } else
        o += recv;
    if (needed)
        *needed = o.size() + 1;
    if (outbuf && cap) {
        size_t k = std::min(cap - 1, o.size());
        std::memcpy(outbuf, o.data(), k);
        outbuf[k] = 0;
    }
    return FNEC_OK;
}

} // extern "C"
