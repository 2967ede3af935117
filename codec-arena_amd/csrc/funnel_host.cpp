// funnel_host.cpp -- host half of libcsfunnel: genome/model parsing, the
// reference heap, the Zig-compatible output writers.
//
// Reference lines restated here (codec-arena snapshot):
//   parseLine            src/genome.zig:116-188
//   Model.fromCompact    src/screen.zig:36-95
//   loadModel            src/screen.zig:115-127
//   elites loading       src/main.zig:201-214
//   Heap.offer           src/screen.zig:166-192
//   output lines/count   src/screen.zig:322-341
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "funnel_internal.h"

namespace csf {

static thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}
void clear_error() { g_err.clear(); }

// ------------------------------------------------------------------ parsing

// std.fmt.parseInt(uN, tok, 10) of Zig std 0.17: optional '+'/'-', '_'
// separators (not first/last), digits only, no overflow; '-' admits only 0.
static bool parse_uint(const char *p, size_t n, uint64_t limit, uint64_t &v) {
    if (n == 0) return false;
    bool minus = false;
    if (*p == '+' || *p == '-') {
        minus = *p == '-';
        ++p;
        --n;
    }
    if (n == 0 || p[0] == '_' || p[n - 1] == '_') return false;
    unsigned __int128 acc = 0;
    for (size_t i = 0; i < n; ++i) {
        if (p[i] == '_') continue;
        unsigned d = (unsigned)(p[i] - '0');
        if (d > 9) return false;
        acc = acc * 10 + d;
        if (acc > limit) return false;
    }
    if (minus && acc != 0) return false;
    v = (uint64_t)acc;
    return true;
}

// std.fmt.parseFloat(f32, tok): correctly rounded decimal (or hex) parse.
static bool parse_float(const char *p, size_t n, float &v) {
    if (n == 0 || n > 120 || p[0] == '_' || p[n - 1] == '_') return false;
    char buf[128];
    size_t k = 0;
    for (size_t i = 0; i < n; ++i) {
        char c = p[i];
        if (c == '_') continue;
        if (c == '(' || c == ')' || c == ' ') return false;
        buf[k++] = c;
    }
    buf[k] = '\0';
    char *end = nullptr;
    v = std::strtof(buf, &end);
    return end == buf + k;
}

struct Tok {
    const char *p;
    size_t n;
};

static bool sep(char c) { return c == ' ' || c == '\t' || c == '\r'; }

static int parse_line(const char *line, size_t len, cs_genome &g) {
    Tok t[34];
    int n = 0;
    for (size_t i = 0; i < len;) {
        if (sep(line[i])) {
            ++i;
            continue;
        }
        size_t b = i;
        while (i < len && !sep(line[i])) ++i;
        if (n == 34) return CS_E_BADLINE;  // genome.zig:121
        t[n++] = Tok{line + b, i - b};
    }
    switch (n) {
    case 21: case 27: case 29: case 32: case 33: case 34: break;
    default: return 0;  // genome.zig:125
    }
    if (t[0].n > 0 && t[0].p[0] == '#') return 0;  // genome.zig:126
    std::memset(&g, 0, sizeof g);
    uint64_t u;
    if (!parse_uint(t[0].p, t[0].n, UINT64_MAX, u)) return CS_E_PARSE;
    g.id = u;
    if (!parse_uint(t[1].p, t[1].n, UINT32_MAX, u)) return CS_E_PARSE;
    g.gen = (uint32_t)u;
    if (!parse_uint(t[2].p, t[2].n, UINT64_MAX, u)) return CS_E_PARSE;
    g.pa = u;
    if (!parse_uint(t[3].p, t[3].n, UINT64_MAX, u)) return CS_E_PARSE;
    g.pb = u;
    if (!parse_uint(t[4].p, t[4].n, UINT32_MAX, u)) return CS_E_PARSE;
    g.tribe = (uint32_t)u;
    // Token slot of each flat gene 0..22 for this line width; -1 = migrated 0.
    // genome.zig:135-163 -- note 34-token lines never fill genes 16..21.
    for (int k = 0; k < CS_NG; ++k) {
        int slot = -1;
        if (k < 16) slot = 5 + k;
        else if (k <= 17) slot = (n == 29 || n == 32 || n == 33) ? 5 + k : -1;
        else if (k <= 20) slot = (n == 32 || n == 33) ? 5 + k : -1;
        else if (k == 21) slot = (n == 33) ? 26 : -1;
        else slot = (n == 34) ? 27 : -1;
        if (slot < 0) {
            g.g[k] = 0.0f;
        } else if (!parse_float(t[slot].p, t[slot].n, g.g[k])) {
            return CS_E_PARSE;
        }
    }
    int rfirst = n == 27 ? 21 : n == 29 ? 23 : n == 32 ? 26 : n == 33 ? 27 : n == 34 ? 28 : -1;
    for (int k = 0; k < CS_NR; ++k) {
        if (rfirst < 0) g.reg[k] = 0.5f;  // genome.zig:185
        else if (!parse_float(t[rfirst + k].p, t[rfirst + k].n, g.reg[k])) return CS_E_PARSE;
    }
    return 1;
}

static bool model_len_ok(uint64_t n) {
    return n == 153 || n == 159 || n == 161 || n == 164 || n == 165 || n == 166;
}

static int from_compact(const float *w, int n, cs_model &m) {
    if (!model_len_ok((uint64_t)n)) return fail(CS_E_BADMODEL, "BadModel: unsupported length " + std::to_string(n));
    std::memset(&m, 0, sizeof m);
    m.w0 = w[0];
    for (int i = 0; i < 16; ++i) m.lin[i] = w[1 + i];
    // (first flat linear slot, count) copied from w[17..], then regulators.
    int extra = 0, regs_at = -1, quad_at = 17;
    switch (n) {
    case 153: extra = 0; quad_at = 17; break;
    case 159: extra = 6; quad_at = 23; break;           // R4: regs land on 16..21
    case 161: extra = 2; regs_at = 19; quad_at = 25; break;
    case 164: extra = 5; regs_at = 22; quad_at = 28; break;
    case 165: extra = 6; regs_at = 23; quad_at = 29; break;
    case 166: extra = 7; regs_at = 24; quad_at = 30; break;
    }
    for (int i = 0; i < extra; ++i) m.lin[16 + i] = w[17 + i];
    if (regs_at >= 0)
        for (int i = 0; i < CS_NR; ++i) m.lin[CS_NG + i] = w[regs_at + i];
    int idx = quad_at;
    for (int i = 0; i < 16; ++i)
        for (int j = 0; j <= i; ++j) m.quad[i * 16 + j] = w[idx++];
    return CS_OK;
}

static int slurp(const char *path, size_t limit, std::string &out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return fail(CS_E_IO, std::string("cannot open ") + path);
    out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    if (out.size() > limit) return fail(CS_E_IO, std::string("file too large: ") + path);
    return CS_OK;
}

// ------------------------------------------------------------------- heap

// src/screen.zig:166-192, verbatim semantics: fill phase sifts up while the
// parent is strictly greater; full phase rejects score <= root, otherwise
// replaces the root and sifts down toward the strictly smaller child (left
// first).
void StreamHeap::offer(const cs_survivor &s, double key) {
    if (slots.size() < k) {
        slots.push_back(HeapItem{key, s});
        size_t i = slots.size() - 1;
        while (i > 0) {
            size_t up = (i - 1) >> 1;
            if (slots[up].key <= slots[i].key) break;
            std::swap(slots[up], slots[i]);
            i = up;
        }
        return;
    }
    if (key <= slots[0].key) return;
    slots[0] = HeapItem{key, s};
    size_t i = 0;
    const size_t n = slots.size();
    for (;;) {
        size_t l = 2 * i + 1, r = l + 1, m = i;
        if (l < n && slots[l].key < slots[m].key) m = l;
        if (r < n && slots[r].key < slots[m].key) m = r;
        if (m == i) break;
        std::swap(slots[m], slots[i]);
        i = m;
    }
}

// --------------------------------------------------------------- printing

// Minimal unsigned big integer for exact shortest-digit generation.
struct Big {
    std::vector<uint32_t> d;  // little-endian base 2^32
    explicit Big(uint64_t v = 0) {
        while (v) {
            d.push_back((uint32_t)v);
            v >>= 32;
        }
    }
    void trim() {
        while (!d.empty() && d.back() == 0) d.pop_back();
    }
    void mul_small(uint32_t m) {
        uint64_t carry = 0;
        for (auto &x : d) {
            uint64_t t = (uint64_t)x * m + carry;
            x = (uint32_t)t;
            carry = t >> 32;
        }
        if (carry) d.push_back((uint32_t)carry);
    }
    void shl(unsigned b) {
        if (d.empty()) return;
        unsigned w = b / 32, r = b % 32;
        if (r) {
            uint32_t carry = 0;
            for (auto &x : d) {
                uint32_t nx = (x << r) | carry;
                carry = x >> (32 - r);
                x = nx;
            }
            if (carry) d.push_back(carry);
        }
        d.insert(d.begin(), w, 0u);
    }
    static int cmp(const Big &a, const Big &b) {
        if (a.d.size() != b.d.size()) return a.d.size() < b.d.size() ? -1 : 1;
        for (size_t i = a.d.size(); i-- > 0;)
            if (a.d[i] != b.d[i]) return a.d[i] < b.d[i] ? -1 : 1;
        return 0;
    }
    void add(const Big &b) {
        uint64_t carry = 0;
        if (d.size() < b.d.size()) d.resize(b.d.size(), 0);
        for (size_t i = 0; i < d.size(); ++i) {
            uint64_t t = (uint64_t)d[i] + (i < b.d.size() ? b.d[i] : 0) + carry;
            d[i] = (uint32_t)t;
            carry = t >> 32;
        }
        if (carry) d.push_back((uint32_t)carry);
    }
    void sub(const Big &b) {  // requires *this >= b
        int64_t borrow = 0;
        for (size_t i = 0; i < d.size(); ++i) {
            int64_t t = (int64_t)d[i] - (i < b.d.size() ? b.d[i] : 0) - borrow;
            borrow = t < 0;
            d[i] = (uint32_t)(t + (borrow ? ((int64_t)1 << 32) : 0));
        }
        trim();
    }
};

// Shortest digit string that reads back as the f32 `v` (> 0), by the
// free-format algorithm of Burger & Dybvig (PLDI 1996) with IEEE
// round-half-even bounds (inclusive when the significand is even).  Returns
// the digits and the decimal exponent K such that v ~= 0.d1d2... * 10^K.
static void shortest_digits(float v, std::string &digits, int &K) {
    uint32_t bits;
    std::memcpy(&bits, &v, 4);
    uint32_t frac = bits & 0x7fffff, bexp = (bits >> 23) & 0xff;
    uint64_t f;
    int e;
    if (bexp == 0) {
        f = frac;
        e = -149;
    } else {
        f = frac | 0x800000;
        e = (int)bexp - 150;
    }
    const bool even = (f & 1) == 0;
    const bool lower_closer = (frac == 0 && bexp > 1);
    Big r, s, mp, mm;
    if (e >= 0) {
        r = Big(f);
        r.shl((unsigned)e + (lower_closer ? 2 : 1));
        s = Big(lower_closer ? 4 : 2);
        mp = Big(1);
        mp.shl((unsigned)e + (lower_closer ? 1 : 0));
        mm = Big(1);
        mm.shl((unsigned)e);
    } else {
        r = Big(f);
        r.shl(lower_closer ? 2 : 1);
        s = Big(1);
        s.shl((unsigned)(-e) + (lower_closer ? 2 : 1));
        mp = Big(lower_closer ? 2 : 1);
        mm = Big(1);
    }
    // Scale so that v/10^K is in [0.1, 1): estimate K then fix up.
    int k = (int)std::ceil(std::log10((double)v) - 1e-10);
    if (k >= 0) {
        for (int i = 0; i < k; ++i) s.mul_small(10);
    } else {
        for (int i = 0; i < -k; ++i) {
            r.mul_small(10);
            mp.mul_small(10);
            mm.mul_small(10);
        }
    }
    auto high_ok = [&](const Big &rr) {  // rr + mp vs s
        Big t = rr;
        t.add(mp);
        int c = Big::cmp(t, s);
        return even ? c >= 0 : c > 0;
    };
    while (high_ok(r)) {  // estimate too low
        s.mul_small(10);
        ++k;
    }
    for (;;) {  // estimate too high?
        Big t = r, mpx = mp;
        t.mul_small(10);
        mpx.mul_small(10);
        t.add(mpx);
        int c = Big::cmp(t, s);
        bool reach = even ? c >= 0 : c > 0;
        if (reach) break;
        r.mul_small(10);
        mp.mul_small(10);
        mm.mul_small(10);
        --k;
    }
    K = k;
    digits.clear();
    for (;;) {
        r.mul_small(10);
        mp.mul_small(10);
        mm.mul_small(10);
        int dgt = 0;
        while (Big::cmp(r, s) >= 0) {
            r.sub(s);
            ++dgt;
        }
        int cl = Big::cmp(r, mm);
        bool tc1 = even ? cl <= 0 : cl < 0;
        bool tc2 = high_ok(r);
        if (!tc1 && !tc2) {
            digits.push_back((char)('0' + dgt));
            continue;
        }
        if (tc1 && tc2) {
            Big r2 = r;
            r2.shl(1);
            int c = Big::cmp(r2, s);
            if (c > 0 || (c == 0 && (dgt & 1))) ++dgt;
        } else if (tc2) {
            ++dgt;
        }
        digits.push_back((char)('0' + dgt));
        break;
    }
    // A terminal carry (digit 10) propagates; it cannot happen for shortest
    // output except as "10" -> fix generically.
    for (size_t i = digits.size(); i-- > 0;) {
        if (digits[i] <= '9') break;
        digits[i] = '0';
        if (i == 0) {
            digits.insert(digits.begin(), '1');
            ++K;
        } else {
            ++digits[i - 1];
        }
    }
    while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
}

// Zig std 0.17 float formatting in decimal mode with precision 4:
// format_float.formatDecimal(round(shortest, .decimal, 4)).  The shortest
// digit string is cut after `round_digit` significant digits (those up to the
// 4th decimal place, or none when the value is below 1e-4) and rounded half-up
// on the first removed digit only; the result is printed with the decimal
// point and zero padding to 4 places.
void format_d4(float v, std::string &out) {
    if (std::isnan(v)) {
        out += "nan";
        return;
    }
    if (std::isinf(v)) {
        out += v < 0 ? "-inf" : "inf";
        return;
    }
    if (std::signbit(v)) out += '-';
    float a = std::fabs(v);
    std::string dg;
    int exp10;  // value = int(dg) * 10^exp10
    if (a == 0.0f) {
        dg = "0";
        exp10 = 0;
    } else {
        int K;
        shortest_digits(a, dg, K);
        exp10 = K - (int)dg.size();
    }
    const int prec = 4;
    int ol = (int)dg.size();
    int keep;  // round_digit of format_float.round
    if (exp10 > 0) keep = (ol - 1) + prec + exp10;
    else keep = (prec + ol > -exp10) ? prec + ol + exp10 : 0;
    if (keep < ol) {
        char first_removed = dg[keep];
        std::string kept = dg.substr(0, keep);
        exp10 += ol - keep;
        if (first_removed >= '5') {
            // increment the kept digit string (empty string counts as 0)
            int i = (int)kept.size() - 1;
            while (i >= 0 && kept[i] == '9') kept[i--] = '0';
            if (i >= 0) ++kept[i];
            else kept.insert(kept.begin(), '1');
        }
        if (kept.empty()) kept = "0";
        // drop leading zeros produced by an all-zero kept string
        size_t nz = kept.find_first_not_of('0');
        kept = nz == std::string::npos ? "0" : kept.substr(nz);
        dg = kept;
    }
    ol = (int)dg.size();
    int point = exp10 + ol;  // digits before the decimal point
    if (point <= 0) {
        out += "0.";
        out.append((size_t)(-point), '0');
        out += dg;
        int written = -point + ol;
        if (written < prec) out.append((size_t)(prec - written), '0');
    } else if (point >= ol) {
        out += dg;
        out.append((size_t)(point - ol), '0');
        out += '.';
        out.append((size_t)prec, '0');
    } else {
        out += dg.substr(0, (size_t)point);
        out += '.';
        out += dg.substr((size_t)point);
        int written = ol - point;
        if (written < prec) out.append((size_t)(prec - written), '0');
    }
}

}  // namespace csf

using namespace csf;

extern "C" {

int cs_version(void) { return 1; }

const char *cs_last_error(void) { return g_err.c_str(); }

int cs_parse_u64(const char *tok, size_t len, uint64_t *out) {
    if (!tok || !out) return fail(CS_E_BADARGS, "null argument");
    uint64_t v;
    if (!parse_uint(tok, len, UINT64_MAX, v)) return fail(CS_E_PARSE, "InvalidCharacter or Overflow");
    *out = v;
    return CS_OK;
}

int cs_parse_f32(const char *tok, size_t len, float *out) {
    if (!tok || !out) return fail(CS_E_BADARGS, "null argument");
    float v;
    if (!parse_float(tok, len, v)) return fail(CS_E_PARSE, "InvalidCharacter");
    *out = v;
    return CS_OK;
}

int cs_parse_pop_line(const char *line, size_t len, cs_genome *out) {
    if (!line || !out) return fail(CS_E_BADARGS, "null argument");
    cs_genome g;
    int rc = parse_line(line, len, g);
    if (rc == 1) *out = g;
    if (rc == CS_E_BADLINE) fail(rc, "BadLine: more than 34 tokens");
    if (rc == CS_E_PARSE) fail(rc, "parse error in pop line");
    return rc;
}

int cs_model_from_compact(const float *w, int n, cs_model *out) {
    if (!w || !out) return fail(CS_E_BADARGS, "null argument");
    return from_compact(w, n, *out);
}

int cs_load_model(const char *path, cs_model *out) {
    std::string txt;
    int rc = slurp(path, 1u << 20, txt);
    if (rc) return rc;
    std::vector<Tok> toks;
    for (size_t i = 0; i < txt.size() && toks.size() < 200;) {
        char c = txt[i];
        if (c == ' ' || c == '\t' || c == '\r' || c == '\n') {
            ++i;
            continue;
        }
        size_t b = i;
        while (i < txt.size() && !(txt[i] == ' ' || txt[i] == '\t' || txt[i] == '\r' || txt[i] == '\n')) ++i;
        toks.push_back(Tok{txt.data() + b, i - b});
    }
    if (toks.empty()) return fail(CS_E_BADMODEL, "BadModel: empty model file");
    uint64_t n;
    if (!parse_uint(toks[0].p, toks[0].n, UINT64_MAX, n)) return fail(CS_E_PARSE, "model length is not an integer");
    if (!model_len_ok(n)) return fail(CS_E_BADMODEL, "BadModel: unsupported length " + std::to_string(n));
    float w[166];
    for (uint64_t i = 0; i < n; ++i) {
        if (1 + i >= toks.size()) return fail(CS_E_BADMODEL, "BadModel: too few coefficients");
        if (!parse_float(toks[1 + i].p, toks[1 + i].n, w[i])) return fail(CS_E_PARSE, "model coefficient parse error");
    }
    return from_compact(w, (int)n, *out);
}

int cs_load_elites(const char *path, cs_genome *out, int cap, int *n_out) {
    std::string txt;
    int rc = slurp(path, 1u << 22, txt);
    if (rc) return rc;
    int n = 0;
    size_t i = 0;
    while (i < txt.size()) {
        size_t b = i, e = txt.find('\n', i);
        if (e == std::string::npos) e = txt.size();
        i = e + 1;
        while (b < e && sep(txt[b])) ++b;
        while (e > b && sep(txt[e - 1])) --e;
        if (b == e || txt[b] == '#') continue;
        cs_genome g;
        if (parse_line(txt.data() + b, e - b, g) == 1) {
            if (out && n < cap) out[n] = g;
            ++n;
        }
    }
    *n_out = n;
    if (n > cap) return fail(CS_E_CAP, "elites buffer too small");
    return CS_OK;
}

int cs_format_gene(float v, char *buf, size_t cap) {
    std::string s;
    format_d4(v, s);
    if (buf && cap) {
        size_t n = s.size() < cap - 1 ? s.size() : cap - 1;
        std::memcpy(buf, s.data(), n);
        buf[n] = '\0';
    }
    return (int)s.size();
}

}  // extern "C"

// "0 0 pa pb tribe" and the first `ngenes` genes at {d:.4}, one line per survivor.
static int write_survivor_lines(const char *path, const cs_survivor *s, uint64_t n, int ngenes) {
    std::string txt;
    txt.reserve(n * 220);
    for (uint64_t i = 0; i < n; ++i) {
        txt += "0 0 ";
        txt += std::to_string(s[i].pa);
        txt += ' ';
        txt += std::to_string(s[i].pb);
        txt += ' ';
        txt += std::to_string(s[i].tribe);
        for (int k = 0; k < ngenes; ++k) {
            txt += ' ';
            format_d4(s[i].g[k], txt);
        }
        txt += '\n';
    }
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(CS_E_IO, std::string("cannot create ") + path);
    size_t w = std::fwrite(txt.data(), 1, txt.size(), f);
    std::fclose(f);
    return w == txt.size() ? CS_OK : fail(CS_E_IO, "short write");
}

extern "C" {

// src/screen.zig:326-337 writes 24 genes (29 tokens per line)
int cs_write_promoted(const char *path, const cs_survivor *s, uint64_t n) {
    return write_survivor_lines(path, s, n, 24);
}

int cs_write_promoted_v6(const char *path, const cs_survivor *s, uint64_t n) {
    return write_survivor_lines(path, s, n, CS_NTOT);
}

int cs_write_count(const char *path, uint64_t total) {
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(CS_E_IO, std::string("cannot create ") + path);
    std::fprintf(f, "%llu\n", (unsigned long long)total);
    std::fclose(f);
    return CS_OK;
}

}  // extern "C"
