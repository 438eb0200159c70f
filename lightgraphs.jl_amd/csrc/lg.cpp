// lg.cpp -- the edge lines of the LG text format, parsed on the host.
//
// Restates the line loop of _lg_read_one_graph (src/persistence/lg.jl:35-52):
// a graph's header line is followed by ne lines; each is chomped (one
// trailing "\n" or "\r\n" removed); an empty line is skipped but still
// counts as one of the ne; any other line is split at its commas (blanks
// around a comma belong to the separator, r"\s*,\s*") and its first two
// fields are parsed as integers -- Julia's two-name destructuring ignores
// any further field, and a line with one field fails.  The integer syntax
// is that of Julia's parse(Int64, s): blanks around the number, an optional
// sign, decimal digits or a 0x / 0o / 0b prefix with digits of that base.
// The caller checks the values against the header's vertex count and
// element type (add_edge! drops out-of-range endpoints).
#include "../../include/lgmi_gen.h"

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

inline bool is_blank(char c) {
    return c == ' ' || c == '\t' || c == '\r' || c == '\n' || c == '\v' || c == '\f';
}

// Parses text[b, e) as an Int64; 0, -2 (not an integer) or -3 (overflow).
int parse_i64(const char* t, int64_t b, int64_t e, int64_t* out) {
    while (b < e && is_blank(t[b])) b++;
    while (e > b && is_blank(t[e - 1])) e--;
    if (b >= e) return -2;
    bool neg = false;
    if (t[b] == '+' || t[b] == '-') {
        neg = t[b] == '-';
        b++;
    }
    int base = 10;
    if (e - b > 2 && t[b] == '0' && (t[b + 1] == 'x' || t[b + 1] == 'o' || t[b + 1] == 'b')) {
        base = t[b + 1] == 'x' ? 16 : (t[b + 1] == 'o' ? 8 : 2);
        b += 2;
    }
    if (b >= e) return -2;
    // magnitude up to 2^63 (the negative limit)
    const uint64_t lim = neg ? (uint64_t)1 << 63 : ((uint64_t)1 << 63) - 1;
    uint64_t x = 0;
    for (int64_t i = b; i < e; i++) {
        const char c = t[i];
        int d;
        if (c >= '0' && c <= '9') d = c - '0';
        else if (c >= 'a' && c <= 'f') d = c - 'a' + 10;
        else if (c >= 'A' && c <= 'F') d = c - 'A' + 10;
        else return -2;
        if (d >= base) return -2;
        if (x > (lim - (uint64_t)d) / (uint64_t)base) return -3;
        x = x * (uint64_t)base + (uint64_t)d;
    }
    *out = neg ? (int64_t)(0 - x) : (int64_t)x;
    return 0;
}

// Parses lines [first_line, first_line + nl) starting at text[pos] (the
// lines are known to end before `len`); pairs go to src/dst from index 0.
// Returns 0 or an error code with *bad = the 1-based line number.
int parse_lines(const char* text, int64_t len, int64_t pos, int64_t first_line, int64_t nl, int64_t* src,
                int64_t* dst, int64_t* np_out, int64_t* bad) {
    int64_t np = 0;
    for (int64_t line = first_line; line < first_line + nl; line++) {
        if (pos >= len) break;   // readline at the end of the stream gives ""
        const char* nlp = static_cast<const char*>(std::memchr(text + pos, '\n', (size_t)(len - pos)));
        const int64_t e = nlp ? (int64_t)(nlp - text) : len;
        const int64_t next = e < len ? e + 1 : e;
        int64_t le = e;
        if (e < len && le > pos && text[le - 1] == '\r') le--;   // chomp removes "\r\n" whole
        if (le > pos) {
            // the first comma ends field one, the next comma (if any) field two
            int64_t c1 = pos;
            while (c1 < le && text[c1] != ',') c1++;
            if (c1 >= le) {
                *bad = line + 1;
                *np_out = np;
                return -1;
            }
            int64_t c2 = c1 + 1;
            while (c2 < le && text[c2] != ',') c2++;
            int64_t sv = 0, dv = 0;
            int rc = parse_i64(text, pos, c1, &sv);
            if (rc == 0) rc = parse_i64(text, c1 + 1, c2, &dv);
            if (rc != 0) {
                *bad = line + 1;
                *np_out = np;
                return rc;
            }
            src[np] = sv;
            dst[np] = dv;
            np++;
        }
        pos = next;
    }
    *np_out = np;
    return 0;
}

}  // namespace

// Small inputs are parsed by the calling thread.  Larger ones are cut by
// bytes into up to 16 ranges that start at line starts; each thread counts
// the line breaks of its range, a prefix sum gives each range's first line,
// and each range parses the lines it holds among the first nlines into its
// own part of src/dst.  The parts are then moved together in line order; an
// error reports the first bad line.
extern "C" int lgmi_lg_parse_edges(const char* text, int64_t len, int64_t nlines, int64_t* src, int64_t* dst,
                                   int64_t* npairs, int64_t* consumed, int64_t* bad_line) {
    if (!npairs || !consumed || !bad_line || len < 0 || nlines < 0 || (len > 0 && !text) ||
        (nlines > 0 && (!src || !dst)))
        return -4;
    *npairs = 0;
    *consumed = 0;
    *bad_line = 0;
    const int64_t kMinBytes = (int64_t)1 << 20;   // bytes per thread at least
    const unsigned hw = std::max(1u, std::min(std::thread::hardware_concurrency(), 16u));
    const int nt = (int)std::min<int64_t>((int64_t)hw, std::max<int64_t>(1, len / kMinBytes));
    if (nt == 1 || nlines < 4096) {
        int64_t np = 0, bad = 0;
        const int rc = parse_lines(text, len, 0, 0, nlines, src, dst, &np, &bad);
        *npairs = np;
        *bad_line = bad;
        if (rc != 0) return rc;
        // where the nlines lines end
        int64_t pos = 0;
        for (int64_t line = 0; line < nlines && pos < len; line++) {
            const char* nlp = static_cast<const char*>(std::memchr(text + pos, '\n', (size_t)(len - pos)));
            pos = nlp ? (int64_t)(nlp - text) + 1 : len;
        }
        *consumed = pos;
        return 0;
    }
    // range starts: the first line start at or after r * len / nt
    std::vector<int64_t> rs((size_t)nt + 1, len);
    rs[0] = 0;
    for (int r = 1; r < nt; r++) {
        int64_t p = std::max(rs[(size_t)r - 1], (int64_t)((double)len * r / nt));
        if (p > 0 && p < len && text[p - 1] != '\n') {
            const char* nlp = static_cast<const char*>(std::memchr(text + p, '\n', (size_t)(len - p)));
            p = nlp ? (int64_t)(nlp - text) + 1 : len;
        }
        rs[(size_t)r] = p;
    }
    // line breaks per range; a range's last line may lack one only at the end
    std::vector<int64_t> cnt((size_t)nt, 0);
    {
        std::vector<std::thread> th;
        for (int r = 0; r < nt; r++)
            th.emplace_back([&, r] {
                cnt[(size_t)r] = (int64_t)std::count(text + rs[(size_t)r], text + rs[(size_t)r + 1], '\n');
            });
        for (auto& t : th) t.join();
    }
    std::vector<int64_t> first((size_t)nt + 1, 0);   // first line of each range
    for (int r = 0; r < nt; r++) {
        const bool tail = rs[(size_t)r + 1] == len && len > rs[(size_t)r] && text[len - 1] != '\n';
        first[(size_t)r + 1] = first[(size_t)r] + cnt[(size_t)r] + (tail ? 1 : 0);
    }
    std::vector<int64_t> got((size_t)nt, 0), bad((size_t)nt, 0);
    std::vector<int> rcs((size_t)nt, 0);
    {
        std::vector<std::thread> th;
        for (int r = 0; r < nt; r++) {
            const int64_t f = first[(size_t)r];
            if (f >= nlines || rs[(size_t)r] >= rs[(size_t)r + 1]) continue;
            const int64_t nl = std::min(first[(size_t)r + 1], nlines) - f;
            th.emplace_back([&, r, f, nl] {
                rcs[(size_t)r] = parse_lines(text, rs[(size_t)r + 1], rs[(size_t)r], f, nl, src + f, dst + f,
                                             &got[(size_t)r], &bad[(size_t)r]);
            });
        }
        for (auto& t : th) t.join();
    }
    // pairs of range r were written from index first[r]; move them down
    int64_t np = 0;
    for (int r = 0; r < nt; r++) {
        const int64_t off = first[(size_t)r];
        if (off >= nlines) break;
        if (np != off && got[(size_t)r] > 0) {
            std::memmove(src + np, src + off, sizeof(int64_t) * (size_t)got[(size_t)r]);
            std::memmove(dst + np, dst + off, sizeof(int64_t) * (size_t)got[(size_t)r]);
        }
        np += got[(size_t)r];
        if (rcs[(size_t)r] != 0) {
            *npairs = np;
            *consumed = rs[(size_t)r];
            *bad_line = bad[(size_t)r];
            return rcs[(size_t)r];
        }
    }
    *npairs = np;
    // where the nlines lines end: inside the range that holds line nlines - 1
    int64_t pos = len;
    for (int r = 0; r < nt; r++) {
        if (first[(size_t)r + 1] >= nlines) {
            pos = rs[(size_t)r];
            for (int64_t line = first[(size_t)r]; line < nlines && pos < len; line++) {
                const char* nlp = static_cast<const char*>(std::memchr(text + pos, '\n', (size_t)(len - pos)));
                pos = nlp ? (int64_t)(nlp - text) + 1 : len;
            }
            break;
        }
    }
    *consumed = pos;
    return 0;
}
