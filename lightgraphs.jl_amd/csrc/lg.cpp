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

#include <cstdint>

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

}  // namespace

extern "C" int lgmi_lg_parse_edges(const char* text, int64_t len, int64_t nlines, int64_t* src, int64_t* dst,
                                   int64_t* npairs, int64_t* consumed, int64_t* bad_line) {
    if (!npairs || !consumed || !bad_line || len < 0 || nlines < 0 || (len > 0 && !text) ||
        (nlines > 0 && (!src || !dst)))
        return -4;
    *npairs = 0;
    *consumed = 0;
    *bad_line = 0;
    int64_t pos = 0, np = 0;
    for (int64_t line = 1; line <= nlines; line++) {
        if (pos >= len) break;   // readline at the end of the stream gives ""
        int64_t e = pos;
        while (e < len && text[e] != '\n') e++;
        const int64_t next = e < len ? e + 1 : e;
        int64_t le = e;
        if (e < len && le > pos && text[le - 1] == '\r') le--;   // chomp removes "\r\n" whole
        if (le > pos) {
            // the first comma ends field one, the next comma (if any) field two
            int64_t c1 = pos;
            while (c1 < le && text[c1] != ',') c1++;
            if (c1 >= le) {
                *bad_line = line;
                *consumed = pos;
                *npairs = np;
                return -1;
            }
            int64_t c2 = c1 + 1;
            while (c2 < le && text[c2] != ',') c2++;
            int64_t s = 0, d = 0;
            int rc = parse_i64(text, pos, c1, &s);
            if (rc == 0) rc = parse_i64(text, c1 + 1, c2, &d);
            if (rc != 0) {
                *bad_line = line;
                *consumed = pos;
                *npairs = np;
                return rc;
            }
            src[np] = s;
            dst[np] = d;
            np++;
        }
        pos = next;
    }
    *npairs = np;
    *consumed = pos;
    return 0;
}
