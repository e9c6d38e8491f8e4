// mcm_textdeck.h -- reader of the reference's version-2 text input deck
// (host code of libmcm_gpu.so).
//
// Follows the reference's input chain for format version 2
// (mcm_getinput.f90:89-127): control cards (mcm_control_v2.f90:86-124,
// defaults :185-198), material cards (mcm_matin.f90:38-203), particle cards
// (mcm_nodein.f90:34-125), time-history node and transducer cards (skipped:
// output only), initial velocity cards with the generation rule
// (mcm_Inputvel.f90:29-129) and boundary planes (mcm_inp_symmetry_v2.f90:
// 23-98).  Cards are read as the reference reads them: a line is the first
// 80 characters of a record (character(len=80), mcm_gttxsg.f90:28-43), lines
// starting with '*' or '$' are comments, and fields are fixed-width Fortran
// I and E edit descriptors (blanks ignored, an empty field is zero, E input
// accepts 'D' and sign-only exponents).
//
// Penalty contact cards (mcm_contact_in.f90:22-45) and the base
// acceleration card (mcm_baseaccel_in.f90:25-31) follow the boundary planes.
// 1-D and 2-D problems are rejected with an error.
#pragma once

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace textdeck {

// One card source: the next non-comment record, cut or blank-padded to 80
// columns.
struct Cards {
    FILE* f = nullptr;
    long line = 0;
    bool next(std::string& out) {
        char buf[4096];
        while (fgets(buf, sizeof buf, f)) {
            line++;
            size_t n = strlen(buf);
            bool whole = n > 0 && buf[n - 1] == '\n';
            while (n > 0 && (buf[n - 1] == '\n' || buf[n - 1] == '\r')) buf[--n] = 0;
            if (!whole) {   // record longer than the buffer: drop the rest of it
                int ch;
                while ((ch = fgetc(f)) != EOF && ch != '\n') {}
            }
            if (buf[0] == '*' || buf[0] == '$') continue;
            out.assign(buf, n < 80 ? n : 80);
            out.resize(80, ' ');
            return true;
        }
        return false;
    }
};

// field [col, col + w) of a card, blanks removed
inline std::string field(const std::string& card, int col, int w) {
    std::string s;
    for (int k = col; k < col + w && k < (int)card.size(); k++)
        if (card[k] != ' ' && card[k] != '\t') s += card[k];
    return s;
}

// Fortran Iw input
inline bool fint(const std::string& card, int col, int w, long* out) {
    std::string s = field(card, col, w);
    if (s.empty()) { *out = 0; return true; }
    char* end = nullptr;
    long v = strtol(s.c_str(), &end, 10);
    if (*end != 0) return false;
    *out = v;
    return true;
}

// Fortran Ew.d input: mantissa with or without a point, exponent introduced by
// E, D or by its sign alone; a mantissa without a point has its last d
// digits after the implied point (Ew.0: none)
inline bool freal(const std::string& card, int col, int w, double* out, int d = 0) {
    std::string s = field(card, col, w);
    if (s.empty()) { *out = 0.0; return true; }
    if (d > 0) {
        size_t e = 1;   // end of the mantissa: first exponent letter or sign after the first character
        while (e < s.size() && !strchr("eEdDqQ+-", s[e])) e++;
        std::string mant = s.substr(0, e), rest = s.substr(e);
        if (mant.find('.') == std::string::npos) {
            std::string sign;
            if (!mant.empty() && (mant[0] == '+' || mant[0] == '-')) { sign = mant.substr(0, 1); mant = mant.substr(1); }
            while ((int)mant.size() <= d) mant = "0" + mant;
            mant = sign + mant.substr(0, mant.size() - d) + "." + mant.substr(mant.size() - d);
        }
        s = mant + rest;
    }
    std::string t;
    for (size_t k = 0; k < s.size(); k++) {
        char ch = s[k];
        if (ch == 'd' || ch == 'D' || ch == 'q' || ch == 'Q') ch = 'e';
        if ((ch == '+' || ch == '-') && k > 0) {
            char prev = (char)tolower((unsigned char)t.back());
            if (prev != 'e') t += 'e';
        }
        t += ch;
    }
    char* end = nullptr;
    double v = strtod(t.c_str(), &end);
    if (*end != 0) return false;
    *out = v;
    return true;
}

// A card read by a list of fixed-width fields: 'i' integer, 'e' real
struct Fmt {
    const std::string& card;
    int col = 0;
    bool ok = true;
    explicit Fmt(const std::string& c) : card(c) {}
    long i(int w) { long v = 0; ok = ok && fint(card, col, w, &v); col += w; return v; }
    double e(int w) { double v = 0.0; ok = ok && freal(card, col, w, &v); col += w; return v; }
    double ed(int w, int d) { double v = 0.0; ok = ok && freal(card, col, w, &v, d); col += w; return v; }
    void skip(int w) { col += w; }
};

}  // namespace textdeck
