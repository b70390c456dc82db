// Readers of laboetie's two input files, host code behind the C-ABI.
//
// lb.in (module_input.f90): every line is left-adjusted and cut to 100
// characters, everything from a '#' on is blanked, and blank lines are
// dropped (put_input_in_character_array, :561-633).  A tag is found on the
// first line whose first len(tag) characters are the tag and whose next
// character is a blank; its value is read list-directed from columns
// len(tag)+4 .. len(tag)+50 (input_dp and its siblings, :62-75).  So
// `tag = value` with one blank either side of `=` is the accepted form, the
// first match wins and tags are case-sensitive.  Defaults live with the
// caller: a tag that is absent leaves the output untouched and reports
// found = 0.
//
// geom.in (construct_custom, module_geometry.f90:380-426): one solid node
// per line as 1-based `i j k`; everything else is fluid.  An index outside
// 1..lx, 1..ly, 1..lz stops the reference; here it returns LBX_ERR_ARG.
//
// supercell.xsf (print_supercell_xsf, module_io.f90:86-160): the same bytes
// as the reference program's file, so that the two can be diffed.  The
// comment lines are part of that file; node lines are x slowest, z fastest,
// in the list-directed form flang writes (a blank before each item,
// integers at their width, integral REAL(4) values as `N.`).

#include <cctype>
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/lbx.h"

struct lbx_lbin {
    std::vector<std::string> lines;  // the cleaned lines, each 100 characters
};

namespace {

constexpr size_t kLineLen = 100;

// read '(a)' into character(len=100) keeps the first 100 characters;
// adjustl / trim shift blanks only (a tab stays, so a tag followed by a tab
// is not found, as in the reference)
std::string clean_line(std::string raw) {
    if (!raw.empty() && raw.back() == '\r') raw.pop_back();  // CRLF line end
    if (raw.size() > kLineLen) raw.resize(kLineLen);
    size_t a = raw.find_first_not_of(' ');
    std::string s = a == std::string::npos ? std::string() : raw.substr(a);
    const size_t hash = s.find('#');
    if (hash != std::string::npos) s.resize(hash);
    a = s.find_first_not_of(' ');
    s = a == std::string::npos ? std::string() : s.substr(a);
    const size_t b = s.find_last_not_of(' ');
    s = b == std::string::npos ? std::string() : s.substr(0, b + 1);
    return s;
}

// the value field of a tag, or nullptr-equivalent (false) when absent
bool field_of(const lbx_lbin* h, const char* tag, std::string* out) {
    const size_t j = std::strlen(tag);
    if (j == 0) return false;
    for (const std::string& line : h->lines) {
        if (line.compare(0, j, tag) == 0 && line.size() > j && line[j] == ' ') {
            const size_t a = j + 3, b = j + 50;
            *out = a < line.size() ? line.substr(a, b - a) : std::string();
            return true;
        }
    }
    return false;
}

// list-directed tokens: separated by blanks, tabs or commas, '/' ends the record
std::vector<std::string> tokens(const std::string& field) {
    std::vector<std::string> out;
    std::string cur;
    for (char ch : field) {
        if (ch == '/') break;
        if (ch == ' ' || ch == '\t' || ch == '\r' || ch == ',') {
            if (!cur.empty()) out.push_back(cur), cur.clear();
        } else {
            cur.push_back(ch);
        }
    }
    if (!cur.empty()) out.push_back(cur);
    return out;
}

bool to_double(std::string t, double* v) {
    for (char& ch : t)
        if (ch == 'd' || ch == 'D') ch = 'e';
    char* end = nullptr;
    const double x = std::strtod(t.c_str(), &end);
    if (end == t.c_str() || *end != '\0') return false;
    *v = x;
    return true;
}

bool to_int(const std::string& t, int64_t* v) {
    if (t.find_first_of(".eEdD") != std::string::npos) {  // a real value read into an integer
        double x;
        if (!to_double(t, &x)) return false;
        *v = (int64_t)x;
        return true;
    }
    char* end = nullptr;
    const long long x = std::strtoll(t.c_str(), &end, 10);
    if (end == t.c_str() || *end != '\0') return false;
    *v = x;
    return true;
}

}  // namespace

extern "C" {

int lbx_lbin_parse(const char* text, int64_t len, lbx_lbin** out) {
    if (!out || (!text && len > 0) || len < 0) return LBX_ERR_ARG;
    *out = nullptr;
    lbx_lbin* h = new (std::nothrow) lbx_lbin;
    if (!h) return LBX_ERR_NOMEM;
    std::string raw;
    for (int64_t k = 0; k <= len; ++k) {
        if (k == len || text[k] == '\n') {
            std::string s = clean_line(raw);
            if (!s.empty()) {
                s.resize(kLineLen, ' ');
                h->lines.push_back(s);
            }
            raw.clear();
        } else {
            raw.push_back(text[k]);
        }
    }
    *out = h;
    return LBX_OK;
}

int lbx_lbin_free(lbx_lbin* h) {
    delete h;
    return LBX_OK;
}

int lbx_lbin_dp(const lbx_lbin* h, const char* tag, int32_t count, double* out, int32_t* found) {
    if (!h || !tag || !out || !found || count < 1) return LBX_ERR_ARG;
    std::string f;
    *found = field_of(h, tag, &f) ? 1 : 0;
    if (!*found) return LBX_OK;
    const std::vector<std::string> t = tokens(f);
    if ((int64_t)t.size() < count) return LBX_ERR_ARG;
    for (int32_t k = 0; k < count; ++k)
        if (!to_double(t[k], &out[k])) return LBX_ERR_ARG;
    return LBX_OK;
}

int lbx_lbin_int(const lbx_lbin* h, const char* tag, int32_t count, int64_t* out, int32_t* found) {
    if (!h || !tag || !out || !found || count < 1) return LBX_ERR_ARG;
    std::string f;
    *found = field_of(h, tag, &f) ? 1 : 0;
    if (!*found) return LBX_OK;
    const std::vector<std::string> t = tokens(f);
    if ((int64_t)t.size() < count) return LBX_ERR_ARG;
    for (int32_t k = 0; k < count; ++k)
        if (!to_int(t[k], &out[k])) return LBX_ERR_ARG;
    return LBX_OK;
}

int lbx_lbin_log(const lbx_lbin* h, const char* tag, int32_t* out, int32_t* found) {
    if (!h || !tag || !out || !found) return LBX_ERR_ARG;
    std::string f;
    *found = field_of(h, tag, &f) ? 1 : 0;
    if (!*found) return LBX_OK;
    const std::vector<std::string> t = tokens(f);
    if (t.empty()) return LBX_ERR_ARG;
    const size_t a = t[0].find_first_not_of('.');
    if (a == std::string::npos) return LBX_ERR_ARG;
    const char c = (char)std::toupper((unsigned char)t[0][a]);
    if (c != 'T' && c != 'F') return LBX_ERR_ARG;
    *out = c == 'T' ? 1 : 0;
    return LBX_OK;
}

int lbx_lbin_char(const lbx_lbin* h, const char* tag, char* out, int64_t cap, int32_t* found) {
    if (!h || !tag || !out || cap < 1 || !found) return LBX_ERR_ARG;
    std::string f;
    *found = field_of(h, tag, &f) ? 1 : 0;
    if (!*found) return LBX_OK;
    const std::vector<std::string> t = tokens(f);
    std::string v = t.empty() ? std::string() : t[0];
    const size_t a = v.find_first_not_of("'\""), b = v.find_last_not_of("'\"");
    v = a == std::string::npos ? std::string() : v.substr(a, b - a + 1);
    if ((int64_t)v.size() >= cap) return LBX_ERR_ARG;
    std::memcpy(out, v.c_str(), v.size() + 1);
    return LBX_OK;
}

int lbx_geom_in_parse(const char* text, int64_t len, int32_t lx, int32_t ly, int32_t lz, int8_t* nature,
                      int64_t* nsolid) {
    if ((!text && len > 0) || len < 0 || !nature || lx < 1 || ly < 1 || lz < 1) return LBX_ERR_ARG;
    const int64_t N = (int64_t)lx * ly * lz;
    std::memset(nature, 0, (size_t)N);
    int64_t marked = 0;
    std::string raw;
    for (int64_t k = 0; k <= len; ++k) {
        if (k < len && text[k] != '\n') {
            raw.push_back(text[k]);
            continue;
        }
        const std::vector<std::string> t = tokens(raw);
        raw.clear();
        if (t.size() < 3) continue;
        int64_t v[3];
        for (int a = 0; a < 3; ++a)
            if (!to_int(t[a], &v[a])) return LBX_ERR_ARG;
        if (v[0] < 1 || v[0] > lx || v[1] < 1 || v[1] > ly || v[2] < 1 || v[2] > lz) return LBX_ERR_ARG;
        const int64_t r = (v[0] - 1) + (int64_t)lx * ((v[1] - 1) + (int64_t)ly * (v[2] - 1));
        if (!nature[r]) ++marked;
        nature[r] = 1;
    }
    if (nsolid) *nsolid = marked;
    return LBX_OK;
}

// The lines before the node list: each `(xA)` record is a blank and the
// text; an empty record is an empty line.
static const char* const kXsfHead1[] = {
    "# this is the specification file of the supercell",
    "# lines beginning with # are commented. There cannot be comment lines within the sections",
    "# XSF format specifications can be found at http://www.xcrysden.org/doc/XSF.html",
    "# I strongly recommends to read this documentation.",
    nullptr,
    "# for periodic structures one has to begin with word CRYSTAL",
    "CRYSTAL",
    nullptr,
    "# Then one needs to specify the lattice vectors",
    "# specification of PRIMVEC (in ANGSTROMS) like:",
    "#         ax, ay, az    (first lattice vector)",
    "#         bx, by, bz    (second lattice vector)",
    "#         cx, cy, cz    (third lattice vector)",
    "# pay attention to vectors as they are written in horizontal way which is quite unusual",
    "# for now only orthorhombic structures allowed (a/=b/=c, all angles are 90 degrees)",
    "PRIMVEC",
};
static const char* const kXsfHead2[] = {
    nullptr,
    "# Then one needs to specify the atoms belonging to the unit cell. ",
    "# First number stands for number of atoms in the primitive cell (2 in this case).",
    "# The second number is always 1 for PRIMCOORD coordinates.",
    "# in angstroms and cartesian coordinates",
    "PRIMCOORD",
};

int lbx_supercell_xsf(int32_t lx, int32_t ly, int32_t lz, const int8_t* nature, const int8_t* interfacial,
                      char* out, int64_t cap, int64_t* len) {
    // flang writes an integral REAL(4) of 9.5e5 or more in E form, which this
    // writer does not reproduce: coordinates stay below 950000
    constexpr int32_t kMaxAxis = 950000;
    if (!nature || !interfacial || !len || lx < 1 || ly < 1 || lz < 1 || lx > kMaxAxis || ly > kMaxAxis ||
        lz > kMaxAxis || cap < 0 || (cap > 0 && !out))
        return LBX_ERR_ARG;
    std::string s;
    auto rec = [&](const char* t) {
        if (t) s.append(" ").append(t);
        s.push_back('\n');
    };
    for (const char* t : kXsfHead1) rec(t);
    char buf[96];
    const int32_t dims[3] = {lx, ly, lz};
    for (int a = 0; a < 3; ++a) {  // (3(xF10.5)) of (lx,0,0), (0,ly,0), (0,0,lz)
        for (int b = 0; b < 3; ++b) {
            std::snprintf(buf, sizeof buf, " %10.5f", a == b ? (double)(float)dims[a] : 0.0);
            s.append(buf);
        }
        s.push_back('\n');
    }
    for (const char* t : kXsfHead2) rec(t);
    std::snprintf(buf, sizeof buf, " %lld 1\n", (long long)lx * ly * lz);
    s.append(buf);
    // VMD colours: solid 4, interfacial fluid 45, other fluid 1
    for (int32_t i = 0; i < lx; ++i)
        for (int32_t j = 0; j < ly; ++j)
            for (int32_t k = 0; k < lz; ++k) {
                const int64_t r = i + (int64_t)lx * (j + (int64_t)ly * k);
                if (nature[r] != 0 && nature[r] != 1) return LBX_ERR_ARG;
                const int colour = nature[r] == 1 ? 4 : (interfacial[r] ? 45 : 1);
                std::snprintf(buf, sizeof buf, " %d %d. %d. %d.\n", colour, i, j, k);
                s.append(buf);
            }
    *len = (int64_t)s.size();
    if (cap == 0) return LBX_OK;
    if (cap < (int64_t)s.size()) return LBX_ERR_ARG;
    std::memcpy(out, s.data(), s.size());
    return LBX_OK;
}

}  // extern "C"
