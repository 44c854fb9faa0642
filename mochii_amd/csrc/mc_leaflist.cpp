// Leaf-list input files (generic_amr_read, src/read_generic_amr.f90:34-237):
// the text form and the FITS binary table, gzip-compressed or not, read on
// the host into the arrays mc_grid_octree takes.  HDF5 files are refused
// (no HDF5 library on this path).
//
// Text: a header line "nleaf boxlen", then one leaf per line
// "x y z level nH [T vx vy vz ...]"; columns past the fifth are ignored.
// FITS: the first extension is a BINTABLE; columns are found by name (x, y,
// z, level, nH | dens | gasDen; case-insensitive) with the index fallback
// 1..5 of the reference, keywords NAXIS2, BOXLEN and ORIGINX/Y/Z (default
// -BOXLEN/2); optional columns metallicity, xHI, ndust.  Column types D, E,
// J, K, I, B (repeat 1), with TSCALn / TZEROn applied.
#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "mc_internal.h"

namespace {

constexpr size_t kFitsBlock = 2880, kCard = 80;

bool ends_with(const std::string &s, const char *suf)
{
    const size_t n = std::strlen(suf);
    if (s.size() < n) return false;
    for (size_t i = 0; i < n; ++i)
        if (std::tolower((unsigned char)s[s.size() - n + i]) != suf[i]) return false;
    return true;
}

bool file_exists(const std::string &p)
{
    FILE *f = std::fopen(p.c_str(), "rb");
    if (!f) return false;
    std::fclose(f);
    return true;
}

// io_resolve_filename (src/iofile_mod.f90:138-170): a name without an
// extension after its last '/' gets .h5, .fits.gz, .fits, .dat, the first
// that exists
bool resolve(const std::string &name, std::string &out)
{
    const size_t slash = name.find_last_of('/');
    const std::string base = slash == std::string::npos ? name : name.substr(slash + 1);
    if (base.find('.') != std::string::npos) {
        out = name;
        return true;
    }
    for (const char *ext : {".h5", ".fits.gz", ".fits", ".dat"}) {
        if (file_exists(name + ext)) {
            out = name + ext;
            return true;
        }
    }
    return false;
}

// whole file, gunzipped when it is gzip data (gzread passes plain files through)
bool slurp(const std::string &path, std::string &data)
{
    gzFile f = gzopen(path.c_str(), "rb");
    if (!f) return false;
    data.clear();
    std::vector<char> buf(1 << 20);
    for (;;) {
        const int n = gzread(f, buf.data(), (unsigned)buf.size());
        if (n < 0) {
            gzclose(f);
            return false;
        }
        if (n == 0) break;
        data.append(buf.data(), (size_t)n);
    }
    gzclose(f);
    return true;
}

struct Columns {
    int64_t n = 0;
    double boxlen = 0.0, origin[3] = {0.0, 0.0, 0.0};
    std::vector<double> x, y, z, nH, opt[3];
    std::vector<int32_t> level;
    bool has[3] = {false, false, false};
};

// ---- text ----------------------------------------------------------------
bool read_text(const std::string &data, Columns &c, std::string &err)
{
    std::istringstream in(data);
    std::string line;
    if (!std::getline(in, line)) {
        err = "empty leaf-list file";
        return false;
    }
    {
        std::istringstream h(line);
        double nl = 0.0;
        if (!(h >> nl >> c.boxlen) || nl < 0.0) {
            err = "leaf-list header is not 'nleaf boxlen'";
            return false;
        }
        c.n = (int64_t)nl;
    }
    c.x.resize(c.n); c.y.resize(c.n); c.z.resize(c.n); c.nH.resize(c.n); c.level.resize(c.n);
    int64_t k = 0;
    while (k < c.n && std::getline(in, line)) {
        std::istringstream r(line);
        double x, y, z, lv, nh;
        if (!(r >> x >> y >> z >> lv >> nh)) continue;     // blank or short line
        c.x[k] = x; c.y[k] = y; c.z[k] = z; c.level[k] = (int32_t)lv; c.nH[k] = nh;
        ++k;
    }
    if (k != c.n) {
        err = "leaf-list file ends after " + std::to_string(k) + " of " + std::to_string(c.n) +
              " leaves";
        return false;
    }
    for (int d = 0; d < 3; ++d) c.origin[d] = -0.5 * c.boxlen;
    return true;
}

// ---- FITS ----------------------------------------------------------------
struct Header {
    std::vector<std::pair<std::string, std::string>> cards;
    size_t end = 0;      // offset of the data after the header blocks
};

std::string trim(const std::string &s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

bool parse_header(const std::string &d, size_t at, Header &h)
{
    for (size_t p = at; p + kCard <= d.size(); p += kCard) {
        const std::string card = d.substr(p, kCard);
        const std::string key = trim(card.substr(0, 8));
        if (key == "END") {
            h.end = (p + kCard + kFitsBlock - 1) / kFitsBlock * kFitsBlock;
            return h.end <= d.size();
        }
        if (card.size() >= 10 && card[8] == '=') {
            std::string v = card.substr(10);
            // a quoted string keeps its text; else the value ends at a '/'
            const std::string tv = trim(v);
            if (!tv.empty() && tv[0] == '\'') {
                const size_t q = tv.find('\'', 1);
                v = tv.substr(1, q == std::string::npos ? std::string::npos : q - 1);
            } else {
                const size_t sl = v.find('/');
                if (sl != std::string::npos) v = v.substr(0, sl);
            }
            h.cards.emplace_back(key, trim(v));
        }
    }
    return false;
}

const std::string *card(const Header &h, const std::string &key)
{
    for (const auto &kv : h.cards)
        if (kv.first == key) return &kv.second;
    return nullptr;
}

bool card_num(const Header &h, const std::string &key, double &v)
{
    const std::string *s = card(h, key);
    if (!s) return false;
    std::string t = *s;
    std::replace(t.begin(), t.end(), 'D', 'E');       // Fortran exponent letter
    char *e = nullptr;
    v = std::strtod(t.c_str(), &e);
    return e != t.c_str();
}

std::string lower(std::string s)
{
    for (char &ch : s) ch = (char)std::tolower((unsigned char)ch);
    return s;
}

struct ColInfo {
    char type = 0;
    size_t off = 0, width = 0;
    double scale = 1.0, zero = 0.0;
};

uint64_t be_bits(const unsigned char *p, size_t w)
{
    uint64_t v = 0;
    for (size_t i = 0; i < w; ++i) v = (v << 8) | p[i];
    return v;
}

double cell(const unsigned char *row, const ColInfo &ci)
{
    const unsigned char *p = row + ci.off;
    double v = 0.0;
    switch (ci.type) {
    case 'D': {
        const uint64_t b = be_bits(p, 8);
        std::memcpy(&v, &b, 8);
        break;
    }
    case 'E': {
        const uint32_t b = (uint32_t)be_bits(p, 4);
        float f;
        std::memcpy(&f, &b, 4);
        v = f;
        break;
    }
    case 'K': v = (double)(int64_t)be_bits(p, 8); break;
    case 'J': v = (double)(int32_t)(uint32_t)be_bits(p, 4); break;
    case 'I': v = (double)(int16_t)(uint16_t)be_bits(p, 2); break;
    case 'B': v = (double)p[0]; break;
    default: break;
    }
    return ci.zero + ci.scale * v;
}

bool read_fits(const std::string &d, Columns &c, std::string &err)
{
    Header h0, h1;
    if (d.compare(0, 9, "SIMPLE  =") != 0 || !parse_header(d, 0, h0)) {
        err = "not a FITS file";
        return false;
    }
    // skip the primary data unit (NAXIS = 0 for a leaf list)
    double naxis = 0.0, bitpix = 8.0;
    card_num(h0, "NAXIS", naxis);
    card_num(h0, "BITPIX", bitpix);
    size_t pdata = naxis > 0 ? (size_t)std::fabs(bitpix) / 8 : 0;
    for (int a = 1; a <= (int)naxis; ++a) {
        double na = 0.0;
        card_num(h0, "NAXIS" + std::to_string(a), na);
        pdata *= (size_t)na;
    }
    const size_t ext = h0.end + (pdata + kFitsBlock - 1) / kFitsBlock * kFitsBlock;
    if (!parse_header(d, ext, h1)) {
        err = "FITS file has no binary-table extension";
        return false;
    }
    const std::string *xt = card(h1, "XTENSION");
    if (!xt || trim(*xt) != "BINTABLE") {
        err = "first FITS extension is not a BINTABLE";
        return false;
    }
    double rowlen = 0.0, nrow = 0.0, nfield = 0.0;
    if (!card_num(h1, "NAXIS1", rowlen) || !card_num(h1, "NAXIS2", nrow) ||
        !card_num(h1, "TFIELDS", nfield)) {
        err = "BINTABLE header lacks NAXIS1 / NAXIS2 / TFIELDS";
        return false;
    }
    if (!card_num(h1, "BOXLEN", c.boxlen)) {
        err = "leaf-list table has no BOXLEN keyword";
        return false;
    }
    const char *okeys[3] = {"ORIGINX", "ORIGINY", "ORIGINZ"};
    for (int k = 0; k < 3; ++k)
        if (!card_num(h1, okeys[k], c.origin[k])) c.origin[k] = -0.5 * c.boxlen;
    std::vector<ColInfo> cols((size_t)nfield);
    std::vector<std::string> names((size_t)nfield);
    size_t off = 0;
    for (int f = 0; f < (int)nfield; ++f) {
        const std::string fs = std::to_string(f + 1);
        const std::string *form = card(h1, "TFORM" + fs);
        if (!form) {
            err = "BINTABLE column " + fs + " has no TFORM";
            return false;
        }
        std::string t = trim(*form);
        size_t i = 0;
        long rep = 0;
        while (i < t.size() && std::isdigit((unsigned char)t[i])) rep = rep * 10 + (t[i++] - '0');
        if (i == 0) rep = 1;
        const char ty = i < t.size() ? (char)std::toupper((unsigned char)t[i]) : 0;
        size_t w = 0;
        switch (ty) {
        case 'D': case 'K': w = 8; break;
        case 'E': case 'J': w = 4; break;
        case 'I': w = 2; break;
        case 'B': case 'L': case 'A': case 'X': w = 1; break;
        default:
            err = "BINTABLE column " + fs + " has unsupported TFORM " + t;
            return false;
        }
        ColInfo &ci = cols[f];
        ci.type = rep == 1 ? ty : 0;          // only scalar columns are read
        ci.off = off;
        ci.width = w;
        off += (ty == 'X' ? (size_t)(rep + 7) / 8 : w * (size_t)rep);
        card_num(h1, "TSCAL" + fs, ci.scale);
        card_num(h1, "TZERO" + fs, ci.zero);
        const std::string *tt = card(h1, "TTYPE" + fs);
        names[f] = tt ? lower(trim(*tt)) : std::string();
    }
    if (off != (size_t)rowlen) {
        err = "BINTABLE column widths do not add up to NAXIS1";
        return false;
    }
    c.n = (int64_t)nrow;
    const size_t need = h1.end + (size_t)rowlen * (size_t)nrow;
    if (need > d.size()) {
        err = "FITS file is shorter than its table";
        return false;
    }
    // a column by candidate names, else the 1-based fallback index
    auto find = [&](std::initializer_list<const char *> cand, int fallback) -> int {
        for (const char *nm : cand)
            for (int f = 0; f < (int)nfield; ++f)
                if (names[f] == nm) return f;
        return fallback - 1;
    };
    const int ix = find({"x"}, 1), iy = find({"y"}, 2), iz = find({"z"}, 3);
    const int il = find({"level"}, 4), in = find({"nh", "dens", "gasden"}, 5);
    for (int f : {ix, iy, iz, il, in})
        if (f < 0 || f >= (int)nfield || cols[f].type == 0 || cols[f].type == 'L') {
            err = "leaf-list table lacks a readable x / y / z / level / nH column";
            return false;
        }
    const int io[3] = {find({"metallicity"}, 0), find({"xhi"}, 0), find({"ndust"}, 0)};
    const unsigned char *base = (const unsigned char *)d.data() + h1.end;
    c.x.resize(c.n); c.y.resize(c.n); c.z.resize(c.n); c.nH.resize(c.n); c.level.resize(c.n);
    for (int k = 0; k < 3; ++k) {
        c.has[k] = io[k] >= 0 && cols[io[k]].type != 0 && cols[io[k]].type != 'L';
        if (c.has[k]) c.opt[k].resize(c.n);
    }
    for (int64_t r = 0; r < c.n; ++r) {
        const unsigned char *row = base + (size_t)r * (size_t)rowlen;
        c.x[r] = cell(row, cols[ix]);
        c.y[r] = cell(row, cols[iy]);
        c.z[r] = cell(row, cols[iz]);
        c.level[r] = (int32_t)cell(row, cols[il]);
        c.nH[r] = cell(row, cols[in]);
        for (int k = 0; k < 3; ++k)
            if (c.has[k]) c.opt[k][r] = cell(row, cols[io[k]]);
    }
    return true;
}

}  // namespace

extern "C" int mc_leaf_list_read(const char *path, int64_t maxn, int64_t *nleaf, double *half,
                                 double *x, double *y, double *z, int32_t *level, double *nH,
                                 double *metallicity, double *xHI, double *ndust,
                                 int32_t *has_opt)
{
    if (!path || !nleaf || !half) {
        mc_set_error("mc_leaf_list_read: null argument");
        return -1;
    }
    std::string file;
    if (!resolve(path, file)) {
        mc_set_error(std::string("leaf-list file not found (tried .h5 .fits.gz .fits .dat): ") +
                     path);
        return -1;
    }
    if (ends_with(file, ".h5") || ends_with(file, ".hdf5")) {
        mc_set_error("HDF5 leaf lists are not read by this library (FITS or text): " + file);
        return -1;
    }
    std::string data;
    if (!slurp(file, data)) {
        mc_set_error("cannot read leaf-list file " + file);
        return -1;
    }
    Columns c;
    std::string err;
    const bool binary = ends_with(file, ".fits") || ends_with(file, ".fits.gz");
    if (!(binary ? read_fits(data, c, err) : read_text(data, c, err))) {
        mc_set_error(err + " (" + file + ")");
        return -1;
    }
    if (!(c.boxlen > 0.0)) {
        mc_set_error("leaf-list box length must be positive (" + file + ")");
        return -1;
    }
    *nleaf = c.n;
    *half = 0.5 * c.boxlen;
    if (has_opt)
        for (int k = 0; k < 3; ++k) has_opt[k] = c.has[k] ? 1 : 0;
    if (!x || maxn < c.n) return 0;           // size query
    if (!y || !z || !level || !nH) {
        mc_set_error("mc_leaf_list_read: null output array");
        return -1;
    }
    // the box recentred on the origin (src/grid_mod_amr.f90:57-60)
    const double cx = c.origin[0] + 0.5 * c.boxlen, cy = c.origin[1] + 0.5 * c.boxlen,
                 cz = c.origin[2] + 0.5 * c.boxlen;
    for (int64_t i = 0; i < c.n; ++i) {
        x[i] = c.x[i] - cx;
        y[i] = c.y[i] - cy;
        z[i] = c.z[i] - cz;
        level[i] = c.level[i];
        nH[i] = c.nH[i];
    }
    double *opt[3] = {metallicity, xHI, ndust};
    for (int k = 0; k < 3; ++k)
        if (opt[k] && c.has[k]) std::memcpy(opt[k], c.opt[k].data(), sizeof(double) * c.n);
    return 0;
}
