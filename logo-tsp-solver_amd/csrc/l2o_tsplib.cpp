// l2o_tsplib.cpp -- TSPLIB .tsp / .tour files and host tour lengths (host C++).
//
// Replaces the reference's readFileCoords (src/utils/tsp_loader.c:49-162),
// routeInitFromFile (src/utils/common.c:109-129) and saveRouteToAFile
// (src/utils/common.c:153-192).  The reference's reader skips three lines,
// takes DIMENSION from the fourth, expects NODE_COORD_SECTION on the sixth and
// treats every file as EUC_2D; this one parses the header by keyword (TSPLIB95
// "KEY : value" lines in any order), reads EDGE_WEIGHT_TYPE (EUC_2D, CEIL_2D,
// ATT) and reports malformed files as errors instead of exiting.
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/logo2opt.h"
#include "l2o_launch.h"

namespace {

bool read_file(const char *path, std::string *out)
{
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    char buf[1 << 16];
    size_t got;
    out->clear();
    while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) out->append(buf, got);
    const bool ok = !std::ferror(f);
    std::fclose(f);
    return ok;
}

std::vector<std::string> split_lines(const std::string &s)
{
    std::vector<std::string> lines;
    size_t a = 0;
    while (a <= s.size()) {
        size_t b = s.find('\n', a);
        if (b == std::string::npos) b = s.size();
        std::string l = s.substr(a, b - a);
        if (!l.empty() && l.back() == '\r') l.pop_back();
        lines.push_back(l);
        a = b + 1;
    }
    return lines;
}

std::string trim(const std::string &s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

std::vector<std::string> tokens(const std::string &s)
{
    std::vector<std::string> t;
    size_t k = 0;
    while (k < s.size()) {
        while (k < s.size() && std::isspace((unsigned char)s[k])) ++k;
        const size_t a = k;
        while (k < s.size() && !std::isspace((unsigned char)s[k])) ++k;
        if (k > a) t.push_back(s.substr(a, k - a));
    }
    return t;
}

// leading keyword [A-Z0-9_]+ of a trimmed line; empty if the line has none
std::string keyword(const std::string &l)
{
    size_t k = 0;
    while (k < l.size() && (std::isupper((unsigned char)l[k]) || std::isdigit((unsigned char)l[k]) || l[k] == '_'))
        ++k;
    if (k == 0 || !std::isupper((unsigned char)l[0])) return std::string();
    return l.substr(0, k);
}

// "KEY : value" (spaces optional); false if the line is not of that form
bool key_value(const std::string &l, std::string *key, std::string *value)
{
    *key = keyword(l);
    if (key->empty()) return false;
    size_t k = key->size();
    while (k < l.size() && std::isspace((unsigned char)l[k])) ++k;
    if (k >= l.size() || l[k] != ':') return false;
    *value = trim(l.substr(k + 1));
    return true;
}

bool parse_u64(const std::string &t, uint64_t *v)
{
    if (t.empty() || !std::isdigit((unsigned char)t[0])) return false;
    errno = 0;
    char *end = nullptr;
    const unsigned long long r = std::strtoull(t.c_str(), &end, 10);
    if (errno || *end) return false;
    *v = r;
    return true;
}

bool parse_double(const std::string &t, double *v)
{
    if (t.empty()) return false;
    char *end = nullptr;
    const double r = std::strtod(t.c_str(), &end);
    if (*end || !std::isfinite(r)) return false;
    *v = r;
    return true;
}

bool is_section(const std::string &kw)
{
    const size_t n = kw.size(), m = sizeof("_SECTION") - 1;
    return n > m && kw.compare(n - m, m, "_SECTION") == 0;
}

}  // namespace

extern "C" int logo2opt_tsp_read(const char *path, uint32_t flags, logo2opt_tsp_info *info,
                                 logo2opt_city_coords *coords, uint32_t cap)
{
    if (!path || !info) return LOGO2OPT_EINVAL;
    std::memset(info, 0, sizeof(*info));
    info->metric = -1;
    std::string text;
    if (!read_file(path, &text)) return LOGO2OPT_EIO;
    const std::vector<std::string> lines = split_lines(text);

    uint64_t n = 0;
    bool have_dim = false, have_type = false, in_skip = false;
    std::vector<double> xs, ys;
    bool have_coords = false;
    for (size_t li = 0; li < lines.size(); ++li) {
        const std::string l = trim(lines[li]);
        if (l.empty()) continue;
        const std::string kw = keyword(l);
        if (in_skip) {  // data lines of a section this reader does not use
            if (kw.empty()) continue;
            in_skip = false;
        }
        if (kw == "EOF" && trim(l.substr(3)).empty()) break;
        if (kw == "NODE_COORD_SECTION" && trim(l.substr(kw.size())).empty()) {
            if (!have_dim || have_coords) return LOGO2OPT_EFORMAT;
            xs.assign(n, 0.0);
            ys.assign(n, 0.0);
            std::vector<unsigned char> seen(n, 0);
            uint64_t got = 0;
            while (got < n) {
                if (++li >= lines.size()) return LOGO2OPT_EFORMAT;  // too few entries
                const std::vector<std::string> t = tokens(lines[li]);
                if (t.empty()) continue;
                uint64_t id;
                double x, y;
                if (t.size() != 3 || !parse_u64(t[0], &id) || !parse_double(t[1], &x) || !parse_double(t[2], &y))
                    return LOGO2OPT_EFORMAT;
                if (id < 1 || id > n || seen[id - 1]) return LOGO2OPT_EFORMAT;
                seen[id - 1] = 1;
                xs[id - 1] = x;
                ys[id - 1] = y;
                ++got;
            }
            have_coords = true;
            continue;
        }
        if (is_section(kw) && trim(l.substr(kw.size())).empty()) {
            in_skip = true;  // e.g. DISPLAY_DATA_SECTION
            continue;
        }
        std::string key, value;
        if (!key_value(l, &key, &value)) return LOGO2OPT_EFORMAT;
        if (key == "NAME") {
            std::snprintf(info->name, sizeof(info->name), "%s", value.c_str());
        } else if (key == "TYPE") {
            if (value != "TSP") return LOGO2OPT_EFORMAT;
        } else if (key == "DIMENSION") {
            if (have_coords || !parse_u64(value, &n) || n < 3 || n > LOGO2OPT_MAX_N) return LOGO2OPT_EFORMAT;
            have_dim = true;
        } else if (key == "EDGE_WEIGHT_TYPE") {
            if (value == "EUC_2D") info->metric = LOGO2OPT_METRIC_EUC_2D;
            else if (value == "CEIL_2D") info->metric = LOGO2OPT_METRIC_CEIL_2D;
            else if (value == "ATT") info->metric = LOGO2OPT_METRIC_ATT;
            else return LOGO2OPT_EFORMAT;  // GEO, EXPLICIT, EUC_3D, MAN_2D, ...
            have_type = true;
        } else if (key == "NODE_COORD_TYPE") {
            if (value != "TWOD_COORDS") return LOGO2OPT_EFORMAT;
        }
        // COMMENT, DISPLAY_DATA_TYPE and other keywords carry nothing used here
    }
    if (!have_dim || !have_type || !have_coords) return LOGO2OPT_EFORMAT;
    info->n = (uint32_t)n;

    bool integral = true;
    for (uint64_t k = 0; k < n; ++k) integral &= xs[k] == std::floor(xs[k]) && ys[k] == std::floor(ys[k]);
    info->integral = integral ? 1u : 0u;
    if (!integral) {
        if (!(flags & LOGO2OPT_TSP_ROUND)) return LOGO2OPT_ECOORD;
        for (uint64_t k = 0; k < n; ++k) {
            xs[k] = std::round(xs[k]);
            ys[k] = std::round(ys[k]);
        }
    }
    double mnx = xs[0], mny = ys[0], mxx = xs[0], mxy = ys[0];
    for (uint64_t k = 1; k < n; ++k) {
        mnx = std::fmin(mnx, xs[k]);
        mny = std::fmin(mny, ys[k]);
        mxx = std::fmax(mxx, xs[k]);
        mxy = std::fmax(mxy, ys[k]);
    }
    info->offset_x = mnx;
    info->offset_y = mny;
    const double lim = (double)(1u << 22);
    if (mxx - mnx >= lim || mxy - mny >= lim) return LOGO2OPT_ECOORD;
    if (coords) {
        if (cap < n) return LOGO2OPT_EINVAL;
        for (uint64_t k = 0; k < n; ++k) {
            coords[k].x = (float)(xs[k] - mnx);  // integers below 2^22: exact
            coords[k].y = (float)(ys[k] - mny);
        }
    }
    return LOGO2OPT_OK;
}

extern "C" int logo2opt_tour_read(const char *path, uint32_t n, uint32_t *route)
{
    if (!path || !route || n < 3 || n > LOGO2OPT_MAX_N) return LOGO2OPT_EINVAL;
    std::string text;
    if (!read_file(path, &text)) return LOGO2OPT_EIO;
    const std::vector<std::string> lines = split_lines(text);
    // TSPLIB form when a TOUR_SECTION line exists, else the header-less form
    size_t first = 0;
    bool tsplib = false;
    for (size_t li = 0; li < lines.size(); ++li)
        if (trim(lines[li]) == "TOUR_SECTION") {
            tsplib = true;
            first = li + 1;
            break;
        }
    if (tsplib) {
        for (size_t li = 0; li + 1 < first; ++li) {
            const std::string l = trim(lines[li]);
            if (l.empty()) continue;
            std::string key, value;
            if (!key_value(l, &key, &value)) return LOGO2OPT_EFORMAT;
            if (key == "TYPE" && value != "TOUR") return LOGO2OPT_EFORMAT;
            if (key == "DIMENSION") {
                uint64_t d;
                if (!parse_u64(value, &d) || d != n) return LOGO2OPT_EFORMAT;
            }
        }
    }
    std::vector<uint32_t> ids;
    ids.reserve(n);
    bool ended = false;
    for (size_t li = first; li < lines.size() && !ended; ++li) {
        for (const std::string &t : tokens(lines[li])) {
            if (tsplib && (t == "-1" || t == "EOF")) {
                ended = true;  // the first tour only
                break;
            }
            // the header-less form reads the first n ids and ignores what
            // follows, as routeInitFromFile's n fscanf calls do (a trailing
            // "-1", say); a TOUR_SECTION holds exactly n ids before its -1
            if (!tsplib && ids.size() == n) {
                ended = true;
                break;
            }
            uint64_t v;
            if (!parse_u64(t, &v) || ids.size() >= n) return LOGO2OPT_EFORMAT;
            if (v < 1 || v > n) return LOGO2OPT_ETOUR;
            ids.push_back((uint32_t)(v - 1));
        }
    }
    if (ids.size() != n) return LOGO2OPT_EFORMAT;
    std::vector<unsigned char> seen(n, 0);
    for (uint32_t v : ids) {
        if (seen[v]) return LOGO2OPT_ETOUR;
        seen[v] = 1;
    }
    std::memcpy(route, ids.data(), sizeof(uint32_t) * n);
    return LOGO2OPT_OK;
}

extern "C" int logo2opt_tour_write(const char *path, const char *name, uint32_t n, const uint32_t *route,
                                   uint64_t length, int headerless)
{
    if (!path || !route || n < 3 || n > LOGO2OPT_MAX_N) return LOGO2OPT_EINVAL;
    std::vector<unsigned char> seen(n, 0);
    for (uint32_t k = 0; k < n; ++k) {
        if (route[k] >= n || seen[route[k]]) return LOGO2OPT_ETOUR;
        seen[route[k]] = 1;
    }
    FILE *f = std::fopen(path, "w");
    if (!f) return LOGO2OPT_EIO;
    bool ok = true;
    if (!headerless) {
        ok &= std::fprintf(f, "NAME : %s\n", name && *name ? name : "tour") > 0;
        if (length) ok &= std::fprintf(f, "COMMENT : length %llu\n", (unsigned long long)length) > 0;
        ok &= std::fprintf(f, "TYPE : TOUR\nDIMENSION : %u\nTOUR_SECTION\n", n) > 0;
    }
    for (uint32_t k = 0; k < n && ok; ++k) ok &= std::fprintf(f, "%u\n", route[k] + 1) > 0;
    if (!headerless) ok &= std::fprintf(f, "-1\nEOF\n") > 0;
    ok &= std::fclose(f) == 0;
    return ok ? LOGO2OPT_OK : LOGO2OPT_EIO;
}

extern "C" int logo2opt_host_route_length(uint32_t n, const logo2opt_city_coords *coords, int metric,
                                          const uint32_t *route, uint64_t *length)
{
    if (!coords || !route || !length || n < 3 || n > LOGO2OPT_MAX_N) return LOGO2OPT_EINVAL;
    if (metric < LOGO2OPT_METRIC_EUC_2D || metric > LOGO2OPT_METRIC_ATT) return LOGO2OPT_EINVAL;
    float mnx = coords[0].x, mny = coords[0].y, mxx = mnx, mxy = mny;
    for (uint32_t k = 0; k < n; ++k) {
        const float x = coords[k].x, y = coords[k].y;
        if (!std::isfinite(x) || !std::isfinite(y) || x != std::floor(x) || y != std::floor(y))
            return LOGO2OPT_ECOORD;
        mnx = std::fmin(mnx, x);
        mny = std::fmin(mny, y);
        mxx = std::fmax(mxx, x);
        mxy = std::fmax(mxy, y);
    }
    if ((double)mxx - mnx >= (double)(1u << 22) || (double)mxy - mny >= (double)(1u << 22)) return LOGO2OPT_ECOORD;
    std::vector<unsigned char> seen(n, 0);
    for (uint32_t k = 0; k < n; ++k) {
        if (route[k] >= n || seen[route[k]]) return LOGO2OPT_ETOUR;
        seen[route[k]] = 1;
    }
    uint64_t L = 0;
    for (uint32_t k = 0; k < n; ++k) {
        const logo2opt_city_coords &a = coords[route[k == 0 ? n - 1 : k - 1]], &b = coords[route[k]];
        const long long dx = (long long)a.x - (long long)b.x, dy = (long long)a.y - (long long)b.y;
        L += (uint64_t)l2o::metric_dist(dx * dx + dy * dy, (uint32_t)metric);
    }
    *length = L;
    return LOGO2OPT_OK;
}
