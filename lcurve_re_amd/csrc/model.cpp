// model.cpp -- Lcurve::Model JSON parameter API over lc_model (host only).
//
// Behaviour follows src/lcurve_base/model.cpp of the reference:
//   - the ctor (model.cpp:10-360): "model_parameters" object of string
//     values, Pparam strings "value range dstep vary [defined]"
//     (lcurve.h:169-176), computational keys as scalars in strings,
//     pdot/third/temp_edge/absorb_edge defaulting to "0 1.e-10 1.e-10 0 0",
//     star-spot groups all-or-none;
//   - nvary / set_param / get_param / get_range / get_dstep / get_limit /
//     is_not_legal / get_name share one ordering (model.cpp:363-1660),
//     expressed here as one table instead of ten hand-written chains.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/lcurve_amd.h"
#include "lc_host.h"

namespace {

// ------------------------------------------------------------ tiny JSON
struct JVal {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    bool b = false;
    double num = 0.;
    std::string str;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;
};

struct JParser {
    const char *p, *end;
    [[noreturn]] void fail(const char *what) {
        throw std::runtime_error(std::string("JSON parse error: ") + what);
    }
    void ws() {
        while (p < end && (*p == ' ' || *p == '\n' || *p == '\t' || *p == '\r')) p++;
    }
    bool lit(const char *s) {
        size_t n = std::strlen(s);
        if ((size_t)(end - p) >= n && std::strncmp(p, s, n) == 0) {
            p += n;
            return true;
        }
        return false;
    }
    static void put_utf8(std::string &o, unsigned cp) {
        if (cp < 0x80) {
            o += (char)cp;
        } else if (cp < 0x800) {
            o += (char)(0xC0 | (cp >> 6));
            o += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            o += (char)(0xE0 | (cp >> 12));
            o += (char)(0x80 | ((cp >> 6) & 0x3F));
            o += (char)(0x80 | (cp & 0x3F));
        } else {
            o += (char)(0xF0 | (cp >> 18));
            o += (char)(0x80 | ((cp >> 12) & 0x3F));
            o += (char)(0x80 | ((cp >> 6) & 0x3F));
            o += (char)(0x80 | (cp & 0x3F));
        }
    }
    unsigned hex4() {
        if (end - p < 4) fail("short \\u escape");
        unsigned v = 0;
        for (int i = 0; i < 4; i++) {
            char c = *p++;
            v <<= 4;
            if (c >= '0' && c <= '9') v |= c - '0';
            else if (c >= 'a' && c <= 'f') v |= c - 'a' + 10;
            else if (c >= 'A' && c <= 'F') v |= c - 'A' + 10;
            else fail("bad \\u escape");
        }
        return v;
    }
    std::string string() {
        if (p >= end || *p != '"') fail("expected string");
        p++;
        std::string o;
        while (true) {
            if (p >= end) fail("unterminated string");
            char c = *p++;
            if (c == '"') break;
            if (c == '\\') {
                if (p >= end) fail("bad escape");
                char e = *p++;
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
                    unsigned cp = hex4();
                    if (cp >= 0xD800 && cp < 0xDC00 && end - p >= 6 && p[0] == '\\' && p[1] == 'u') {
                        p += 2;
                        unsigned lo = hex4();
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    put_utf8(o, cp);
                    break;
                }
                default: fail("bad escape");
                }
            } else {
                o += c;
            }
        }
        return o;
    }
    JVal value(int depth) {
        if (depth > 256) fail("nesting too deep");
        ws();
        if (p >= end) fail("unexpected end");
        JVal v;
        char c = *p;
        if (c == '{') {
            p++;
            v.kind = JVal::OBJ;
            ws();
            if (p < end && *p == '}') {
                p++;
                return v;
            }
            while (true) {
                ws();
                std::string k = string();
                ws();
                if (p >= end || *p != ':') fail("expected ':'");
                p++;
                JVal item = value(depth + 1);
                bool replaced = false;  // later duplicates win, as in nlohmann
                for (auto &kv : v.obj)
                    if (kv.first == k) {
                        kv.second = item;
                        replaced = true;
                    }
                if (!replaced) v.obj.emplace_back(k, item);
                ws();
                if (p < end && *p == ',') {
                    p++;
                    continue;
                }
                if (p < end && *p == '}') {
                    p++;
                    break;
                }
                fail("expected ',' or '}'");
            }
        } else if (c == '[') {
            p++;
            v.kind = JVal::ARR;
            ws();
            if (p < end && *p == ']') {
                p++;
                return v;
            }
            while (true) {
                v.arr.push_back(value(depth + 1));
                ws();
                if (p < end && *p == ',') {
                    p++;
                    continue;
                }
                if (p < end && *p == ']') {
                    p++;
                    break;
                }
                fail("expected ',' or ']'");
            }
        } else if (c == '"') {
            v.kind = JVal::STR;
            v.str = string();
        } else if (lit("true")) {
            v.kind = JVal::BOOL;
            v.b = true;
        } else if (lit("false")) {
            v.kind = JVal::BOOL;
        } else if (lit("null")) {
            v.kind = JVal::NUL;
        } else {
            const char *s = p;
            if (p < end && (*p == '-' || *p == '+')) p++;
            while (p < end && (std::isdigit((unsigned char)*p) || *p == '.' || *p == 'e' || *p == 'E' || *p == '-' ||
                               *p == '+'))
                p++;
            if (p == s) fail("unexpected character");
            v.kind = JVal::NUM;
            v.num = std::strtod(std::string(s, p).c_str(), nullptr);
        }
        return v;
    }
};

// --------------------------------------------------------- Pparam parse
// Pparam(const string&) (lcurve.h:169-176)
lc_pparam parse_pparam(const std::string &entry) {
    std::istringstream istr(entry);
    double value = 0., range = 0., dstep = 0.;
    bool vary = false, defined = false;
    istr >> value >> range >> dstep >> vary;
    if (!istr) throw std::runtime_error("Pparam: could not read entry = " + entry);
    istr >> defined;
    if (!istr) defined = true;
    lc_pparam pp;
    pp.value = value;
    pp.range = range;
    pp.dstep = dstep;
    pp.vary = vary ? 1 : 0;
    pp.defined = defined ? 1 : 0;
    return pp;
}

// Subs::string_to_bool (new_subs.h:48-57)
bool string_to_bool(const std::string &entry) {
    std::string t = entry;
    for (auto &c : t) c = (char)std::toupper((unsigned char)c);
    if (t == "T" || t == "TRUE" || t == "Y" || t == "YES" || t == "1") return true;
    if (t == "F" || t == "FALSE" || t == "N" || t == "NO" || t == "0") return false;
    throw std::runtime_error("to_stringing_to_bool: could not translate entry = " + entry + "to a bool");
}

// ----------------------------------------------------- parameter tables
struct PDesc {
    const char *name;
    size_t off;     // offsetof(lc_model, field)
    int group;      // 0 always, 1 use_radii, 2 !use_radii, 3 add_disc, 4 add_spot, 5 spot (needs defined)
    int lo_kind;    // 0 none, 1 v < lo illegal, 2 v <= lo illegal
    double lo;
    int hi_kind;    // 0 none, 1 v > hi illegal, 2 v >= hi illegal, 3 r1: >= xl11, 4 r2: > 1-xl12
    double hi;
    double lim_lo, lim_hi;  // get_limit
};

const double INF = std::numeric_limits<double>::infinity();
#define OFF(f) offsetof(lc_model, f)

// Reference ordering of the variable parameters (model.cpp:363-461) with
// the legality rules of is_not_legal (model.cpp:788-1158) and the limits of
// get_limit (model.cpp:1286-1410).
const PDesc PTAB[] = {
    {"q", OFF(q), 0, 2, 0., 1, 100., 0.01, 100},
    {"iangle", OFF(iangle), 0, 1, 0., 1, 90., 0, 90},
    {"r1", OFF(r1), 1, 2, 0., 3, 0., 0, INF},
    {"r2", OFF(r2), 1, 2, 0., 4, 0., 0, INF},
    {"cphi3", OFF(cphi3), 2, 2, 0., 1, 0.25, 0, INF},
    {"cphi4", OFF(cphi4), 2, 2, 0., 1, 0.25, 0, INF},
    {"spin1", OFF(spin1), 0, 2, 0., 1, 1000., 0, INF},
    {"spin2", OFF(spin2), 0, 2, 0., 1, 1000., 0, INF},
    {"t1", OFF(t1), 0, 2, 500., 1, 1.e6, 0, INF},
    {"t2", OFF(t2), 0, 2, 500., 1, 1.e6, 0, INF},
    {"ldc1_1", OFF(ldc1_1), 0, 1, -1., 1, 1., 0, INF},
    {"ldc1_2", OFF(ldc1_2), 0, 1, -1., 1, 1., 0, INF},
    {"ldc1_3", OFF(ldc1_3), 0, 1, -1., 1, 1., 0, INF},
    {"ldc1_4", OFF(ldc1_4), 0, 1, -1., 1, 1., 0, INF},
    {"ldc2_1", OFF(ldc2_1), 0, 1, -1., 1, 1., 0, INF},
    {"ldc2_2", OFF(ldc2_2), 0, 1, -1., 1, 1., 0, INF},
    {"ldc2_3", OFF(ldc2_3), 0, 1, -1., 1, 1., 0, INF},
    {"ldc2_4", OFF(ldc2_4), 0, 1, -1., 1, 1., 0, INF},
    {"velocity_scale", OFF(velocity_scale), 0, 0, 0., 0, 0., 0, INF},
    {"beam_factor1", OFF(beam_factor1), 0, 0, 0., 0, 0., 0, INF},
    {"beam_factor2", OFF(beam_factor2), 0, 0, 0., 0, 0., 0, INF},
    {"t0", OFF(t0), 0, 0, 0., 0, 0., -INF, INF},
    {"period", OFF(period), 0, 2, 1.e-3, 1, 100., 0, INF},
    {"pdot", OFF(pdot), 0, 0, 0., 0, 0., -INF, INF},
    {"deltat", OFF(deltat), 0, 2, -1., 1, 1., 0, INF},
    {"gravity_dark1", OFF(gravity_dark1), 0, 1, -1., 1, 1., 0, INF},
    {"gravity_dark2", OFF(gravity_dark2), 0, 1, -1., 1, 1., 0, INF},
    {"absorb", OFF(absorb), 0, 1, 0., 1, 10., 0, 1},
    {"slope", OFF(slope), 0, 1, -2., 1, 2., 0, INF},
    {"quad", OFF(quad), 0, 1, -2., 1, 2., 0, INF},
    {"cube", OFF(cube), 0, 1, -2., 1, 2., 0, INF},
    {"third", OFF(third), 0, 0, 0., 0, 0., 0, INF},
    {"rdisc1", OFF(rdisc1), 3, 1, 0., 1, 1., 0, INF},
    {"rdisc2", OFF(rdisc2), 3, 1, 0., 1, 1., 0, INF},
    {"height_disc", OFF(height_disc), 3, 1, 0., 0, 0., 0, INF},
    {"beta_disc", OFF(beta_disc), 3, 1, 1., 1, 100., 0, INF},
    {"temp_disc", OFF(temp_disc), 3, 1, 500., 1, 1.e6, 0, INF},
    {"texp_disc", OFF(texp_disc), 3, 1, -100., 1, 100., 0, INF},
    {"lin_limb_disc", OFF(lin_limb_disc), 3, 1, 0., 1, 1., 0, INF},
    {"quad_limb_disc", OFF(quad_limb_disc), 3, 1, 0., 1, 1., 0, INF},
    {"temp_edge", OFF(temp_edge), 3, 2, 0., 1, 1.e6, 0, INF},
    {"absorb_edge", OFF(absorb_edge), 3, 1, 0., 1, 1000., 0, INF},
    {"radius_spot", OFF(radius_spot), 4, 1, 0., 1, 1., 0, INF},
    {"length_spot", OFF(length_spot), 4, 1, 0., 1, 1., 0, INF},
    {"height_spot", OFF(height_spot), 4, 1, 0., 1, 1., 0, INF},
    {"expon_spot", OFF(expon_spot), 4, 2, 0., 1, 100., 0, INF},
    {"epow_spot", OFF(epow_spot), 4, 2, 0., 1, 10., 0, INF},
    {"angle_spot", OFF(angle_spot), 4, 1, -1000., 1, 1000., 0, INF},
    {"yaw_spot", OFF(yaw_spot), 4, 1, -1000., 1, 1000., 0, INF},
    {"temp_spot", OFF(temp_spot), 4, 1, 0., 1, 1.e10, 0, INF},
    {"tilt_spot", OFF(tilt_spot), 4, 1, -1000., 1, 1000., 0, INF},
    {"cfrac_spot", OFF(cfrac_spot), 4, 1, 0., 1, 1., 0, INF},
    {"stsp11_long", OFF(stsp11_long), 5, 1, -400., 1, 400., 0, INF},
    {"stsp11_lat", OFF(stsp11_lat), 5, 1, -90., 1, 90., 0, INF},
    {"stsp11_fwhm", OFF(stsp11_fwhm), 5, 2, 0., 1, 180., 0, INF},
    {"stsp11_tcen", OFF(stsp11_tcen), 5, 2, 0., 0, 0., 0, INF},
    {"stsp12_long", OFF(stsp12_long), 5, 1, -400., 1, 400., 0, INF},
    {"stsp12_lat", OFF(stsp12_lat), 5, 1, -90., 1, 90., 0, INF},
    {"stsp12_fwhm", OFF(stsp12_fwhm), 5, 2, 0., 1, 180., 0, INF},
    {"stsp12_tcen", OFF(stsp12_tcen), 5, 2, 0., 0, 0., 0, INF},
    {"stsp13_long", OFF(stsp13_long), 5, 1, -400., 1, 400., 0, INF},
    {"stsp13_lat", OFF(stsp13_lat), 5, 1, -90., 1, 90., 0, INF},
    {"stsp13_fwhm", OFF(stsp13_fwhm), 5, 2, 0., 1, 180., 0, INF},
    {"stsp13_tcen", OFF(stsp13_tcen), 5, 2, 0., 0, 0., 0, INF},
    {"stsp21_long", OFF(stsp21_long), 5, 1, -400., 1, 400., 0, INF},
    {"stsp21_lat", OFF(stsp21_lat), 5, 1, -90., 1, 90., 0, INF},
    {"stsp21_fwhm", OFF(stsp21_fwhm), 5, 2, 0., 1, 180., 0, INF},
    {"stsp21_tcen", OFF(stsp21_tcen), 5, 2, 0., 0, 0., 0, INF},
    {"stsp22_long", OFF(stsp22_long), 5, 1, -400., 1, 400., 0, INF},
    {"stsp22_lat", OFF(stsp22_lat), 5, 1, -90., 1, 90., 0, INF},
    {"stsp22_fwhm", OFF(stsp22_fwhm), 5, 2, 0., 1, 180., 0, INF},
    {"stsp22_tcen", OFF(stsp22_tcen), 5, 2, 0., 0, 0., 0, INF},
    {"uesp_long1", OFF(uesp_long1), 5, 1, -360., 1, 360., 0, INF},
    {"uesp_long2", OFF(uesp_long2), 5, 1, -360., 1, 720., 0, INF},
    {"uesp_lathw", OFF(uesp_lathw), 5, 2, 0., 1, 80., 0, INF},
    {"uesp_taper", OFF(uesp_taper), 5, 2, 0., 0, 0., 0, INF},
    {"uesp_temp", OFF(uesp_temp), 5, 2, 0., 0, 0., 0, INF},
};
const size_t NPTAB = sizeof(PTAB) / sizeof(PTAB[0]);

inline const lc_pparam &pp_at(const lc_model *m, size_t off) {
    return *reinterpret_cast<const lc_pparam *>(reinterpret_cast<const char *>(m) + off);
}
inline lc_pparam &pp_at(lc_model *m, size_t off) {
    return *reinterpret_cast<lc_pparam *>(reinterpret_cast<char *>(m) + off);
}

bool active(const lc_model *m, const PDesc &d) {
    const lc_pparam &p = pp_at(m, d.off);
    switch (d.group) {
    case 1: if (!m->use_radii) return false; break;
    case 2: if (m->use_radii) return false; break;
    case 3: if (!m->add_disc) return false; break;
    case 4: if (!m->add_spot) return false; break;
    case 5: return p.defined && p.vary;
    default: break;
    }
    return p.vary != 0;
}

// Names every model JSON must carry (model.cpp:14-135); star spots optional.
const char *REQUIRED_COMP[] = {"delta_phase", "nlat1f", "nlat2f", "nlat1c", "nlat2c", "npole", "nlatfill",
                               "nlngfill", "lfudge", "llo", "lhi", "phase1", "phase2", "wavelength", "roche1",
                               "roche2", "eclipse1", "eclipse2", "glens1", "use_radii", "tperiod", "gdark_bolom1",
                               "gdark_bolom2", "mucrit1", "mucrit2", "limb1", "limb2", "mirror", "add_disc", "nrad",
                               "opaque", "add_spot", "nspot", "iscale"};

void model_from_json(const char *text, size_t len, lc_model *out) {
    JParser jp{text, text + len};
    JVal root = jp.value(0);
    jp.ws();
    if (jp.p != jp.end) throw std::runtime_error("JSON parse error: trailing characters");
    const JVal *mp = nullptr;
    if (root.kind == JVal::OBJ)
        for (auto &kv : root.obj)
            if (kv.first == "model_parameters") mp = &kv.second;
    if (!mp || mp->kind != JVal::OBJ) throw std::runtime_error("Missing or invalid 'model_parameters'");

    std::map<std::string, std::string> pv;
    for (auto &kv : mp->obj) {
        if (kv.second.kind != JVal::STR)
            throw std::runtime_error("Malformed model parameter value for parameter '" + kv.first + "'");
        pv[kv.first] = kv.second.str;
    }
    auto has = [&](const char *k) { return pv.count(k) != 0; };

    std::string missing;
    for (size_t i = 0; i < NPTAB; i++) {
        const char *n = PTAB[i].name;
        if (has(n) || PTAB[i].group == 5) continue;
        if (!std::strcmp(n, "pdot") || !std::strcmp(n, "third") || !std::strcmp(n, "temp_edge") ||
            !std::strcmp(n, "absorb_edge")) {
            std::fprintf(stderr, "%s was not defined; will be set = 0\n", n);
            pv[n] = "0 1.e-10 1.e-10 0 0";
        } else {
            missing += std::string(missing.empty() ? "" : ", ") + n;
        }
    }
    for (const char *n : REQUIRED_COMP)
        if (!has(n)) missing += std::string(missing.empty() ? "" : ", ") + n;
    if (!missing.empty()) throw std::runtime_error("One or more parameters were not initialised: " + missing);

    lc_model m;
    std::memset(&m, 0, sizeof(m));
    for (size_t i = 0; i < NPTAB; i++) {
        if (PTAB[i].group == 5) continue;
        pp_at(&m, PTAB[i].off) = parse_pparam(pv[PTAB[i].name]);
    }
    // star-spot groups: all or none (model.cpp:232-301)
    static const char *groups[6][5] = {
        {"stsp11_long", "stsp11_lat", "stsp11_fwhm", "stsp11_tcen", nullptr},
        {"stsp12_long", "stsp12_lat", "stsp12_fwhm", "stsp12_tcen", nullptr},
        {"stsp13_long", "stsp13_lat", "stsp13_fwhm", "stsp13_tcen", nullptr},
        {"stsp21_long", "stsp21_lat", "stsp21_fwhm", "stsp21_tcen", nullptr},
        {"stsp22_long", "stsp22_lat", "stsp22_fwhm", "stsp22_tcen", nullptr},
        {"uesp_long1", "uesp_long2", "uesp_lathw", "uesp_taper", "uesp_temp"}};
    for (int g = 0; g < 6; g++) {
        bool any = false, all = true;
        for (int k = 0; k < 5 && groups[g][k]; k++) {
            any = any || has(groups[g][k]);
            all = all && has(groups[g][k]);
        }
        if (!any) continue;
        if (!all) {
            if (g < 5)
                throw std::runtime_error(std::string("One or more of the star spot ") + (groups[g][0] + 4) +
                                         " parameters were not initialised");
            throw std::runtime_error("One or more of uniform equatorial spot parameters were not initialised");
        }
        for (int k = 0; k < 5 && groups[g][k]; k++)
            for (size_t i = 0; i < NPTAB; i++)
                if (!std::strcmp(PTAB[i].name, groups[g][k])) pp_at(&m, PTAB[i].off) = parse_pparam(pv[groups[g][k]]);
    }

    m.delta_phase = std::stod(pv["delta_phase"]);
    m.nlat1f = std::stoi(pv["nlat1f"]);
    m.nlat2f = std::stoi(pv["nlat2f"]);
    m.nlat1c = std::stoi(pv["nlat1c"]);
    m.nlat2c = std::stoi(pv["nlat2c"]);
    m.npole = string_to_bool(pv["npole"]);
    m.nlatfill = std::stoi(pv["nlatfill"]);
    m.nlngfill = std::stoi(pv["nlngfill"]);
    m.lfudge = std::stod(pv["lfudge"]);
    m.llo = std::stod(pv["llo"]);
    m.lhi = std::stod(pv["lhi"]);
    m.phase1 = std::stod(pv["phase1"]);
    m.phase2 = std::stod(pv["phase2"]);
    m.wavelength = std::stod(pv["wavelength"]);
    m.roche1 = string_to_bool(pv["roche1"]);
    m.roche2 = string_to_bool(pv["roche2"]);
    m.eclipse1 = string_to_bool(pv["eclipse1"]);
    m.eclipse2 = string_to_bool(pv["eclipse2"]);
    m.glens1 = string_to_bool(pv["glens1"]);
    if (m.glens1 && m.roche1)
        throw std::runtime_error("For reasons of simplicity, glens1 = 1 and roche1 = 1 are not simultaneously allowed");
    m.use_radii = string_to_bool(pv["use_radii"]);
    m.tperiod = std::stod(pv["tperiod"]);
    m.gdark_bolom1 = string_to_bool(pv["gdark_bolom1"]);
    m.gdark_bolom2 = string_to_bool(pv["gdark_bolom2"]);
    m.mucrit1 = std::stod(pv["mucrit1"]);
    m.mucrit2 = std::stod(pv["mucrit2"]);
    for (int s = 0; s < 2; s++) {
        const std::string &v = pv[s == 0 ? "limb1" : "limb2"];
        int t;
        if (v == "Poly")
            t = LC_LDC_POLY;
        else if (v == "Claret")
            t = LC_LDC_CLARET;
        else
            throw std::runtime_error(std::string("Could not recognize the value of ") + (s == 0 ? "limb1" : "limb2") +
                                     " = " + v + "; should be 'Poly' or 'Claret'");
        (s == 0 ? m.limb1 : m.limb2) = t;
    }
    m.mirror = string_to_bool(pv["mirror"]);
    m.add_disc = string_to_bool(pv["add_disc"]);
    m.nrad = std::stoi(pv["nrad"]);
    m.opaque = string_to_bool(pv["opaque"]);
    m.add_spot = string_to_bool(pv["add_spot"]);
    m.nspot = std::stoi(pv["nspot"]);
    m.iscale = string_to_bool(pv["iscale"]);
    *out = m;
}

thread_local std::string g_thread_err;

}  // namespace

namespace lch {
void set_thread_error(const std::string &s) { g_thread_err = s; }
}  // namespace lch

extern "C" {

const char *lc_thread_error(void) { return g_thread_err.c_str(); }

size_t lc_sizeof(const char *name) {
    if (!name) return 0;
    struct {
        const char *n;
        size_t s;
    } t[] = {{"lc_pparam", sizeof(lc_pparam)},       {"lc_model", sizeof(lc_model)},
             {"lc_datum", sizeof(lc_datum)},         {"lc_scalars", sizeof(lc_scalars)},
             {"lc_eval_opts", sizeof(lc_eval_opts)}, {"lc_ldc", sizeof(lc_ldc)},
             {"lc_flat_grid", sizeof(lc_flat_grid)}, {"lc_phase_batch", sizeof(lc_phase_batch)},
             {"lc_star_faces", sizeof(lc_star_faces)}, {"lc_profile", sizeof(lc_profile)}};
    for (auto &e : t)
        if (!std::strcmp(e.n, name)) return e.s;
    return 0;
}

lc_status lc_model_from_json(const char *json_text, size_t len, lc_model *out) {
    if (!json_text || !out) {
        g_thread_err = "lc_model_from_json: null argument";
        return LC_E_ARG;
    }
    try {
        model_from_json(json_text, len, out);
        return LC_OK;
    } catch (const std::exception &e) {
        g_thread_err = e.what();
        return LC_E_PARAM;
    }
}

int lc_model_nvary(const lc_model *m) {
    int n = 0;
    for (size_t i = 0; i < NPTAB; i++) n += active(m, PTAB[i]) ? 1 : 0;
    return n;
}

lc_status lc_model_set_param(lc_model *m, const double *vpar, size_t n) {
    if ((size_t)lc_model_nvary(m) > n) {
        g_thread_err = "set_param: parameter vector too short";
        return LC_E_ARG;
    }
    size_t k = 0;
    for (size_t i = 0; i < NPTAB; i++)
        if (active(m, PTAB[i])) pp_at(m, PTAB[i].off).value = vpar[k++];
    return LC_OK;
}

static lc_status get_field(const lc_model *m, double *v, size_t n, int which) {
    if ((size_t)lc_model_nvary(m) > n) {
        g_thread_err = "parameter vector too short";
        return LC_E_ARG;
    }
    size_t k = 0;
    for (size_t i = 0; i < NPTAB; i++) {
        if (!active(m, PTAB[i])) continue;
        const lc_pparam &p = pp_at(m, PTAB[i].off);
        v[k++] = which == 0 ? p.value : (which == 1 ? p.range : p.dstep);
    }
    return LC_OK;
}

lc_status lc_model_get_param(const lc_model *m, double *vpar, size_t n) { return get_field(m, vpar, n, 0); }
lc_status lc_model_get_range(const lc_model *m, double *v, size_t n) { return get_field(m, v, n, 1); }
lc_status lc_model_get_dstep(const lc_model *m, double *v, size_t n) { return get_field(m, v, n, 2); }

lc_status lc_model_get_limit(const lc_model *m, double *lo, double *hi, size_t n) {
    if ((size_t)lc_model_nvary(m) > n) {
        g_thread_err = "get_limit: output too short";
        return LC_E_ARG;
    }
    size_t k = 0;
    for (size_t i = 0; i < NPTAB; i++) {
        if (!active(m, PTAB[i])) continue;
        lo[k] = PTAB[i].lim_lo;
        hi[k] = PTAB[i].lim_hi;
        k++;
    }
    return LC_OK;
}

int lc_model_is_not_legal(const lc_model *m, const double *vpar, size_t n) {
    if ((size_t)lc_model_nvary(m) > n) return 1;
    int st = 0;
    // Both computed from the current (not the proposed) q (model.cpp:801-802)
    double xl11 = lcx_host_l1(1, m->q.value, m->spin1.value, &st);
    double xl12 = 1. - lcx_host_l1(2, m->q.value, m->spin2.value, &st);
    size_t k = 0;
    for (size_t i = 0; i < NPTAB; i++) {
        const PDesc &d = PTAB[i];
        if (!active(m, d)) continue;
        double v = vpar[k++];
        if ((d.lo_kind == 1 && v < d.lo) || (d.lo_kind == 2 && v <= d.lo)) return 1;
        if ((d.hi_kind == 1 && v > d.hi) || (d.hi_kind == 2 && v >= d.hi)) return 1;
        if (d.hi_kind == 3 && v >= xl11) return 1;
        if (d.hi_kind == 4 && v > xl12) return 1;
    }
    return 0;
}

lc_status lc_model_is_not_legal_batch(const lc_model *m, const double *vpar, size_t nvec, size_t n, int *out) {
    if (!m || (!vpar && nvec) || (!out && nvec)) return LC_E_ARG;
    const size_t nv = (size_t)lc_model_nvary(m);
    if (nv > n) {
        for (size_t k = 0; k < nvec; k++) out[k] = 1;
        return LC_OK;
    }
    int st = 0;
    // the limits depend on the current model only: computed once
    const double xl11 = lcx_host_l1(1, m->q.value, m->spin1.value, &st);
    const double xl12 = 1. - lcx_host_l1(2, m->q.value, m->spin2.value, &st);
    for (size_t r = 0; r < nvec; r++) {
        const double *vp = vpar + r * n;
        int bad = 0;
        size_t k = 0;
        for (size_t i = 0; i < NPTAB && !bad; i++) {
            const PDesc &d = PTAB[i];
            if (!active(m, d)) continue;
            const double v = vp[k++];
            bad = (d.lo_kind == 1 && v < d.lo) || (d.lo_kind == 2 && v <= d.lo) || (d.hi_kind == 1 && v > d.hi) ||
                  (d.hi_kind == 2 && v >= d.hi) || (d.hi_kind == 3 && v >= xl11) || (d.hi_kind == 4 && v > xl12);
        }
        out[r] = bad;
    }
    return LC_OK;
}

lc_status lc_model_get_name(const lc_model *m, int i, char *buf, size_t buflen) {
    if (!buf || buflen == 0) return LC_E_ARG;
    int k = 0;
    const char *name = "UNKNOWN";
    for (size_t j = 0; j < NPTAB; j++) {
        if (!active(m, PTAB[j])) continue;
        if (k == i) {
            name = PTAB[j].name;
            break;
        }
        k++;
    }
    std::snprintf(buf, buflen, "%s", name);
    return LC_OK;
}

}  // extern "C"
