// lcurve_re/lcurve.h -- C++ drop-in for the light-curve path of Fabmat1/lcurve_re.
//
// Header-only layer over the C ABI of include/lcurve_amd.h.  It keeps the
// reference's names, argument meanings and error behaviour for the path a
// fitting program calls:
//
//   Lcurve::Lcurve_Error                      lcurve.h:34-43 (runtime_error)
//   Lcurve::Model(const json&)                lcurve.h:319-330, model.cpp:10-360
//   Model::nvary/set_param/is_not_legal/
//         get_param/get_range/get_dstep/
//         get_limit/get_r1r2                  lcurve.h:321-412
//   Lcurve::Datum, Lcurve::Data (+ rasc,
//         wrasc, Data(file), >> and <<)       lcurve.h:427-472, lcurve.cpp:14-100
//   Lcurve::light_curve_comp                  lcurve.h:669-674
//   Lcurve::light_curve_comp_fast             lcurve.h:474-483
//
// plus Lcurve::light_curve_comp_batch, which evaluates many parameter
// vectors against one data set in one launch sequence.
//
// Differences a caller can see:
//   * Model(const json&) accepts any object with a dump() method returning
//     the JSON text (nlohmann::json has one), or the text itself, so this
//     header does not depend on a JSON library;
//   * set_param / is_not_legal take any container with size() and
//     operator[] (Subs::Array1D<double>, std::vector<double>); the getters
//     return std::vector<double>;
//   * Data::rasc does not print the number of lines read;
//   * every evaluation runs on the GPU through a thread_local engine context
//     (device from LCURVE_AMD_DEVICE, default 0).  There is no CPU path: a
//     missing GPU or library is an Lcurve_Error;
//   * light_curve_comp / _fast compute each face in FP64 (the reference's
//     CPU arithmetic, comp_light_flat.cpp:81-168) unless the caller asks for
//     FP32 per face with FP64 accumulation, the arithmetic of the
//     reference's CUDA flux path: Lcurve::set_precision(LC_PREC_MIXED) for
//     the calling thread, or the environment variable LCURVE_AMD_PRECISION
//     (or the reference's LCURVE_CUDA_PRECISION, comp_light_cuda.cu:523-533)
//     set to float / fp32 / 32 / mixed.  double / fp64 / 64 select FP64.
#ifndef LCURVE_RE_LCURVE_H
#define LCURVE_RE_LCURVE_H

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../lcurve_amd.h"

namespace Lcurve {

// ------------------------------------------------------------------ errors
class Lcurve_Error : public std::runtime_error {
public:
    Lcurve_Error() : std::runtime_error("Lcurve error") {}
    explicit Lcurve_Error(const std::string &err) : std::runtime_error(err) {}
};

namespace detail {

[[noreturn]] inline void fail(const char *where, const char *msg) {
    throw Lcurve_Error(std::string(where) + ": " + (msg ? msg : "unknown error"));
}

// One engine context per host thread, created on first use (the reference
// keeps its grid cache and CUDA contexts thread_local the same way).
struct ThreadCtx {
    lc_ctx *ctx = nullptr;
    ~ThreadCtx() {
        if (ctx) lc_ctx_destroy(ctx);
    }
};

inline lc_ctx *thread_ctx() {
    thread_local ThreadCtx holder;
    if (!holder.ctx) {
        int dev = 0;
        if (const char *d = std::getenv("LCURVE_AMD_DEVICE")) dev = std::atoi(d);
        if (lc_ctx_create(dev, &holder.ctx) != LC_OK) {
            holder.ctx = nullptr;
            fail("Lcurve::light_curve_comp", lc_thread_error());
        }
    }
    return holder.ctx;
}

// Precision of the per-face arithmetic for light_curve_comp(_fast):
// LC_PREC_FP64 unless LCURVE_AMD_PRECISION, else LCURVE_CUDA_PRECISION,
// names FP32 (case-insensitive).  Read once per process.
inline int env_precision() {
    static const int value = [] {
        const char *env = std::getenv("LCURVE_AMD_PRECISION");
        if (!env) env = std::getenv("LCURVE_CUDA_PRECISION");
        if (!env) return int(LC_PREC_FP64);
        std::string mode(env);
        for (char &c : mode) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
        if (mode == "float" || mode == "fp32" || mode == "32" || mode == "single" || mode == "mixed")
            return int(LC_PREC_MIXED);
        return int(LC_PREC_FP64);
    }();
    return value;
}

// Per-thread choice made with Lcurve::set_precision (-1: none made).
inline int &thread_precision() {
    thread_local int p = -1;
    return p;
}

inline int precision() { return thread_precision() >= 0 ? thread_precision() : env_precision(); }

template <class Vec> std::vector<double> to_vector(const Vec &v) {
    std::vector<double> out(static_cast<size_t>(v.size()));
    for (size_t i = 0; i < out.size(); i++) out[i] = v[static_cast<int>(i)];
    return out;
}

// std::void_t stand-in for "has a dump() method"
template <class T, class = void> struct has_dump : std::false_type {};
template <class T>
struct has_dump<T, decltype((void)std::declval<const T &>().dump(), void())> : std::true_type {};

}  // namespace detail

// ------------------------------------------------------------------ Model
// Lcurve::Model: the fields are those of the reference (q, iangle, ...,
// delta_phase, nlat1f, ...), each physical one an lc_pparam with value,
// range, dstep, vary and defined.
struct Model : lc_model {
    // From the JSON text of a configuration with a "model_parameters" object.
    explicit Model(const std::string &json_text) : lc_model() {
        if (lc_model_from_json(json_text.data(), json_text.size(), this) != LC_OK)
            throw Lcurve_Error(lc_thread_error());
    }
    explicit Model(const char *json_text) : Model(std::string(json_text)) {}
    // From a JSON object (nlohmann::json or anything with dump()).
    template <class Json, typename std::enable_if<detail::has_dump<Json>::value, int>::type = 0>
    explicit Model(const Json &config) : Model(std::string(config.dump())) {}
    explicit Model(const lc_model &pod) : lc_model(pod) {}

    int nvary() const { return lc_model_nvary(this); }

    template <class Vec> void set_param(const Vec &vpar) {
        const std::vector<double> v = detail::to_vector(vpar);
        if (lc_model_set_param(this, v.data(), v.size()) != LC_OK) throw Lcurve_Error(lc_thread_error());
    }

    template <class Vec> bool is_not_legal(const Vec &vpar) const {
        const std::vector<double> v = detail::to_vector(vpar);
        return lc_model_is_not_legal(this, v.data(), v.size()) != 0;
    }

    std::vector<double> get_param() const { return get(lc_model_get_param); }
    std::vector<double> get_range() const { return get(lc_model_get_range); }
    std::vector<double> get_dstep() const { return get(lc_model_get_dstep); }

    std::vector<std::pair<double, double>> get_limit() const {
        const size_t n = static_cast<size_t>(nvary());
        std::vector<double> lo(n ? n : 1), hi(n ? n : 1);
        if (lc_model_get_limit(this, lo.data(), hi.data(), n) != LC_OK) throw Lcurve_Error(lc_thread_error());
        std::vector<std::pair<double, double>> out(n);
        for (size_t i = 0; i < n; i++) out[i] = {lo[i], hi[i]};
        return out;
    }

    // Name of the i-th variable parameter (model.cpp:562).
    std::string get_name(int i) const {
        char buf[64];
        if (lc_model_get_name(this, i, buf, sizeof(buf)) != LC_OK) throw Lcurve_Error(lc_thread_error());
        return buf;
    }

    // Relative radii for either parameterisation (lcurve.h:399-409).
    void get_r1r2(double &rr1, double &rr2) const {
        if (use_radii) {
            rr1 = r1.value;
            rr2 = r2.value;
        } else {
            const double pi = 3.14159265358979323846;
            const double sini = std::sin(pi * iangle.value / 180.);
            const double r2pr1 = std::sqrt(1.0 - std::pow(sini * std::cos(2.0 * pi * cphi4.value), 2));
            const double r2mr1 = std::sqrt(1.0 - std::pow(sini * std::cos(2.0 * pi * cphi3.value), 2));
            rr1 = (r2pr1 - r2mr1) / 2.0;
            rr2 = (r2pr1 + r2mr1) / 2.0;
        }
    }

    // Number of faces of one star grid (set_star_grid.cpp:110-162).
    size_t grid_size(int star, bool fine) const {
        size_t n = 0;
        if (lc_star_grid_size(this, star, fine ? 1 : 0, &n) != LC_OK) throw Lcurve_Error(lc_thread_error());
        return n;
    }

    // Model::wrasc (model.cpp:1812-1818): the parameter listing of
    // operator<< written to a file.
    inline void wrasc(const std::string &file) const;

private:
    std::vector<double> get(lc_status (*fn)(const lc_model *, double *, size_t)) const {
        const size_t n = static_cast<size_t>(nvary());
        std::vector<double> out(n ? n : 1);
        if (fn(this, out.data(), n) != LC_OK) throw Lcurve_Error(lc_thread_error());
        out.resize(n);
        return out;
    }
};

// Pparam output (model.cpp:1664-1670): value, range and dstep in fixed
// notation, then vary and defined.  The stream keeps the last precision
// and the fixed flag, as the reference's does.
inline std::ostream &print_pparam(std::ostream &s, const lc_pparam &p) {
    s << std::setw(17) << std::setprecision(10) << std::fixed << p.value << " " << std::setw(8)
      << std::setprecision(4) << std::fixed << p.range << " " << std::setw(8) << std::setprecision(4) << std::fixed
      << p.dstep << " " << (p.vary ? 1 : 0) << " " << (p.defined ? 1 : 0);
    return s;
}

// Model output (model.cpp:1673-1806): one "name = value" line per field in
// the reference's order and blank-line grouping.
inline std::ostream &operator<<(std::ostream &s, const Model &m) {
    const auto P = [&s](const char *label, const lc_pparam &p) -> std::ostream & {
        s << label;
        return print_pparam(s, p);
    };
    P("q              = ", m.q) << "\n";
    P("iangle         = ", m.iangle) << "\n";
    P("r1             = ", m.r1) << "\n";
    P("r2             = ", m.r2) << "\n";
    P("cphi3          = ", m.cphi3) << "\n";
    P("cphi4          = ", m.cphi4) << "\n";
    P("spin1          = ", m.spin1) << "\n";
    P("spin2          = ", m.spin2) << "\n";
    P("t1             = ", m.t1) << "\n";
    P("t2             = ", m.t2) << "\n";
    P("ldc1_1         = ", m.ldc1_1) << "\n";
    P("ldc1_2         = ", m.ldc1_2) << "\n";
    P("ldc1_3         = ", m.ldc1_3) << "\n";
    P("ldc1_4         = ", m.ldc1_4) << "\n";
    P("ldc2_1         = ", m.ldc2_1) << "\n";
    P("ldc2_2         = ", m.ldc2_2) << "\n";
    P("ldc2_3         = ", m.ldc2_3) << "\n";
    P("ldc2_4         = ", m.ldc2_4) << "\n";
    P("velocity_scale = ", m.velocity_scale) << "\n";
    P("beam_factor1   = ", m.beam_factor1) << "\n";
    P("beam_factor2   = ", m.beam_factor2) << "\n\n";
    P("t0             = ", m.t0) << "\n";
    P("period         = ", m.period) << "\n";
    P("pdot           = ", m.pdot) << "\n";
    P("deltat         = ", m.deltat) << "\n";
    P("gravity_dark1  = ", m.gravity_dark1) << "\n";
    P("gravity_dark2  = ", m.gravity_dark2) << "\n";
    P("absorb         = ", m.absorb) << "\n";
    P("slope          = ", m.slope) << "\n\n";
    P("quad           = ", m.quad) << "\n\n";
    P("cube           = ", m.cube) << "\n\n";
    P("third          = ", m.third) << "\n\n";
    P("rdisc1         = ", m.rdisc1) << "\n";
    P("rdisc2         = ", m.rdisc2) << "\n";
    P("height_disc    = ", m.height_disc) << "\n";
    P("beta_disc      = ", m.beta_disc) << "\n";
    P("temp_disc      = ", m.temp_disc) << "\n";
    P("texp_disc      = ", m.texp_disc) << "\n";
    P("lin_limb_disc  = ", m.lin_limb_disc) << "\n";
    P("quad_limb_disc = ", m.quad_limb_disc) << "\n";
    P("temp_edge      = ", m.temp_edge) << "\n";
    P("absorb_edge    = ", m.absorb_edge) << "\n\n";
    P("radius_spot    = ", m.radius_spot) << "\n";
    P("length_spot    = ", m.length_spot) << "\n";
    P("height_spot    = ", m.height_spot) << "\n";
    P("expon_spot     = ", m.expon_spot) << "\n";
    P("epow_spot      = ", m.epow_spot) << "\n";
    P("angle_spot     = ", m.angle_spot) << "\n";
    P("yaw_spot       = ", m.yaw_spot) << "\n";
    P("temp_spot      = ", m.temp_spot) << "\n";
    P("tilt_spot      = ", m.tilt_spot) << "\n";
    P("cfrac_spot     = ", m.cfrac_spot) << "\n\n";
    const lc_pparam *sp[5][4] = {{&m.stsp11_long, &m.stsp11_lat, &m.stsp11_fwhm, &m.stsp11_tcen},
                                 {&m.stsp12_long, &m.stsp12_lat, &m.stsp12_fwhm, &m.stsp12_tcen},
                                 {&m.stsp13_long, &m.stsp13_lat, &m.stsp13_fwhm, &m.stsp13_tcen},
                                 {&m.stsp21_long, &m.stsp21_lat, &m.stsp21_fwhm, &m.stsp21_tcen},
                                 {&m.stsp22_long, &m.stsp22_lat, &m.stsp22_fwhm, &m.stsp22_tcen}};
    const char *sn[5] = {"11", "12", "13", "21", "22"};
    for (int k = 0; k < 5; k++) {
        s << "stsp" << sn[k] << "_long    = ";
        print_pparam(s, *sp[k][0]) << "\n";
        s << "stsp" << sn[k] << "_lat     = ";
        print_pparam(s, *sp[k][1]) << "\n";
        s << "stsp" << sn[k] << "_fwhm    = ";
        print_pparam(s, *sp[k][2]) << "\n";
        s << "stsp" << sn[k] << "_tcen    = ";
        print_pparam(s, *sp[k][3]) << "\n\n";
    }
    P("uesp_long1     = ", m.uesp_long1) << "\n";
    P("uesp_long2     = ", m.uesp_long2) << "\n";
    P("uesp_lathw     = ", m.uesp_lathw) << "\n";
    P("uesp_taper     = ", m.uesp_taper) << "\n";
    P("uesp_temp      = ", m.uesp_temp) << "\n\n";
    s << "delta_phase    = " << m.delta_phase << "\n";
    s << "nlat1f         = " << m.nlat1f << "\n";
    s << "nlat2f         = " << m.nlat2f << "\n";
    s << "nlat1c         = " << m.nlat1c << "\n";
    s << "nlat2c         = " << m.nlat2c << "\n";
    s << "npole          = " << (m.npole ? 1 : 0) << "\n";
    s << "nlatfill       = " << m.nlatfill << "\n";
    s << "nlngfill       = " << m.nlngfill << "\n";
    s << "lfudge         = " << m.lfudge << "\n";
    s << "llo            = " << m.llo << "\n";
    s << "lhi            = " << m.lhi << "\n";
    s << "phase1         = " << m.phase1 << "\n";
    s << "phase2         = " << m.phase2 << "\n";
    s << "wavelength     = " << m.wavelength << "\n";
    s << "roche1         = " << (m.roche1 ? 1 : 0) << "\n";
    s << "roche2         = " << (m.roche2 ? 1 : 0) << "\n";
    s << "eclipse1       = " << (m.eclipse1 ? 1 : 0) << "\n";
    s << "eclipse2       = " << (m.eclipse2 ? 1 : 0) << "\n";
    s << "glens1         = " << (m.glens1 ? 1 : 0) << "\n";
    s << "use_radii      = " << (m.use_radii ? 1 : 0) << "\n";
    s << "tperiod        = " << m.tperiod << "\n";
    s << "gdark_bolom1   = " << (m.gdark_bolom1 ? 1 : 0) << "\n";
    s << "gdark_bolom2   = " << (m.gdark_bolom2 ? 1 : 0) << "\n";
    s << "mucrit1        = " << m.mucrit1 << "\n";
    s << "mucrit2        = " << m.mucrit2 << "\n";
    s << (m.limb1 == LC_LDC_CLARET ? "limb1          = Claret\n" : "limb1          = Poly\n");
    s << (m.limb2 == LC_LDC_CLARET ? "limb2          = Claret\n" : "limb2          = Poly\n");
    s << "mirror         = " << (m.mirror ? 1 : 0) << "\n";
    s << "add_disc       = " << (m.add_disc ? 1 : 0) << "\n";
    s << "nrad           = " << m.nrad << "\n";
    s << "opaque         = " << (m.opaque ? 1 : 0) << "\n";
    s << "add_spot       = " << (m.add_spot ? 1 : 0) << "\n";
    s << "nspot          = " << m.nspot << "\n";
    s << "iscale         = " << (m.iscale ? 1 : 0) << "\n";
    return s;
}

inline void Model::wrasc(const std::string &file) const {
    std::ofstream fout(file.c_str());
    fout << *this;
}

// ------------------------------------------------------------- Datum, Data
struct Datum : lc_datum {
    Datum() : lc_datum() {}
    Datum(double t, double e, double f, double fe, double w, int nd) : lc_datum() {
        time = t;
        expose = e;
        flux = f;
        ferr = fe;
        weight = w;
        ndiv = nd;
    }
};
static_assert(sizeof(Datum) == sizeof(lc_datum), "Datum must keep the lc_datum layout");

// One line "time expose flux ferr weight ndiv"; a negative ferr gives
// ferr = 0 and weight = 0 (lcurve.cpp:18-31).
inline std::istream &operator>>(std::istream &s, Datum &d) {
    std::string buff;
    std::getline(s, buff);
    std::istringstream istr(buff);
    istr >> d.time >> d.expose >> d.flux >> d.ferr >> d.weight >> d.ndiv;
    if (!istr) throw std::runtime_error("Lcurve::operator>>: failed to read t,e,f,fe,wgt,ndiv of a Datum");
    if (d.ferr < 0.) {
        d.ferr -= d.ferr;
        d.weight = 0.;
    }
    return s;
}

// lcurve.cpp:35-44
inline std::ostream &operator<<(std::ostream &s, const Datum &d) {
    s << std::setw(17) << std::fixed << std::setprecision(8) << d.time << " " << d.expose << " " << std::setw(10)
      << std::fixed << std::setprecision(5) << d.flux << " " << d.ferr << " " << d.weight << " " << d.ndiv;
    return s;
}

class Data : public std::vector<Datum> {
public:
    Data() = default;
    explicit Data(int n) : std::vector<Datum>(static_cast<size_t>(n)) {}
    explicit Data(const std::string &file) { rasc(file); }

    // Lines starting with '#', blank, space or tab are skipped
    // (lcurve.cpp:50-85).
    void rasc(const std::string &file) {
        std::ifstream fin(file.c_str());
        if (!fin) throw std::runtime_error("Lcurve::Data::Data: failed to open " + file + " for data.");
        clear();
        int n = 0;
        Datum d;
        while (fin && !fin.eof()) {
            n++;
            const int ch = fin.peek();
            if (fin.eof()) break;
            if (ch == '#' || ch == ' ' || ch == '\t' || ch == '\n') {
                fin.ignore(5000, '\n');
            } else if (fin >> d) {
                push_back(d);
            } else if (!fin.eof()) {
                throw std::runtime_error("Data file input failure on line " + std::to_string(n));
            }
        }
    }

    void wrasc(const std::string &file) const {
        std::ofstream fout(file.c_str());
        if (!fout) throw std::runtime_error("Lcurve::Data::wrasc: failed to open " + file + " for output.");
        for (const Datum &d : *this) fout << d << std::endl;
    }

    std::vector<lc_datum> pod() const { return std::vector<lc_datum>(begin(), end()); }
};

// ------------------------------------------------------------ evaluations
namespace detail {

inline void eval(const Model &mdl, const Data &data, bool scale, bool rdata, bool info, std::vector<double> &sfac,
                 std::vector<double> &calc, double &wdwarf, double &chisq, double &wnok, double &logg1,
                 double &logg2, double &rv1, double &rv2, int max_model_points, bool diagnostics,
                 int precision, const char *where) {
    if (data.empty()) fail(where, "no data points");
    calc.resize(data.size());
    lc_scalars io{};
    if (!scale) {
        if (sfac.empty()) fail(where, "sfac must hold the scale factor(s) when scale is false");
        for (size_t k = 0; k < 5; k++) io.sfac[k] = sfac[k < sfac.size() ? k : 0];
    }
    lc_eval_opts o{};
    o.scale = scale;
    o.rdata = rdata;
    o.info = info;
    o.diagnostics = diagnostics;
    o.max_model_points = max_model_points;
    o.precision = precision;
    const std::vector<lc_datum> d = data.pod();
    lc_ctx *ctx = thread_ctx();
    if (lc_light_curve_comp(ctx, &mdl, d.data(), d.size(), &o, calc.data(), &io) != LC_OK)
        throw Lcurve_Error(lc_last_error(ctx));
    if (scale) {
        if (sfac.size() < 5) sfac.resize(5);
        for (size_t k = 0; k < 5; k++) sfac[k] = io.sfac[k];
    }
    wdwarf = io.wdwarf;
    if (scale || rdata) {
        chisq = io.chisq;
        wnok = io.wnok;
    }
    logg1 = io.logg1;
    logg2 = io.logg2;
    rv1 = io.rv1;
    rv2 = io.rv2;
}

}  // namespace detail

// Per-face precision of light_curve_comp(_fast) on the calling thread:
// LC_PREC_FP64 (the reference's CPU arithmetic) or LC_PREC_MIXED (FP32 per
// face, FP64 accumulation).  A negative value returns to the environment's
// choice (FP64 when LCURVE_AMD_PRECISION / LCURVE_CUDA_PRECISION are unset).
inline void set_precision(int precision) {
    if (precision != LC_PREC_FP64 && precision != LC_PREC_MIXED && precision >= 0)
        detail::fail("Lcurve::set_precision", "precision must be LC_PREC_FP64 or LC_PREC_MIXED");
    detail::thread_precision() = precision < 0 ? -1 : precision;
}

inline int get_precision() { return detail::precision(); }

// Lcurve::light_curve_comp (lcurve.h:669-674): the model light curve at the
// times of `data`, optionally rescaled (scale) and compared (rdata).
inline void light_curve_comp(const Model &model, const Data &data, bool scale, bool rdata, bool info,
                             std::vector<double> &sfac, std::vector<double> &calc, double &wdwarf, double &chisq,
                             double &wnok, double &logg1, double &logg2, double &rv1, double &rv2,
                             bool diagnostics = true) {
    detail::eval(model, data, scale, rdata, info, sfac, calc, wdwarf, chisq, wnok, logg1, logg2, rv1, rv2, 0,
                 diagnostics, detail::precision(), "Lcurve::light_curve_comp");
}

// Lcurve::light_curve_comp_fast (lcurve.h:474-483): evaluates at most
// max_model_points evenly spaced points and interpolates in time.
inline void light_curve_comp_fast(const Model &mdl, const Data &data, bool scale, bool rdata, bool info,
                                  std::vector<double> &sfac, std::vector<double> &calc, double &wdwarf,
                                  double &chisq, double &wnok, double &logg1, double &logg2, double &rv1,
                                  double &rv2, int max_model_points = 500, bool diagnostics = true) {
    detail::eval(mdl, data, scale, rdata, info, sfac, calc, wdwarf, chisq, wnok, logg1, logg2, rv1, rv2,
                 max_model_points, diagnostics, detail::precision(), "Lcurve::light_curve_comp_fast");
}

// Results of light_curve_comp_batch: calc is models.size() x data.size(),
// row i belongs to models[i]; status[i] is LC_OK or the lc_status of the
// error that instance i hit (its row is then zero).
struct BatchResult {
    size_t n = 0;
    std::vector<double> calc;
    std::vector<lc_scalars> scalars;
    std::vector<int32_t> status;
    const double *row(size_t i) const { return calc.data() + i * n; }
};

// Many parameter vectors (MCMC proposals, Jacobian columns) against one
// data set in one launch sequence.  Instances fail individually (status);
// an Lcurve_Error is thrown only when the batch itself could not run.
// precision < 0 takes the thread's choice (get_precision()).
inline BatchResult light_curve_comp_batch(const std::vector<Model> &models, const Data &data, bool scale, bool rdata,
                                          double sfac = 1.0, int max_model_points = 0, bool diagnostics = false,
                                          int precision = -1) {
    if (precision < 0) precision = detail::precision();
    if (data.empty()) detail::fail("Lcurve::light_curve_comp_batch", "no data points");
    BatchResult r;
    r.n = data.size();
    const size_t m = models.size();
    r.calc.assign(m * r.n, 0.0);
    r.scalars.assign(m, lc_scalars{});
    r.status.assign(m, 0);
    for (lc_scalars &s : r.scalars)
        for (double &f : s.sfac) f = sfac;
    std::vector<lc_model> pods(models.begin(), models.end());
    const std::vector<lc_datum> d = data.pod();
    lc_eval_opts o{};
    o.scale = scale;
    o.rdata = rdata;
    o.diagnostics = diagnostics;
    o.max_model_points = max_model_points;
    o.precision = precision;
    lc_ctx *ctx = detail::thread_ctx();
    if (lc_light_curve_comp_batch(ctx, pods.data(), m, d.data(), d.size(), &o, r.calc.data(), r.scalars.data(),
                                  r.status.data()) != LC_OK)
        throw Lcurve_Error(lc_last_error(ctx));
    return r;
}

}  // namespace Lcurve

#endif  // LCURVE_RE_LCURVE_H
