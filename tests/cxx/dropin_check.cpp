// Exercises the C++ drop-in header include/lcurve_re/lcurve.h the way a
// fitting program of the reference uses Lcurve::Model / Lcurve::Data /
// Lcurve::light_curve_comp.  Driven by tests/test_cxx_dropin.py.
//
//   dropin_check model <config.json>            parameter API, one value per line
//   dropin_check data  <in.dat> <out.dat>       Data::rasc then Data::wrasc
//   dropin_check eval  <config.json> <in.dat> <out.txt>
//                                               light_curve_comp, _fast and _batch
//   dropin_check error <config.json> <in.dat>   expects an Lcurve_Error
//   dropin_check print <config.json> <out.txt>  Model::wrasc
//   dropin_check jac   <config.json> <in.dat> <out.txt> <j> [fp64|mixed]
//                                               central-difference column j of
//                                               light_curve_comp, as the LM
//                                               solver forms it, in the
//                                               precision of the environment
//                                               or the one named
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "lcurve_re/lcurve.h"

namespace {

std::string slurp(const char *path) {
    std::ifstream f(path);
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

// Minimal object with dump(), standing in for nlohmann::json.
struct JsonText {
    std::string text;
    std::string dump() const { return text; }
};

void print_vec(const char *tag, const std::vector<double> &v) {
    std::printf("%s", tag);
    for (double x : v) std::printf(" %.17g", x);
    std::printf("\n");
}

int run_model(const char *cfg) {
    Lcurve::Model m(JsonText{slurp(cfg)});
    std::printf("nvary %d\n", m.nvary());
    const std::vector<double> p = m.get_param();
    print_vec("param", p);
    print_vec("range", m.get_range());
    print_vec("dstep", m.get_dstep());
    std::printf("names");
    for (int i = 0; i < m.nvary(); i++) std::printf(" %s", m.get_name(i).c_str());
    std::printf("\n");
    std::printf("limits");
    for (const auto &l : m.get_limit()) std::printf(" %.17g:%.17g", l.first, l.second);
    std::printf("\n");
    double r1, r2;
    m.get_r1r2(r1, r2);
    std::printf("r1r2 %.17g %.17g\n", r1, r2);
    std::printf("q %.17g iangle %.17g nlat1f %d delta_phase %.17g\n", double(m.q), m.iangle.value, m.nlat1f,
                m.delta_phase);
    std::printf("grid %zu %zu %zu %zu\n", m.grid_size(0, true), m.grid_size(1, true), m.grid_size(0, false),
                m.grid_size(1, false));
    // set_param round trip with a shifted vector
    std::vector<double> v = p;
    for (double &x : v) x *= 1.0 + 1e-3;
    std::printf("legal %d\n", m.is_not_legal(v) ? 0 : 1);
    m.set_param(v);
    print_vec("param2", m.get_param());
    return 0;
}

int run_data(const char *in, const char *out) {
    Lcurve::Data d(in);
    d.wrasc(out);
    std::printf("n %zu\n", d.size());
    return 0;
}

int run_eval(const char *cfg, const char *dat, const char *out) {
    Lcurve::Model m(slurp(cfg));
    Lcurve::Data data(dat);
    std::vector<double> sfac(5, 1.0), calc;
    double wdwarf, chisq, wnok, logg1, logg2, rv1, rv2;
    Lcurve::light_curve_comp(m, data, false, true, false, sfac, calc, wdwarf, chisq, wnok, logg1, logg2, rv1, rv2);
    FILE *f = std::fopen(out, "w");
    std::fprintf(f, "comp");
    for (double c : calc) std::fprintf(f, " %.17g", c);
    std::fprintf(f, "\nscalars %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", wdwarf, chisq, wnok, logg1, logg2, rv1,
                 rv2);
    std::vector<double> sfac2;
    Lcurve::light_curve_comp(m, data, true, true, false, sfac2, calc, wdwarf, chisq, wnok, logg1, logg2, rv1, rv2,
                             false);
    std::fprintf(f, "scaled %.17g %.17g %.17g %zu\n", sfac2[0], chisq, wnok, sfac2.size());
    Lcurve::light_curve_comp_fast(m, data, false, true, false, sfac, calc, wdwarf, chisq, wnok, logg1, logg2, rv1,
                                  rv2, 50, false);
    std::fprintf(f, "fast");
    for (double c : calc) std::fprintf(f, " %.17g", c);
    std::fprintf(f, "\n");
    std::vector<Lcurve::Model> batch(3, m);
    batch[1].iangle.value += 1.0;
    batch[2].q.value *= 1.1;
    const Lcurve::BatchResult r = Lcurve::light_curve_comp_batch(batch, data, false, true);
    for (size_t i = 0; i < batch.size(); i++) {
        std::fprintf(f, "batch%zu %d", i, r.status[i]);
        for (size_t k = 0; k < r.n; k++) std::fprintf(f, " %.17g", r.row(i)[k]);
        std::fprintf(f, "\n");
    }
    std::fclose(f);
    return 0;
}

int run_error(const char *cfg, const char *dat) {
    try {
        Lcurve::Model m(slurp(cfg));
        Lcurve::Data data(dat);
        std::vector<double> sfac(5, 1.0), calc;
        double wdwarf, chisq, wnok, logg1, logg2, rv1, rv2;
        Lcurve::light_curve_comp(m, data, false, true, false, sfac, calc, wdwarf, chisq, wnok, logg1, logg2, rv1,
                                 rv2);
    } catch (const Lcurve::Lcurve_Error &e) {
        std::printf("Lcurve_Error: %s\n", e.what());
        return 0;
    }
    std::printf("no error\n");
    return 1;
}

int run_print(const char *cfg, const char *out) {
    Lcurve::Model m(slurp(cfg));
    m.wrasc(out);
    return 0;
}

// h = eps^(1/4) |p| + 1e-10, at most half the parameter's range
// (levmarq_solver.cpp:766-778); column = (c(p + h) - c(p - h)) / 2h.
int run_jac(const char *cfg, const char *dat, const char *out, int j, const char *prec) {
    if (prec) Lcurve::set_precision(std::string(prec) == "mixed" ? LC_PREC_MIXED : LC_PREC_FP64);
    Lcurve::Model m(slurp(cfg));
    Lcurve::Data data(dat);
    const std::vector<double> p = m.get_param();
    const auto lim = m.get_limit();
    double h = std::sqrt(std::sqrt(2.220446049250313e-16)) * std::fabs(p[j]) + 1e-10;
    const double span = lim[j].second - lim[j].first;
    if (std::isfinite(span) && span > 0.0) h = std::min(h, 0.5 * span);
    std::vector<double> sfac(5, 1.0), cp, cm;
    double wdwarf, chisq, wnok, logg1, logg2, rv1, rv2;
    std::vector<double> v = p;
    v[j] = p[j] + h;
    m.set_param(v);
    Lcurve::light_curve_comp(m, data, false, false, false, sfac, cp, wdwarf, chisq, wnok, logg1, logg2, rv1, rv2,
                             false);
    v[j] = p[j] - h;
    m.set_param(v);
    Lcurve::light_curve_comp(m, data, false, false, false, sfac, cm, wdwarf, chisq, wnok, logg1, logg2, rv1, rv2,
                             false);
    FILE *f = std::fopen(out, "w");
    std::fprintf(f, "precision %d\nh %.17g\ncol", Lcurve::get_precision(), h);
    for (size_t k = 0; k < cp.size(); k++) std::fprintf(f, " %.17g", (cp[k] - cm[k]) / (2.0 * h));
    std::fprintf(f, "\n");
    std::fclose(f);
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    const std::string mode = argc > 1 ? argv[1] : "";
    try {
        if (mode == "model" && argc == 3) return run_model(argv[2]);
        if (mode == "data" && argc == 4) return run_data(argv[2], argv[3]);
        if (mode == "eval" && argc == 5) return run_eval(argv[2], argv[3], argv[4]);
        if (mode == "error" && argc == 4) return run_error(argv[2], argv[3]);
        if (mode == "print" && argc == 4) return run_print(argv[2], argv[3]);
        if (mode == "jac" && (argc == 6 || argc == 7))
            return run_jac(argv[2], argv[3], argv[4], std::atoi(argv[5]), argc == 7 ? argv[6] : nullptr);
    } catch (const std::exception &e) {
        std::printf("exception: %s\n", e.what());
        return 3;
    }
    std::fprintf(stderr, "usage: dropin_check model|data|eval|error|print|jac ...\n");
    return 2;
}
