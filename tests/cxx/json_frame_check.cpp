// Writes a "lcurve.plot.v1" frame with nlohmann::json (the library the
// reference's stream plot device uses, new_helpers.h:71-109) from
// whitespace-separated input: n, then n lines "time flux ferr fit", then the
// metadata as one JSON text (or {}).  Driven by tests/test_output_formats.py
// to check lcurve_re_amd.data_io.model_frame byte for byte.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "json.hpp"

using json = nlohmann::json;

int main() {
    size_t n = 0;
    std::cin >> n;
    std::vector<double> t(n), f(n), e(n), m(n);
    auto num = [] {  // strtod reads inf and nan, operator>> does not
        std::string tok;
        std::cin >> tok;
        return std::strtod(tok.c_str(), nullptr);
    };
    for (size_t i = 0; i < n; i++) {
        t[i] = num();
        f[i] = num();
        e[i] = num();
        m[i] = num();
    }
    std::string rest, line;
    while (std::getline(std::cin, line)) rest += line;
    json meta = rest.find('{') == std::string::npos ? json::object() : json::parse(rest);
    json frame;
    frame["protocol"] = "lcurve.plot.v1";
    frame["x"] = json::array();
    frame["flux"] = json::array();
    frame["error"] = json::array();
    frame["model"] = json::array();
    frame["residual"] = json::array();
    double x1 = t.front(), x2 = t.front();
    for (double x : t) {
        x1 = std::min(x1, x);
        x2 = std::max(x2, x);
    }
    const double origin = (x2 - x1 < 0.01 * std::abs((x1 + x2) / 2.0)) ? x1 : 0.0;
    for (size_t i = 0; i < n; i++) {
        double sigma = e[i];
        if (!(sigma > 0.0) || !std::isfinite(sigma)) sigma = 1e-3;
        frame["x"].push_back(t[i] - origin);
        frame["flux"].push_back(f[i]);
        frame["error"].push_back(sigma);
        frame["model"].push_back(m[i]);
        frame["residual"].push_back((f[i] - m[i]) / sigma);
    }
    if (!meta.empty()) frame["meta"] = meta;
    std::cout << "@@LCURVE_PLOT@@" << frame.dump() << '\n';
    return 0;
}
