// config.cpp — `.cfg` reader and the canned cases behind the C ABI.
//
// Same file format, keys, defaults and error texts as the reference reader
// (src/config.cpp:243-347): `key = value` lines, `#` comments, comma-separated
// doubles read with std::stod, unknown keys rejected. Parsing starts from the
// default case with its obstacle removed (config.cpp:244-245), and the map
// falls back to the state box only when map_lb[0] and map_ub[0] are both 0
// (config.cpp:343-345) — so a file that omits map_* keeps the default map,
// exactly as in the reference.
//
// Extension keys (no reference counterpart; absent => reference behaviour):
//   target      = lb0,lb1,lb2,lb3,ub0,ub1,ub2,ub3   (repeatable) extra target box
//   target_band = c, w     candidate box cells additionally need |x*y - c| <= w

#include <cctype>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "gsc_amd.h"

struct gsc_config {
    std::string name;
    std::vector<gsc_rect4> obstacles;
    std::vector<gsc_rect4> targets;
    gsc_case_desc desc;

    void bind() {
        desc.name = name.c_str();
        desc.obstacles = obstacles.empty() ? nullptr : obstacles.data();
        desc.obstacle_count = static_cast<uint32_t>(obstacles.size());
        desc.extra_targets = targets.empty() ? nullptr : targets.data();
        desc.extra_target_count = static_cast<uint32_t>(targets.size());
    }
};

namespace {

constexpr double kPiLiteral = 3.141592653589793;

gsc_rect4 rect(const double lb[4], const double ub[4]) {
    gsc_rect4 r;
    for (int i = 0; i < 4; ++i) {
        r.lb[i] = lb[i];
        r.ub[i] = ub[i];
    }
    return r;
}

// CaseConfig member defaults (config.hpp:41-67) + make_default_case (config.cpp:129-139).
void set_default_case(gsc_config* c) {
    std::memset(&c->desc, 0, sizeof(c->desc));
    gsc_case_desc& d = c->desc;
    c->name = "unicycle_default";
    const double slb[4] = {-2.0, -2.0, -kPiLiteral, 0.0};
    const double sub[4] = {2.0, 2.0, kPiLiteral, 2.0};
    const double seta[4] = {0.5, 0.5, 0.7853981633974483, 0.5};
    std::memcpy(d.state_lb, slb, sizeof(slb));
    std::memcpy(d.state_ub, sub, sizeof(sub));
    std::memcpy(d.state_eta, seta, sizeof(seta));
    d.wrap_dim = 2;
    d.input_lb[0] = -0.5;
    d.input_lb[1] = -1.0;
    d.input_ub[0] = 0.5;
    d.input_ub[1] = 1.0;
    d.input_eta[0] = 0.5;
    d.input_eta[1] = 1.0;
    const double clb[4] = {-0.5, -0.5, -kPiLiteral, 0.0};
    const double cub[4] = {0.5, 0.5, kPiLiteral, 1.5};
    d.candidate = rect(clb, cub);
    d.map = rect(d.state_lb, d.state_ub);
    c->obstacles.clear();
    const double olb[4] = {0.5, -0.5, -kPiLiteral, 0.0};
    const double oub[4] = {1.0, 0.5, kPiLiteral, 2.0};
    c->obstacles.push_back(rect(olb, oub));
    c->targets.clear();
    d.constraint_type = GSC_CONSTRAINT_NONE;
    d.hyperbolic_a = 4.0;
    d.hyperbolic_b = 4.0;
    d.hyperbolic_c = 16.0;
    d.elliptic_a = 2.0;
    d.elliptic_b = 2.0;
    const double dw[4] = {0.05, 0.05, 0.02, 0.02};
    std::memcpy(d.disturbance_half_width, dw, sizeof(dw));
    d.sample_time = 0.2;
    d.integration_substeps = 32;
    d.max_iterations = 32;
    d.verbose = 1;
}

// make_wraparound_case, config.cpp:141-154
void set_wraparound_case(gsc_config* c) {
    set_default_case(c);
    gsc_case_desc& d = c->desc;
    c->name = "wraparound_case";
    d.state_lb[2] = -kPiLiteral;
    d.state_ub[2] = kPiLiteral;
    d.state_eta[2] = 0.5235987755982988;
    d.disturbance_half_width[2] = 0.35;
    const double clb[4] = {-0.5, -0.5, 2.6, 0.0};
    const double cub[4] = {0.5, 0.5, kPiLiteral, 1.5};
    d.candidate = rect(clb, cub);
    d.map = rect(d.state_lb, d.state_ub);
    c->obstacles.clear();
}

// make_smoke_case, config.cpp:156-185
void set_smoke_case(gsc_config* c) {
    set_default_case(c);
    gsc_case_desc& d = c->desc;
    c->name = "smoke_case";
    const double slb[4] = {-6.0, -6.0, -kPiLiteral, 0.0};
    const double sub[4] = {6.0, 6.0, kPiLiteral, 4.0};
    const double seta[4] = {0.3, 0.3, 0.39269908169872414, 0.5};
    std::memcpy(d.state_lb, slb, sizeof(slb));
    std::memcpy(d.state_ub, sub, sizeof(sub));
    std::memcpy(d.state_eta, seta, sizeof(seta));
    d.input_lb[0] = -1.0;
    d.input_ub[0] = 1.0;
    d.input_eta[0] = 0.25;
    d.input_lb[1] = -1.5;
    d.input_ub[1] = 1.5;
    d.input_eta[1] = 0.3;
    const double clb[4] = {-1.0, -1.0, -kPiLiteral, 0.0};
    const double cub[4] = {1.0, 1.0, kPiLiteral, 2.0};
    d.candidate = rect(clb, cub);
    d.map = rect(d.state_lb, d.state_ub);
    c->obstacles.clear();
    d.max_iterations = 4;
    d.verbose = 0;
}

// make_hyperbolic_case, config.cpp:187-241 (a fresh CaseConfig, not the default case:
// its obstacle list is empty).
void set_hyperbolic_case(gsc_config* c) {
    set_default_case(c);
    c->obstacles.clear();
    gsc_case_desc& d = c->desc;
    c->name = "hyperbolic_case";
    const double slb[4] = {-3.5, -3.5, -kPiLiteral, 0.2};
    const double sub[4] = {3.5, 3.5, kPiLiteral, 2.0};
    const double seta[4] = {0.1, 0.1, 0.1, 0.2};
    std::memcpy(d.state_lb, slb, sizeof(slb));
    std::memcpy(d.state_ub, sub, sizeof(sub));
    std::memcpy(d.state_eta, seta, sizeof(seta));
    d.wrap_dim = 2;
    d.input_lb[0] = 0.2;
    d.input_ub[0] = 2.0;
    d.input_lb[1] = -1.0;
    d.input_ub[1] = 1.0;
    d.input_eta[0] = 0.3;
    d.input_eta[1] = 0.4;
    const double clb[4] = {-1.0, -1.0, -kPiLiteral, 0.2};
    const double cub[4] = {1.0, 1.0, kPiLiteral, 2.0};
    d.candidate = rect(clb, cub);
    d.map = rect(d.state_lb, d.state_ub);
    d.constraint_type = GSC_CONSTRAINT_HYPERBOLIC;
    d.hyperbolic_a = 4.0;
    d.hyperbolic_b = 4.0;
    d.hyperbolic_c = 16.0;
    const double dw[4] = {0.12, 0.12, 0.12, 0.0};
    std::memcpy(d.disturbance_half_width, dw, sizeof(dw));
    d.sample_time = 1.0;
    d.integration_substeps = 32;
    d.max_iterations = 100;
    d.verbose = 1;
}

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) ++a;
    while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) --b;
    return s.substr(a, b - a);
}

std::vector<double> csv_doubles(const std::string& value) {
    std::vector<double> out;
    std::stringstream ss(value);
    std::string tok;
    while (std::getline(ss, tok, ',')) out.push_back(std::stod(trim(tok)));
    return out;
}

void take(const std::vector<double>& v, size_t n, double* out, const char* key) {
    if (v.size() != n)
        throw std::runtime_error(std::string(key) + (n == 4 ? " expects 4 comma-separated numbers"
                                                            : " expects 2 comma-separated numbers"));
    for (size_t i = 0; i < n; ++i) out[i] = v[i];
}

void parse_stream(std::istream& in, gsc_config* c) {
    set_default_case(c);
    c->obstacles.clear();
    gsc_case_desc& d = c->desc;
    std::string line;
    while (std::getline(in, line)) {
        const auto hash = line.find('#');
        if (hash != std::string::npos) line = line.substr(0, hash);
        line = trim(line);
        if (line.empty()) continue;
        const auto eq = line.find('=');
        if (eq == std::string::npos) throw std::runtime_error("expected key=value line: " + line);
        const std::string key = trim(line.substr(0, eq));
        const std::string value = trim(line.substr(eq + 1));
        if (key == "name") {
            c->name = value;
        } else if (key == "state_lb") {
            take(csv_doubles(value), 4, d.state_lb, "state_lb");
        } else if (key == "state_ub") {
            take(csv_doubles(value), 4, d.state_ub, "state_ub");
        } else if (key == "state_eta") {
            take(csv_doubles(value), 4, d.state_eta, "state_eta");
        } else if (key == "wrap_dim") {
            d.wrap_dim = std::stoi(value);
        } else if (key == "input_lb") {
            take(csv_doubles(value), 2, d.input_lb, "input_lb");
        } else if (key == "input_ub") {
            take(csv_doubles(value), 2, d.input_ub, "input_ub");
        } else if (key == "input_eta") {
            take(csv_doubles(value), 2, d.input_eta, "input_eta");
        } else if (key == "candidate_lb") {
            take(csv_doubles(value), 4, d.candidate.lb, "candidate_lb");
        } else if (key == "candidate_ub") {
            take(csv_doubles(value), 4, d.candidate.ub, "candidate_ub");
        } else if (key == "map_lb") {
            take(csv_doubles(value), 4, d.map.lb, "map_lb");
        } else if (key == "map_ub") {
            take(csv_doubles(value), 4, d.map.ub, "map_ub");
        } else if (key == "disturbance_half_width") {
            take(csv_doubles(value), 4, d.disturbance_half_width, "disturbance_half_width");
        } else if (key == "sample_time") {
            d.sample_time = std::stod(value);
        } else if (key == "integration_substeps") {
            d.integration_substeps = static_cast<uint32_t>(std::stoul(value));
        } else if (key == "max_iterations") {
            d.max_iterations = static_cast<uint32_t>(std::stoul(value));
        } else if (key == "verbose") {
            const std::string v = trim(value);
            d.verbose = (v == "1" || v == "true" || v == "TRUE") ? 1 : 0;
        } else if (key == "obstacle" || key == "target") {
            const auto v = csv_doubles(value);
            if (v.size() != 8) throw std::runtime_error(key + " expects 8 comma-separated numbers");
            gsc_rect4 r;
            for (int i = 0; i < 4; ++i) {
                r.lb[i] = v[i];
                r.ub[i] = v[i + 4];
            }
            (key == "obstacle" ? c->obstacles : c->targets).push_back(r);
        } else if (key == "target_band") {
            const auto v = csv_doubles(value);
            if (v.size() != 2) throw std::runtime_error("target_band expects 2 comma-separated numbers");
            d.target_band_enabled = 1;
            d.target_band_c = v[0];
            d.target_band_w = v[1];
        } else if (key == "constraint_type") {
            if (value == "none")
                d.constraint_type = GSC_CONSTRAINT_NONE;
            else if (value == "hyperbolic")
                d.constraint_type = GSC_CONSTRAINT_HYPERBOLIC;
            else if (value == "elliptic")
                d.constraint_type = GSC_CONSTRAINT_ELLIPTIC;
            else
                throw std::runtime_error("unknown constraint_type: " + value);
        } else if (key == "hyperbolic_a") {
            d.hyperbolic_a = std::stod(value);
        } else if (key == "hyperbolic_b") {
            d.hyperbolic_b = std::stod(value);
        } else if (key == "hyperbolic_c") {
            d.hyperbolic_c = std::stod(value);
        } else if (key == "elliptic_a") {
            d.elliptic_a = std::stod(value);
        } else if (key == "elliptic_b") {
            d.elliptic_b = std::stod(value);
        } else {
            throw std::runtime_error("unknown config key: " + key);
        }
    }
    if (d.map.ub[0] == 0.0 && d.map.lb[0] == 0.0) d.map = rect(d.state_lb, d.state_ub);
}

void put_err(const char* msg, char* err, size_t len) {
    if (err && len) std::snprintf(err, len, "%s", msg);
}

int finish(gsc_config* c, gsc_config** out) {
    c->bind();
    *out = c;
    return GSC_OK;
}

}  // namespace

extern "C" {

int gsc_config_load(const char* path, gsc_config** out, char* err, size_t err_len) {
    if (!path || !out) return GSC_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    std::ifstream in(path);
    if (!in) {
        put_err((std::string("failed to open config: ") + path).c_str(), err, err_len);
        return GSC_ERR_PARSE;
    }
    auto* c = new gsc_config();
    try {
        parse_stream(in, c);
    } catch (const std::exception& ex) {
        put_err(ex.what(), err, err_len);
        delete c;
        return GSC_ERR_PARSE;
    }
    return finish(c, out);
}

int gsc_config_parse(const char* text, gsc_config** out, char* err, size_t err_len) {
    if (!text || !out) return GSC_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    std::istringstream in{std::string(text)};
    auto* c = new gsc_config();
    try {
        parse_stream(in, c);
    } catch (const std::exception& ex) {
        put_err(ex.what(), err, err_len);
        delete c;
        return GSC_ERR_PARSE;
    }
    return finish(c, out);
}

const gsc_case_desc* gsc_config_desc(const gsc_config* c) { return c ? &c->desc : nullptr; }

void gsc_config_free(gsc_config* c) { delete c; }

int gsc_config_builtin(int which, gsc_config** out) {
    if (!out) return GSC_ERR_INVALID_ARGUMENT;
    auto* c = new gsc_config();
    switch (which) {
        case 0: set_default_case(c); break;
        case 1: set_wraparound_case(c); break;
        case 2: set_smoke_case(c); break;
        case 3: set_hyperbolic_case(c); break;
        default: delete c; *out = nullptr; return GSC_ERR_INVALID_ARGUMENT;
    }
    return finish(c, out);
}

}  // extern "C"
