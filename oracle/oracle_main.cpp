// ORACLE — TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
// Command line / input-file handling of the oracle driver, restating the
// reference's configuration semantics (src/main.cpp:86-257,
// src/parameter_input.cpp:24-152): "-i file" plus "block/name=value"
// overrides, last assignment wins, '#' comments, per-problem defaults.
#include <algorithm>
#include <fstream>
#include <map>
#include <sstream>
#include "oracle_core.hpp"

namespace orc {

namespace {
std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}
typedef std::map<std::string, std::map<std::string, std::string>> Params;

bool load_file(const std::string& fn, Params& P, std::string& err) {
    std::ifstream f(fn);
    if (!f.is_open()) { err = "cannot open " + fn; return false; }
    std::string line, block;
    while (std::getline(f, line)) {
        size_t h = line.find('#');
        if (h != std::string::npos) line = line.substr(0, h);
        line = trim(line);
        if (line.empty()) continue;
        if (line[0] == '<') { block = trim(line.substr(1, line.find('>') - 1)); continue; }
        size_t eq = line.find('=');
        if (eq == std::string::npos || block.empty()) { err = "bad line: " + line; return false; }
        P[block][trim(line.substr(0, eq))] = trim(line.substr(eq + 1));
    }
    return true;
}
struct Get {
    Params& P;
    bool has(const char* b, const char* n) { return P.count(b) && P[b].count(n); }
    std::string s(const char* b, const char* n, const std::string& d) { return has(b, n) ? P[b][n] : d; }
    double r(const char* b, const char* n, double d) { return has(b, n) ? std::stod(P[b][n]) : d; }
    int i(const char* b, const char* n, int d) { return has(b, n) ? std::stoi(P[b][n]) : d; }
    bool bo(const char* b, const char* n, bool d) {
        if (!has(b, n)) return d;
        std::string v = P[b][n];
        std::transform(v.begin(), v.end(), v.begin(), ::tolower);
        if (v == "true" || v == "yes" || v == "on" || v == "1") return true;
        if (v == "false" || v == "no" || v == "off" || v == "0") return false;
        throw std::runtime_error("not a boolean: " + v);
    }
};

int bc_of(const std::string& s) {
    if (s == "periodic") return BC_PERIODIC;
    if (s == "gradfree") return BC_GRADFREE;
    if (s == "reflective") return BC_REFLECTIVE;
    return -1;
}
int problem_of(const std::string& s) {
    const char* names[] = {"sine_wave", "sedov", "spherical_blast", "square", "sod_shock_tube", "shu_osher",
                           "kelvin_helmholtz", "implosion", "rti", "user", "sedov_cells"};
    for (int i = 0; i < 11; i++)
        if (s == names[i]) return i;
    return -1;
}
// src/main.cpp:31-65 (field order: amp v1 v2 v3 d0 d1 p0 p1 radius sigma dir)
void problem_defaults(int prob, Problem& pp) {
    auto set = [&](double amp, double v1, double v2, double v3, double d0, double d1, double p0, double p1, double radius,
                   double sigma) {
        pp.amp = amp; pp.v1 = v1; pp.v2 = v2; pp.v3 = v3; pp.d0 = d0; pp.d1 = d1; pp.p0 = p0; pp.p1 = p1;
        pp.radius = radius; pp.sigma = sigma; pp.dir = 0;
    };
    switch (prob) {
        case IC_SINE: set(0.125, 1.0, 1.0, 0.0, 1.0, 0.0, 1.0, 0.0, 0.0, 0.0); break;
        case IC_SQUARE: set(0.0, 1.0, 1.0, 0.0, 1.0, 2.0, 1.0, 0.0, 0.25, 0.0); break;
        case IC_SOD: set(0.0, 0.0, 0.0, 0.0, 1.0, 0.125, 1.0, 0.1, 0.5, 0.0); break;
        case IC_SHU_OSHER: set(0.2, 0.0, 0.0, 0.0, 1.0, 0.0, 1.0, 0.0, 0.0, 0.0); break;
        case IC_KH: set(0.1, 0.5, 0.0, 0.0, 1.0, 2.0, 2.5, 0.0, 0.0, 0.05 / std::sqrt(2.0)); break;
        case IC_IMPLOSION: set(0.0, 0.0, 0.0, 0.0, 1.0, 0.125, 1.0, 0.14, 0.15, 0.0); break;
        case IC_SEDOV: set(0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.1, 0.0); break;
        case IC_SPHERICAL_BLAST: set(0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 10.0, 0.1, 0.1, 0.0); break;
        case IC_RTI: set(0.025, 0.0, 0.0, 0.0, 2.0, 1.0, 1.0, 0.0, 0.5, 0.0); break;
        case IC_USER: set(0.5, 1.0, 0.0, 0.0, 1.0, 0.0, 1.0, 0.0, 0.5, 0.1); break;
        case IC_SEDOV_CELLS: set(1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 1e-5, 0.0, 0.0); break;
    }
}
}  // namespace

int parse_config(int argc, char** argv, Config& cfg, std::string& err) {
    Params P;
    std::vector<std::string> ov;
    for (int n = 1; n < argc; n++) {
        std::string a = argv[n];
        if (a == "-i" && n + 1 < argc) {
            if (!load_file(argv[++n], P, err)) return 1;
        } else if (a.find('/') != std::string::npos && a.find('=') != std::string::npos)
            ov.push_back(a);
        else if (a == "-o" && n + 1 < argc)
            ++n;  // output directory, handled by main
        else { err = "unrecognized argument " + a; return 1; }
    }
    for (auto& o : ov) {
        size_t s = o.find('/'), e = o.find('=');
        P[trim(o.substr(0, s))][trim(o.substr(s + 1, e - s - 1))] = trim(o.substr(e + 1));
    }
    Get g{P};
    try {
        cfg.p = g.i("mesh", "p", 3);
        cfg.Nglob[0] = g.i("mesh", "nx1", 8);
        cfg.Nglob[1] = g.i("mesh", "nx2", 8);
        cfg.Nglob[2] = g.i("mesh", "nx3", 8);
        cfg.L[0] = g.r("mesh", "x1len", 1.0);
        cfg.L[1] = g.r("mesh", "x2len", 1.0);
        cfg.L[2] = g.r("mesh", "x3len", 1.0);
        for (int d = 0; d < 3; d++) {
            cfg.act[d] = cfg.Nglob[d] > 1;
            cfg.ngh[d] = cfg.act[d] ? NGH : 0;
            cfg.fgh[d] = cfg.act[d] ? nGH : 0;
        }
        if (!cfg.act[0]) { err = "the x-direction must be active (nx1>1)"; return 1; }
        cfg.ndim = (int)cfg.act[0] + (int)cfg.act[1] + (int)cfg.act[2];
        cfg.cfl = g.r("time", "cfl", 0.8);
        cfg.gamma = g.r("hydro", "gamma", 1.4);
        cfg.g[0] = g.r("hydro", "g1", 0.0);
        cfg.g[1] = g.r("hydro", "g2", 0.0);
        cfg.g[2] = g.r("hydro", "g3", 0.0);
        cfg.fallback = g.bo("job", "fallback", true);
        cfg.nad_tol = g.r("fallback", "tolerance", 1e-5);
        cfg.nad_delta = g.s("fallback", "NAD", "relative") == "delta";
        cfg.nad_moore = g.s("fallback", "NAD_neighbors", "2nd") == "2nd";
        cfg.sed = g.bo("fallback", "SED", true);
        cfg.blending = g.bo("fallback", "blending", true);
        cfg.problem = problem_of(g.s("problem", "problem", "sine_wave"));
        if (cfg.problem < 0) { err = "unknown problem"; return 1; }
        problem_defaults(cfg.problem, cfg.pp);
        cfg.pp.amp = g.r("problem", "amp", cfg.pp.amp);
        cfg.pp.v1 = g.r("problem", "v1", cfg.pp.v1);
        cfg.pp.v2 = g.r("problem", "v2", cfg.pp.v2);
        cfg.pp.v3 = g.r("problem", "v3", cfg.pp.v3);
        cfg.pp.d0 = g.r("problem", "d0", cfg.pp.d0);
        cfg.pp.d1 = g.r("problem", "d1", cfg.pp.d1);
        cfg.pp.p0 = g.r("problem", "p0", cfg.pp.p0);
        cfg.pp.p1 = g.r("problem", "p1", cfg.pp.p1);
        cfg.pp.radius = g.r("problem", "radius", cfg.pp.radius);
        cfg.pp.sigma = g.r("problem", "sigma", cfg.pp.sigma);
        cfg.pp.dir = g.i("problem", "dir", cfg.pp.dir);
        cfg.pp.perturb = g.s("problem", "perturb", "sine") == "random" ? 1 : 0;
        cfg.pp.seed = (uint64_t)std::stoull(g.s("problem", "seed", "20251202"));
        cfg.pp.noise_cells = g.i("problem", "noise_cells", 0);
        cfg.pp.cx = 0.5 * cfg.L[0];
        cfg.pp.cy = 0.5 * cfg.L[1];
        cfg.pp.cz = 0.5 * cfg.L[2];
        for (int d = 0; d < 3; d++) {
            const char* key[3] = {"x1_bc", "x2_bc", "x3_bc"};
            cfg.bc[d] = bc_of(g.s("mesh", key[d], "periodic"));
            if (cfg.bc[d] < 0) { err = "unknown boundary type"; return 1; }
        }
        cfg.tlim = g.r("time", "tlim", 0.1);
        cfg.outputs = g.has("output", "dt") && std::stod(P["output"]["dt"]) > 0.0;
        cfg.dt_output = cfg.outputs ? std::stod(P["output"]["dt"]) : cfg.tlim;
        std::string integ = g.s("time", "integrator", "ader");
        if (integ == "ader") cfg.integrator = INT_ADER;
        else if (integ == "rk1" || integ == "rk2" || integ == "rk3") { cfg.integrator = INT_RK; cfg.rk_order = integ[2] - '0'; }
        else { err = "unknown integrator " + integ; return 1; }
        cfg.system = g.s("job", "system", "hydro");
        cfg.nu = g.r("hydro", "nu", 0.0);
        cfg.beta = g.r("hydro", "beta", -2. / 3 * cfg.nu);
        cfg.eta = g.r("induction", "nu", 0.0025);
        cfg.vel[0] = g.r("induction", "v1", 1.0);
        cfg.vel[1] = g.r("induction", "v2", 1.0);
        cfg.vel[2] = g.r("induction", "v3", 0.0);
        cfg.loop_c[0] = g.r("induction", "cx", 0.5);
        cfg.loop_c[1] = g.r("induction", "cy", 0.5);
        cfg.loop_R = g.r("induction", "radius", 0.3);
        cfg.loop_A0 = g.r("induction", "amp", 0.001);
        cfg.max_steps = g.i("time", "nlim", -1);
    } catch (std::exception& e) {
        err = e.what();
        return 1;
    }
    return 0;
}

}  // namespace orc

// C entry: run the oracle driver on an argv-style configuration.
extern "C" int orc_run_argv(int argc, char** argv, const char* outdir, double* evol_seconds, int* steps, long long* dofs) {
    orc::Config cfg;
    std::string err;
    if (orc::parse_config(argc, argv, cfg, err)) {
        std::fprintf(stderr, "oracle: %s\n", err.c_str());
        return 1;
    }
    if (cfg.system == "hydro") return orc::run_hydro(cfg, outdir, evol_seconds, steps, dofs);
    if (cfg.system == "induction") return orc::run_induction(cfg, outdir, evol_seconds, steps, dofs);
    std::fprintf(stderr, "oracle: unknown system %s\n", cfg.system.c_str());
    return 1;
}

#ifdef ORC_MAIN
int main(int argc, char** argv) {
    const char* outdir = std::getenv("SPD_OUTPUT_DIR");
    for (int n = 1; n + 1 < argc; n++)
        if (std::string(argv[n]) == "-o") outdir = argv[n + 1];
    double sec = 0;
    int steps = 0;
    long long dofs = 0;
    int rc = orc_run_argv(argc, argv, outdir ? outdir : "output", &sec, &steps, &dofs);
    if (rc) return rc;
    std::printf("evolution: %d steps, %.6f s (%.4f ms/step), %.6e zone-cycles/s\n", steps, sec,
                steps > 0 ? sec / steps * 1e3 : 0.0, sec > 0 ? (double)dofs * steps / sec : 0.0);
    return 0;
}
#endif
