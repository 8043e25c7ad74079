#include "setup.hpp"
#include <sys/stat.h>
#include <algorithm>
#include <cmath>
#include <fstream>
#include <stdexcept>
#include <thread>

namespace spdk_host {

static const double kPi = 3.141592653589793;  // src/define.hpp:13

// ---------------------------------------------------------------------------
// Configuration (src/main.cpp:5-84, :126-189)
static int bc_from(const std::string& s) {
    if (s == "periodic") return BC_PERIODIC;
    if (s == "gradfree") return BC_GRADFREE;
    if (s == "reflective") return BC_REFLECTIVE;
    return -1;
}
static int problem_from(const std::string& s) {
    static const char* names[] = {"sine_wave", "sedov", "spherical_blast", "square", "sod_shock_tube", "shu_osher",
                                  "kelvin_helmholtz", "implosion", "rti", "user", "sedov_cells"};
    for (int i = 0; i < 11; i++)
        if (s == names[i]) return i;
    return -1;
}
// Per-problem defaults of the <problem> block (src/main.cpp:31-65).
static void problem_defaults(int prob, ProblemParams& pp) {
    struct D { double amp, v1, v2, v3, d0, d1, p0, p1, radius, sigma; };
    D d{pp.amp, pp.v1, pp.v2, pp.v3, pp.d0, pp.d1, pp.p0, pp.p1, pp.radius, pp.sigma};
    switch (prob) {
        case P_SINE: d = {0.125, 1.0, 1.0, 0.0, 1.0, 0.0, 1.0, 0.0, 0.0, 0.0}; break;
        case P_SQUARE: d = {0.0, 1.0, 1.0, 0.0, 1.0, 2.0, 1.0, 0.0, 0.25, 0.0}; break;
        case P_SOD: d = {0.0, 0.0, 0.0, 0.0, 1.0, 0.125, 1.0, 0.1, 0.5, 0.0}; break;
        case P_SHU_OSHER: d = {0.2, 0.0, 0.0, 0.0, 1.0, 0.0, 1.0, 0.0, 0.0, 0.0}; break;
        case P_KH: d = {0.1, 0.5, 0.0, 0.0, 1.0, 2.0, 2.5, 0.0, 0.0, 0.05 / std::sqrt(2.0)}; break;
        case P_IMPLOSION: d = {0.0, 0.0, 0.0, 0.0, 1.0, 0.125, 1.0, 0.14, 0.15, 0.0}; break;
        case P_SEDOV: d = {0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.1, 0.0}; break;
        case P_SPHERICAL_BLAST: d = {0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 10.0, 0.1, 0.1, 0.0}; break;
        case P_RTI: d = {0.025, 0.0, 0.0, 0.0, 2.0, 1.0, 1.0, 0.0, 0.5, 0.0}; break;
        case P_USER: d = {0.5, 1.0, 0.0, 0.0, 1.0, 0.0, 1.0, 0.0, 0.5, 0.1}; break;
        case P_SEDOV_CELLS: d = {1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 1e-5, 0.0, 0.0}; break;
    }
    pp.amp = d.amp; pp.v1 = d.v1; pp.v2 = d.v2; pp.v3 = d.v3; pp.d0 = d.d0; pp.d1 = d.d1;
    pp.p0 = d.p0; pp.p1 = d.p1; pp.radius = d.radius; pp.sigma = d.sigma; pp.dir = 0;
}

std::string configure(Params& pin, RunConfig& c) {
    c.p = pin.get_or<int>("mesh/p", 3);
    c.nx[0] = pin.get_or<int>("mesh/nx1", 8);
    c.nx[1] = pin.get_or<int>("mesh/nx2", 8);
    c.nx[2] = pin.get_or<int>("mesh/nx3", 8);
    c.len[0] = pin.get_or<double>("mesh/x1len", 1.0);
    c.len[1] = pin.get_or<double>("mesh/x2len", 1.0);
    c.len[2] = pin.get_or<double>("mesh/x3len", 1.0);
    for (int d = 0; d < 3; d++) c.act[d] = c.nx[d] > 1;
    if (!c.act[0]) return "the x-direction must be active (nx1>1)";
    if (c.p < 1 || c.p > 8) return "mesh/p must be in 1..8";
    c.ndim = (int)c.act[0] + (int)c.act[1] + (int)c.act[2];
    c.cfl = pin.get_or<double>("time/cfl", 0.8);
    c.gamma = pin.get_or<double>("hydro/gamma", 1.4);
    c.g[0] = pin.get_or<double>("hydro/g1", 0.0);
    c.g[1] = pin.get_or<double>("hydro/g2", 0.0);
    c.g[2] = pin.get_or<double>("hydro/g3", 0.0);
    c.fallback = pin.get_or<bool>("job/fallback", true);
    c.nad_tol = pin.get_or<double>("fallback/tolerance", 1e-5);
    c.nad_delta = pin.get_or("fallback/NAD", "relative") == "delta";
    c.nad_moore = pin.get_or("fallback/NAD_neighbors", "2nd") == "2nd";
    c.sed = pin.get_or<bool>("fallback/SED", true);
    c.blending = pin.get_or<bool>("fallback/blending", true);
    c.problem = problem_from(pin.get_or("problem/problem", "sine_wave"));
    if (c.problem < 0) return "unknown problem '" + pin.get<std::string>("problem/problem") + "'";
    problem_defaults(c.problem, c.pp);
    ProblemParams& pp = c.pp;
    pp.amp = pin.get_or<double>("problem/amp", pp.amp);
    pp.v1 = pin.get_or<double>("problem/v1", pp.v1);
    pp.v2 = pin.get_or<double>("problem/v2", pp.v2);
    pp.v3 = pin.get_or<double>("problem/v3", pp.v3);
    pp.d0 = pin.get_or<double>("problem/d0", pp.d0);
    pp.d1 = pin.get_or<double>("problem/d1", pp.d1);
    pp.p0 = pin.get_or<double>("problem/p0", pp.p0);
    pp.p1 = pin.get_or<double>("problem/p1", pp.p1);
    pp.radius = pin.get_or<double>("problem/radius", pp.radius);
    pp.sigma = pin.get_or<double>("problem/sigma", pp.sigma);
    pp.dir = pin.get_or<int>("problem/dir", pp.dir);
    if (pin.has("problem/perturb")) pp.random_perturb = pin.get<std::string>("problem/perturb") == "random";
    if (pin.has("problem/seed")) pp.seed = (uint64_t)std::stoull(pin.get<std::string>("problem/seed"));
    // noise lattice: one sample per cell of an M per axis lattice (M divides
    // every active nx), so runs at different resolutions share one field
    pp.noise_cells = pin.get_or<int>("problem/noise_cells", 0);
    for (int d = 0; d < 3; d++)
        if (pp.noise_cells < 0 || (pp.noise_cells > 0 && c.act[d] && c.nx[d] % pp.noise_cells != 0))
            return "problem/noise_cells must divide mesh/nx1..3";
    pp.cx = 0.5 * c.len[0];
    pp.cy = 0.5 * c.len[1];
    pp.cz = 0.5 * c.len[2];
    const char* keys[3] = {"x1_bc", "x2_bc", "x3_bc"};
    for (int d = 0; d < 3; d++) {
        c.bc[d] = bc_from(pin.get_or(std::string("mesh/") + keys[d], "periodic"));
        if (c.bc[d] < 0) return std::string("unknown boundary type for mesh/") + keys[d];
    }
    c.tlim = pin.get_or<double>("time/tlim", 0.1);
    c.outputs = pin.has("output/dt") && pin.get<double>("output/dt") > 0.0;
    c.dt_output = c.outputs ? pin.get<double>("output/dt") : c.tlim;
    if (pin.has("output/fields")) {
        c.out_troubles = c.out_flux = c.out_wcv = false;
        std::string list = pin.get<std::string>("output/fields") + ",";
        for (size_t a = 0, b; (b = list.find(',', a)) != std::string::npos; a = b + 1) {
            std::string f = list.substr(a, b - a);
            if (f == "troubles") c.out_troubles = true;
            else if (f == "F_ader_fp_x") c.out_flux = true;
            else if (f == "W_cv") c.out_wcv = true;
            else if (!f.empty()) return "output/fields: unknown field '" + f + "' (troubles, F_ader_fp_x, W_cv)";
        }
    }
    std::string integ = pin.get_or("time/integrator", "ader");
    if (integ == "ader") c.integrator = INT_ADER;
    else if (integ.size() == 3 && integ.compare(0, 2, "rk") == 0 && integ[2] >= '1' && integ[2] <= '3') {
        c.integrator = INT_RK;
        c.rk_order = integ[2] - '0';
    } else
        return "unknown integrator '" + integ + "' (ader, rk1, rk2, rk3)";
    c.system = pin.get_or("job/system", "hydro");
    c.nu = pin.get_or<double>("hydro/nu", 0.0);
    c.beta = pin.get_or<double>("hydro/beta", -2. / 3 * c.nu);
    if (c.system == "induction") {
        c.eta = pin.get_or<double>("induction/nu", 0.0025);
        c.vel[0] = pin.has("induction/v1") ? pin.get<double>("induction/v1") : 1.0;
        c.vel[1] = pin.has("induction/v2") ? pin.get<double>("induction/v2") : 1.0;
        c.vel[2] = pin.has("induction/v3") ? pin.get<double>("induction/v3") : 0.0;
        c.loop_c[0] = pin.has("induction/cx") ? pin.get<double>("induction/cx") : 0.5;
        c.loop_c[1] = pin.has("induction/cy") ? pin.get<double>("induction/cy") : 0.5;
        c.loop_R = pin.has("induction/radius") ? pin.get<double>("induction/radius") : 0.3;
        c.loop_A0 = pin.has("induction/amp") ? pin.get<double>("induction/amp") : 0.001;
    }
    c.max_steps = pin.has("time/nlim") ? pin.get<int>("time/nlim") : -1;
    c.path = pin.has("job/path") ? pin.get<std::string>("job/path") : std::string("fused");
    if (c.path != "fused" && c.path != "stages") return "job/path must be fused or stages";
    {
        const std::string fc = pin.get_or<std::string>("job/fv_contraction", "mfma");
        if (fc != "mfma" && fc != "valu") return "job/fv_contraction must be mfma or valu";
        c.fv_valu = fc == "valu";
        const std::string fl = pin.get_or<std::string>("job/fv_layout", "fv");
        if (fl != "fv" && fl != "sd") return "job/fv_layout must be fv or sd";
        c.fv_layout_sd = fl == "sd";
    }
    c.timing = pin.has("job/timing") && pin.get<bool>("job/timing");
    if (pin.has("job/precision")) c.precision = pin.get<std::string>("job/precision");
    if (c.precision != "fp64" && c.precision != "mixed") return "job/precision must be fp64 or mixed";
    if (pin.has("job/picard")) c.picard = pin.get<std::string>("job/picard");
    if (c.picard != "fixed" && c.picard != "residual") return "job/picard must be fixed or residual";
    if (pin.has("job/picard_tol")) c.picard_tol = pin.get<double>("job/picard_tol");
    if (pin.has("job/picard_max")) c.picard_max = pin.get<int>("job/picard_max");
    {
        const std::string gc = pin.get_or<std::string>("job/guess_cap", "restart");
        if (gc != "restart" && gc != "keep") return "job/guess_cap must be restart or keep";
        c.guess_keep = gc == "keep";
    }
    if (pin.has("job/guess")) c.guess = pin.get<std::string>("job/guess");
    if (c.guess != "zero" && c.guess != "mlp") return "job/guess must be zero or mlp";
    if (pin.has("job/mlp_weights")) c.mlp_weights = pin.get<std::string>("job/mlp_weights");
    if (pin.has("job/mode")) c.mode = pin.get<std::string>("job/mode");
    c.final_output = pin.has("job/final_output") && pin.get<bool>("job/final_output");
    if (c.mode != "run" && c.mode != "train_guess") return "job/mode must be run or train_guess";
    if (c.mode == "train_guess") {
        if (!pin.has("mlp/out")) return "job/mode=train_guess needs mlp/out";
        c.mlp_out = pin.get<std::string>("mlp/out");
        if (pin.has("mlp/snapshots")) c.mlp_snapshots = pin.get<int>("mlp/snapshots");
        if (pin.has("mlp/every")) c.mlp_every = pin.get<int>("mlp/every");
        if (pin.has("mlp/sweeps")) c.mlp_sweeps = pin.get<int>("mlp/sweeps");
        if (pin.has("mlp/epochs")) c.mlp_epochs = pin.get<int>("mlp/epochs");
        if (pin.has("mlp/batch")) c.mlp_batch = pin.get<int>("mlp/batch");
        if (pin.has("mlp/hidden")) c.mlp_hidden = pin.get<int>("mlp/hidden");
        if (pin.has("mlp/lr")) c.mlp_lr = pin.get<double>("mlp/lr");
        if (pin.has("mlp/max_target")) c.mlp_max_target = pin.get<double>("mlp/max_target");
        if (pin.has("mlp/seed")) c.mlp_seed = (uint64_t)std::stoull(pin.get<std::string>("mlp/seed"));
        if (c.mlp_snapshots < 2 || c.mlp_every < 1 || c.mlp_sweeps < 1 || c.mlp_epochs < 1 || c.mlp_batch < 64 ||
            c.mlp_hidden < 64)
            return "mlp/: snapshots >= 2, every >= 1, sweeps >= 1, epochs >= 1, batch >= 64, hidden >= 64";
        if (c.system != "hydro" || c.integrator != INT_ADER)
            return "job/mode=train_guess covers the ADER hydro step";
        c.precision = "mixed";
    }
    if (c.system == "induction" && c.precision == "mixed") {
        if (c.integrator != INT_ADER || c.fallback || c.picard != "fixed" || c.guess != "zero")
            return "job/precision=mixed covers the induction ADER step without the FV update (fixed Picard count)";
        return "";
    }
    // with the fallback, the mixed predictor is followed by the FP64 flux
    // evaluation and FV update with trouble detection
    if (c.precision == "mixed" || c.picard == "residual" || c.guess == "mlp") {
        if (c.system != "hydro" || c.integrator != INT_ADER)
            return "job/precision=mixed, job/picard=residual and job/guess=mlp cover the ADER hydro step";
        if (c.precision != "mixed") return "job/picard=residual and job/guess=mlp run on job/precision=mixed";
        if (c.picard_tol <= 0 || c.picard_max < 0 || c.picard_max > 30) return "job/picard_tol > 0, 0 <= job/picard_max <= 30";
        if (c.guess == "mlp" && (c.picard != "residual" || c.mlp_weights.empty()))
            return "job/guess=mlp needs job/picard=residual and job/mlp_weights";
    }
    return "";
}

// ---------------------------------------------------------------------------
// Basis and quadrature (src/polynomials.cpp)
double lagrange(const double* x, double y, int i, int n) {
    double l = 1.;
    for (int j = 0; j < n; j++)
        if (j != i) l *= (y - x[j]) / (x[i] - x[j]);
    return l;
}
double lagrange_prime(const double* x, double y, int i, int n) {
    double s = 0;
    for (int k = 0; k < n; k++) {
        if (k == i) continue;
        double l = 1;
        for (int j = 0; j < n; j++)
            if (j != i && j != k) l *= (y - x[j]) / (x[i] - x[j]);
        s += l / (x[i] - x[k]);
    }
    return s;
}
// Newton iteration on the Legendre polynomial from cos(3.141592654*(i+3/4)/(n+1/2)),
// stopping at |dz| <= 3e-14 (src/polynomials.cpp:46-79)
void gauss_legendre(double a, double b, int n, double* x, double* w) {
    if (n <= 0) {
        x[0] = 0.5;
        w[0] = 1;
        return;
    }
    double xm = 0.5 * (b + a), xl = 0.5 * (b - a);
    for (int i = 0; i < (n + 1) / 2; i++) {
        double z = std::cos(3.141592654 * (i + 0.75) / (n + 0.5)), z1 = 0.0, p1 = 0, p2 = 0, p3 = 0, pp = 0;
        while (std::fabs(z - z1) > 3E-14) {
            p1 = 1.0;
            p2 = 0.0;
            for (int j = 0; j < n; j++) {
                p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j + 1) * z * p2 - j * p3) / (j + 1);
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            z1 = z;
            z = z1 - p1 / pp;
        }
        x[i] = xm - xl * z;
        x[n - i - 1] = xm + xl * z;
        w[i] = (2.0 * xl) / ((1.0 - z * z) * pp * pp);
        w[n - i - 1] = w[i];
    }
}
static void lagrange_matrix(Matrix& m, const double* xa, const double* xb, int na, int nb) {
    m.resize(nb, na);
    for (int j = 0; j < nb; j++)
        for (int i = 0; i < na; i++) m.at(j, i) = lagrange(xa, xb[j], i, na);
}
static void lagrange_prime_matrix(Matrix& m, const double* xa, const double* xb, int na, int nb) {
    m.resize(nb, na);
    for (int j = 0; j < nb; j++)
        for (int i = 0; i < na; i++) m.at(j, i) = lagrange_prime(xa, xb[j], i, na);
}
// cell averages of the source Lagrange basis over [x_fp[k], x_fp[k+1]] by
// GL(n_src-1) quadrature (src/polynomials.cpp:103-124)
static void integral_matrix(Matrix& m, const double* x_fp, const double* x_src, int n_cv, int n_src) {
    m.resize(n_cv, n_src);
    int q = n_src - 1;
    std::vector<double> x(q > 0 ? q : 1), w(q > 0 ? q : 1);
    for (int k = 0; k < n_cv; k++) {
        if (q > 0) gauss_legendre(x_fp[k], x_fp[k + 1], q, x.data(), w.data());
        for (int j = 0; j < n_src; j++) {
            double s;
            if (q > 0) {
                s = 0.0;
                for (int i = 0; i < q; i++) s += lagrange(x_src, x[i], j, q + 1) * w[i];
            } else
                s = 1.0;
            m.at(k, j) = s / (x_fp[k + 1] - x_fp[k]);
        }
    }
}
// Doolittle LU without pivoting and column-wise substitution (src/polynomials.cpp:126-198)
static void invert(const Matrix& A, Matrix& C) {
    int n = A.rows;
    C.resize(n, n);
    std::vector<double> B(A.v), L((size_t)n * n, 0.0), U((size_t)n * n, 0.0), d(n), x(n), b(n, 0.0);
    // B is A row-major: B[r*n+c] = A(r,c)
    for (int k = 0; k < n - 1; k++)
        for (int i = k + 1; i < n; i++) {
            double f = B[(size_t)i * n + k] / B[(size_t)k * n + k];
            L[(size_t)i * n + k] = f;
            for (int j = k + 1; j < n; j++) B[(size_t)i * n + j] -= f * B[(size_t)k * n + j];
        }
    for (int i = 0; i < n; i++) L[(size_t)i * n + i] = 1.0;
    for (int j = 0; j < n; j++)
        for (int i = 0; i <= j; i++) U[(size_t)i * n + j] = B[(size_t)i * n + j];
    for (int k = 0; k < n; k++) {
        b[k] = 1.0;
        d[0] = b[0];
        for (int i = 1; i < n; i++) {
            d[i] = b[i];
            for (int j = 0; j < i; j++) d[i] -= L[(size_t)i * n + j] * d[j];
        }
        x[n - 1] = d[n - 1] / U[(size_t)n * n - 1];
        for (int i = n - 2; i >= 0; i--) {
            x[i] = d[i];
            for (int j = n - 1; j >= i + 1; j--) x[i] = x[i] - U[(size_t)i * n + j] * x[j];
            x[i] = x[i] / U[(size_t)i * n + i];
        }
        for (int i = 0; i < n; i++) C.at(i, k) = x[i];
        b[k] = 0.0;
    }
}

void Operators::build(int p_, bool ader_integrator) {
    p = p_;
    x_gl.assign(p, 0.0);
    w_gl.assign(p, 0.0);
    gauss_legendre(0.0, 1.0, p, x_gl.data(), w_gl.data());
    x_fp.assign(p + 2, 0.0);  // flux points [0, GL_p, 1] (src/polynomials.cpp:81-87)
    x_fp[0] = 0.0;
    for (int i = 1; i <= p; i++) x_fp[i] = x_gl[i - 1];
    x_fp[p + 1] = 1.0;
    x_sp.assign(p + 1, 0.0);  // Chebyshev-Gauss solution points (src/polynomials.cpp:89-93)
    for (int i = 0; i <= p; i++) x_sp[i] = 0.5 * (1.0 - std::cos((2 * i + 1) / (2.0 * (p + 1)) * kPi));
    xq.assign(p + 1, 0.0);
    wq.assign(p + 1, 0.0);
    gauss_legendre(0.0, 1.0, p + 1, xq.data(), wq.data());
    if (ader_integrator) {
        T = p + 1;
        xt = xq;
        wt = wq;
    } else {
        T = 1;
        xt.assign(1, 0.0);
        wt.assign(1, 1.0);
    }
    lagrange_matrix(sp_to_fp, x_sp.data(), x_fp.data(), p + 1, p + 2);
    lagrange_matrix(fp_to_sp, x_fp.data(), x_sp.data(), p + 2, p + 1);
    lagrange_prime_matrix(dfp_to_sp, x_fp.data(), x_sp.data(), p + 2, p + 1);
    integral_matrix(sp_to_cv, x_fp.data(), x_sp.data(), p + 1, p + 1);
    integral_matrix(fp_to_cv, x_fp.data(), x_fp.data(), p + 1, p + 2);
    invert(sp_to_cv, cv_to_sp);
    ader.resize(T, T);
    invader.resize(T, T);
    if (ader_integrator) {
        // ader(j,i) = l_i(1) l_j(1) - l'_j(x_i) w_i (src/polynomials.cpp:95-101)
        for (int j = 0; j < T; j++)
            for (int i = 0; i < T; i++)
                ader.at(j, i) = lagrange(xt.data(), 1, i, T) * lagrange(xt.data(), 1, j, T) -
                                lagrange_prime(xt.data(), xt[i], j, T) * wt[i];
        invert(ader, invader);
    }
}

// ---------------------------------------------------------------------------
// Geometry: element flux-point coordinates and the FV sub-grid of one axis
// (src/structs.hpp:40-75). The last FV face is never assigned (stays 0).
void Axis::build(int dim_, int n_global_, int n_local, int p_, int start, double length_, const double* x_fp, bool active) {
    dim = dim_;
    n_global = n_global_;
    n = n_local;
    p = p_;
    length = length_;
    h = length / n_global;
    if (active) {
        n_sp = p + 1;
        n_fp = p + 2;
        n_total = n + 2 * kNGH;
        fv_ncells = n * n_sp + 2 * kFGH;
        fv_nfaces = fv_ncells + 1;
        id_left = n_sp - kFGH;
        id_right = fv_ncells + id_left;
    }
    sd_faces.resize(n_total, n_fp);
    sd_centers.resize(n_total, n_sp);
    fv_faces.assign(fv_nfaces, 0.0);
    fv_centers.assign(fv_ncells, 0.0);
    for (int j = 0; j < n_total; j++) {
        for (int i = 0; i < n_fp; i++) {
            sd_faces.at(j, i) = (start + j - kNGH + x_fp[i]) * h;
            int g = i + j * n_sp;
            if (g >= id_left && g < id_right) fv_faces[g - id_left] = sd_faces.at(j, i);
        }
        for (int i = 0; i < n_sp; i++) {
            double right = (i + 1 < n_fp) ? sd_faces.at(j, i + 1) : sd_faces.at(j, i);
            sd_centers.at(j, i) = 0.5 * (right + sd_faces.at(j, i));
            int g = i + j * n_sp;
            if (g >= id_left && g < id_right) fv_centers[g - id_left] = sd_centers.at(j, i);
        }
    }
}

// ---------------------------------------------------------------------------
// Initial conditions: primitive variables at a point (src/initial_conditions.cpp:9-177)
static uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
// Standard normal sample for FV sub-cell `cell`: Box-Muller over two
// splitmix64 draws keyed by (seed, cell).
double normal_sample(uint64_t seed, uint64_t cell) {
    uint64_t a = mix64(seed ^ mix64(2 * cell)), b = mix64(seed ^ mix64(2 * cell + 1));
    double u1 = ((double)(a >> 11) + 1.0) * (1.0 / 9007199254740992.0);
    double u2 = (double)(b >> 11) * (1.0 / 9007199254740992.0);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
}

static double ic_point(const RunConfig& c, int var, double x, double y, double z, double noise) {
    const ProblemParams& pp = c.pp;
    double gm = c.gamma, gy = c.g[1];
    bool ay = c.act[1], az = c.act[2];
    switch (c.problem) {
        case P_SINE:
            if (var == 0) return pp.d0 + pp.amp * (std::sin(2 * kPi * (x + y)));
            return var == 1 ? pp.v1 : var == 2 ? pp.v2 : var == 3 ? pp.v3 : var == 4 ? pp.p0 : 0;
        case P_SEDOV: {
            double r = std::sqrt(std::pow(x - 0.5, 2.0) + std::pow(y - 0.5, 2.0) + (az ? std::pow(z - 0.5, 2.0) : 0.0));
            if (var == 0) return pp.d0;
            if (var == 4) return r < pp.radius ? 1E-6 + (gm - 1) : 1E-6;
            return 0;
        }
        case P_SPHERICAL_BLAST: {
            double r = std::sqrt(std::pow(x - pp.cx, 2) + std::pow(y - pp.cy, 2) + (az ? std::pow(z - pp.cz, 2) : 0.0));
            if (var == 0) return pp.d0;
            if (var == 4) return r < pp.radius ? pp.p0 : pp.p1;
            return 0;
        }
        case P_SQUARE: {
            if (var == 0) {
                bool in = std::fabs(x - 0.5) < pp.radius;
                if (ay) in = in && std::fabs(y - 0.5) < pp.radius;
                if (az) in = in && std::fabs(z - 0.5) < pp.radius;
                return in ? pp.d1 : pp.d0;
            }
            return var == 1 ? pp.v1 : var == 2 ? pp.v2 : var == 3 ? pp.v3 : var == 4 ? pp.p0 : 0;
        }
        case P_SOD: {
            double s = (pp.dir == 0 ? x : (pp.dir == 1 ? y : z));
            bool left = s < pp.radius;
            if (var == 0) return left ? pp.d0 : pp.d1;
            if (var == 4) return left ? pp.p0 : pp.p1;
            return 0;
        }
        case P_SHU_OSHER: {
            double xs = x - 1.0;
            bool left = xs < -0.8;
            if (var == 0) return left ? 3.857143 : 1.0 + pp.amp * std::sin(5 * kPi * xs);
            if (var == 1) return left ? 2.629369 : 0.0;
            if (var == 4) return left ? 10.33333 : 1.0;
            return 0;
        }
        case P_KH: {
            bool inner = (y > 0.25 && y < 0.75);
            if (var == 0) return inner ? pp.d1 : pp.d0;
            if (var == 1) return inner ? pp.v1 : -pp.v1;
            if (var == 2) {
                if (pp.random_perturb) return pp.amp * noise * std::exp(-std::pow(y - 0.25, 2) / (2 * pp.sigma * pp.sigma));
                return pp.amp * std::sin(4 * kPi * x) *
                       (std::exp(-std::pow(y - 0.25, 2) / (2 * pp.sigma * pp.sigma)) +
                        std::exp(-std::pow(y - 0.75, 2) / (2 * pp.sigma * pp.sigma)));
            }
            if (var == 4) return pp.p0;
            return 0;
        }
        case P_IMPLOSION: {
            bool inside = (x + y) < pp.radius;
            if (var == 0) return inside ? pp.d1 : pp.d0;
            if (var == 4) return inside ? pp.p1 : pp.p0;
            return 0;
        }
        case P_RTI: {
            double yc = pp.radius, hi = pp.d0, lo = pp.d1, P0 = pp.p0;
            bool top = y > yc;
            if (var == 0) return top ? lo : hi;
            if (var == 2) {
                double dv = std::sqrt(gm * (hi * yc + 1.0) / hi);
                return -pp.amp * dv * std::cos(8.0 * kPi * x);
            }
            if (var == 4) return top ? (P0 + lo * gy * y + (hi - lo) * gy * yc) : (P0 + hi * gy * y);
            return 0;
        }
        case P_USER: return user_ic(var, x, y, z, gm, ay, az, pp);
        case P_SEDOV_CELLS: {
            // energy pp.amp spread over the central 2^d elements, background pressure p1
            if (var == 0) return pp.d0;
            if (var == 4) {
                bool in = std::fabs(x - pp.cx) < pp.hx;
                double vol = 2 * pp.hx;
                if (ay) { in = in && std::fabs(y - pp.cy) < pp.hy; vol *= 2 * pp.hy; }
                if (az) { in = in && std::fabs(z - pp.cz) < pp.hz; vol *= 2 * pp.hz; }
                return in ? (gm - 1) * pp.amp / vol : pp.p1;
            }
            return 0;
        }
    }
    return 0;
}

// GL(p+1) cv averages of the IC, every element and sub-cell
// (src/initial_conditions.cpp:179-228).
void initialize_hydro(const RunConfig& c, const Axis& X, const Axis& Y, const Axis& Z, const Operators& op,
                      std::vector<double>& W) {
    bool ay = c.act[1], az = c.act[2];
    int px = X.n_sp, py = Y.n_sp, pz = Z.n_sp;
    size_t pts = (size_t)px * py * pz, cells = (size_t)X.n_total * Y.n_total * Z.n_total * pts;
    W.assign(cells * kNVAR, 0.0);
    long long ncx = (long long)X.n * X.n_sp, ncy = (long long)Y.n * Y.n_sp;
    // every (element, point) is independent: element rows are shared out
    // over host threads
    const int rows = Z.n_total * Y.n_total;
    auto work = [&](int r0, int r1) {
    for (int r = r0; r < r1; r++) {
        const int k = r / Y.n_total, j = r % Y.n_total;
            for (int i = 0; i < X.n_total; i++)
                for (int kk = 0; kk < pz; kk++)
                    for (int jj = 0; jj < py; jj++)
                        for (int ii = 0; ii < px; ii++) {
                            double noise = 0;
                            if (c.problem == P_KH && c.pp.random_perturb) {
                                auto wrap = [](int e, int ng, int N, int n, int q) {
                                    return (long long)(((e - ng) % N + N) % N) * n + q;
                                };
                                long long gx = wrap(i, c.act[0] ? kNGH : 0, X.n, X.n_sp, ii);
                                long long gy = wrap(j, c.act[1] ? kNGH : 0, Y.n, Y.n_sp, jj);
                                long long gz = wrap(k, c.act[2] ? kNGH : 0, Z.n, Z.n_sp, kk);
                                long long sx = ncx, sy = ncy;
                                if (c.pp.noise_cells > 0) {  // one sample per cell of an M^d lattice
                                    const long long M = c.pp.noise_cells;
                                    gx = gx / (ncx / M);
                                    gy = c.act[1] ? gy / (ncy / M) : 0;
                                    gz = c.act[2] ? gz / ((long long)Z.n * Z.n_sp / M) : 0;
                                    sx = M;
                                    sy = c.act[1] ? M : 1;
                                }
                                noise = normal_sample(c.pp.seed, (uint64_t)((gz * sy + gy) * sx + gx));
                            }
                            size_t cell = (((((size_t)k * Y.n_total + j) * X.n_total + i) * pz + kk) * py + jj) * px + ii;
                            for (int var = 0; var < kNVAR; var++) {
                                double value = 0, x, y = 0, z = 0;
                                for (int nn = 0; nn < pz; nn++) {
                                    if (az) z = Z.sd_faces.at(k, kk) + op.xq[nn] * (Z.sd_faces.at(k, kk + 1) - Z.sd_faces.at(k, kk));
                                    for (int mm = 0; mm < py; mm++) {
                                        if (ay) y = Y.sd_faces.at(j, jj) + op.xq[mm] * (Y.sd_faces.at(j, jj + 1) - Y.sd_faces.at(j, jj));
                                        for (int ll = 0; ll < px; ll++) {
                                            x = X.sd_faces.at(i, ii) + op.xq[ll] * (X.sd_faces.at(i, ii + 1) - X.sd_faces.at(i, ii));
                                            double s = ic_point(c, var, x, y, z, noise);
                                            s *= op.wq[ll];
                                            if (ay) s *= op.wq[mm];
                                            if (az) s *= op.wq[nn];
                                            value += s;
                                        }
                                    }
                                }
                                W[(size_t)var * cells + cell] = value;
                            }
                        }
    }
    };
    unsigned nth = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
    if (rows < 64) nth = 1;
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nth; t++) pool.emplace_back(work, (int)((long long)rows * t / nth), (int)((long long)rows * (t + 1) / nth));
    for (auto& th : pool) th.join();
}

// ---------------------------------------------------------------------------
// Output (src/output.cpp:4-50)
std::string output_folder(const std::string& dir) {
    std::string f = dir.empty() ? std::string("output") : dir;
    for (size_t pos = 1; pos <= f.size(); pos++)
        if (pos == f.size() || f[pos] == '/') mkdir(f.substr(0, pos).c_str(), 0755);
    return f + "/";
}
void write_raw(const std::string& path, const double* a, size_t n) {
    std::ofstream o(path, std::ios::binary);
    if (!o) throw std::runtime_error("cannot write " + path);
    o.write(reinterpret_cast<const char*>(a), (std::streamsize)(n * sizeof(double)));
    o.close();
    if (!o) throw std::runtime_error("short write to " + path + " (file system full?)");
}
std::string sd_name(const std::string& label, int Nx_total, int nx, int n_out) {
    return label + "_N" + std::to_string(Nx_total - 2 * kNGH) + "p" + std::to_string(nx - 1) + "_" + std::to_string(n_out) + "_0.dat";
}
std::string fv_name(const std::string& label, int Nx, int n_out) {
    return label + "_N" + std::to_string(Nx) + "_" + std::to_string(n_out) + "_0.dat";
}
void write_dims(const std::string& dir, const Axis& X, const Axis& Y, const Axis& Z) {
    const Axis* a[3] = {&X, &Y, &Z};
    const char* nm[3] = {"X", "Y", "Z"};
    std::string f = output_folder(dir);
    for (int d = 0; d < 3; d++)
        write_raw(f + nm[d] + "_N" + std::to_string(a[d]->n) + "p" + std::to_string(a[d]->p) + "_0.dat", a[d]->fv_faces.data(),
                  a[d]->fv_faces.size());
}

}  // namespace spdk_host
