// Host-side set-up of a run: configuration, basis/quadrature and operator
// matrices, element and FV sub-grid geometry, initial conditions and the
// raw output format. Everything here runs once per run on the host.
#pragma once
#include <cstdint>
#include <string>
#include <vector>
#include "params.hpp"

namespace spdk_host {

constexpr int kNGH = 1;  // ghost elements per active axis
constexpr int kFGH = 2;  // ghost FV cells per active axis
constexpr int kNVAR = 6; // rho, m1, m2, m3, E, slot

enum { BC_PERIODIC = 0, BC_GRADFREE = 1, BC_REFLECTIVE = 2 };
enum { INT_ADER = 0, INT_RK = 1 };
enum Problem {
    P_SINE = 0, P_SEDOV, P_SPHERICAL_BLAST, P_SQUARE, P_SOD, P_SHU_OSHER, P_KH, P_IMPLOSION, P_RTI, P_USER,
    P_SEDOV_CELLS
};

struct ProblemParams {  // src/global.hpp:22-37, plus the seeded-noise knobs
    double amp = 0.125, v1 = 1.0, v2 = 1.0, v3 = 0.0, d0 = 1.0, d1 = 0.125, p0 = 1.0, p1 = 0.1, radius = 0.1,
           sigma = 0.05;
    int dir = 0;
    double cx = 0.5, cy = 0.5, cz = 0.5;
    bool random_perturb = false;
    int noise_cells = 0;  // 0: one noise sample per FV sub-cell; M: per cell of an M^d lattice
    uint64_t seed = 20251202ull;
    double hx = 0, hy = 0, hz = 0;  // element widths (sedov_cells deposit region)
};

struct RunConfig {  // src/global.hpp:41-58, src/main.cpp:126-189
    int ndim = 3;
    bool act[3] = {true, true, true};
    int problem = P_SINE;
    bool fallback = true;
    double gamma = 1.4, cfl = 0.8, g[3] = {0, 0, 0};
    double nad_tol = 1e-5;
    bool nad_delta = false, nad_moore = true, sed = true, blending = true;
    int bc[3] = {0, 0, 0};
    int integrator = INT_ADER, rk_order = 3;
    bool outputs = false;
    ProblemParams pp;
    int p = 3, nx[3] = {8, 8, 8};
    double len[3] = {1, 1, 1};
    double tlim = 0.1, dt_output = 0.1;
    std::string system = "hydro";
    double nu = 0, beta = 0, eta = 0.0025;
    double vel[3] = {1.0, 1.0, 0.0};
    double loop_c[2] = {0.5, 0.5}, loop_R = 0.3, loop_A0 = 0.001;
    int max_steps = -1;      // time/nlim: stop after this many steps
    std::string path = "fused";  // job/path: fused | stages
    bool fv_valu = false;        // job/fv_contraction=valu: mixed face integrals on the VALU instead of MFMA
    bool fv_layout_sd = false;   // job/fv_layout=sd: FV update state in the SD layout (U_cv) instead of the FV cell layout
    bool timing = false;         // job/timing: HIP-event timing of the dominant kernel
    std::string precision = "fp64";  // job/precision: fp64 | mixed (FP32 state and contractions)
    std::string picard = "fixed";    // job/picard: fixed (p predictor sweeps) | residual
    double picard_tol = 1e-6;        // job/picard_tol: element residual tolerance
    int picard_max = 0;              // job/picard_max: per-element cap of predictor sweeps (0: the fixed count p)
    bool guess_keep = false;         // job/guess_cap=keep: from the learned guess, elements at the cap keep their state (else restart)
    std::string guess = "zero";      // job/guess: zero | mlp
    std::string mlp_weights;         // job/mlp_weights: trained network file (job/guess=mlp)
    std::string mode = "run";        // job/mode: run | train_guess
    bool final_output = false;       // job/final_output: write the state after the last step
    // output/fields (hydro): comma-separated subset of troubles, F_ader_fp_x,
    // W_cv; the writer's default is all of them
    bool out_troubles = true, out_flux = true, out_wcv = true;
    // <mlp> block (job/mode=train_guess): fixed-point data from this run, then training
    std::string mlp_out;             // mlp/out: weights file written
    int mlp_snapshots = 8, mlp_every = 4, mlp_sweeps = 20, mlp_epochs = 6, mlp_batch = 4096, mlp_hidden = 256;
    double mlp_lr = 1e-3;
    double mlp_max_target = 50.0;    // mlp/max_target: training rows with a larger |target| are left out
    uint64_t mlp_seed = 1;
};

// Parses the configuration from the parameter table (src/main.cpp:126-189);
// returns an error message or "".
std::string configure(Params& pin, RunConfig& cfg);

// ---- basis and operators (src/polynomials.cpp) -----------------------------
double lagrange(const double* x, double y, int i, int n);
double lagrange_prime(const double* x, double y, int i, int n);
void gauss_legendre(double a, double b, int n, double* x, double* w);

struct Matrix {
    int rows = 0, cols = 0;
    std::vector<double> v;
    void resize(int r, int c) { rows = r; cols = c; v.assign((size_t)r * c, 0.0); }
    double& at(int i, int j) { return v[(size_t)i * cols + j]; }
    double at(int i, int j) const { return v[(size_t)i * cols + j]; }
};

struct Operators {
    int p = 0, T = 1;
    std::vector<double> x_gl, w_gl, x_sp, x_fp, xq, wq, xt, wt;
    Matrix sp_to_fp, fp_to_sp, dfp_to_sp, sp_to_cv, cv_to_sp, fp_to_cv, ader, invader;
    void build(int p, bool ader_integrator);
};

// ---- geometry (src/structs.hpp:19-76) --------------------------------------
struct Axis {
    int dim = 0, n_global = 1, n = 1, p = 0, n_total = 1, n_sp = 1, n_fp = 1, fv_ncells = 1, fv_nfaces = 1,
        id_left = 0, id_right = 1;
    double length = 1.0, h = 1.0;
    Matrix sd_faces, sd_centers;
    std::vector<double> fv_faces, fv_centers;
    void build(int dim, int n_global, int n_local, int p, int start, double length, const double* x_fp, bool active);
};

// ---- initial conditions (src/initial_conditions.cpp:9-228, src/user_ic.hpp) --
double user_ic(int var, double x, double y, double z, double gm, bool ay, bool az, const ProblemParams& pp);
// cv averages of the primitive IC over every element (ghosts included),
// layout (1, 6, Nz, Ny, Nx, nz, ny, nx)
void initialize_hydro(const RunConfig& cfg, const Axis& X, const Axis& Y, const Axis& Z, const Operators& op,
                      std::vector<double>& W_cv);
double normal_sample(uint64_t seed, uint64_t cell);

// ---- output (src/output.cpp:4-50) -------------------------------------------
std::string output_folder(const std::string& dir);  // creates it
void write_raw(const std::string& path, const double* a, size_t n);
std::string sd_name(const std::string& label, int Nx_total, int nx, int n_out);
std::string fv_name(const std::string& label, int Nx, int n_out);
void write_dims(const std::string& dir, const Axis& X, const Axis& Y, const Axis& Z);

}  // namespace spdk_host
