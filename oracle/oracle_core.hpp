// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// CPU FP64 restatement of the spd_K spectral-difference update, written to
// follow the reference's arithmetic and loop/summation order so that a
// correct GPU implementation reproduces it bit for bit. Only tests/, the
// smoke check and bench.py's cpu_baseline leg may build, load or run it; the
// product (spd_k_amd/) never links or calls it.
//
// Every function cites the reference file:line it restates (paths relative
// to the reference repository root).
//
// Pinning (which independent check fixes each part) is listed per section
// below; a part with no independent check says "parity unpinned".
#pragma once
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace orc {

constexpr double PI = 3.141592653589793;   // src/define.hpp:13
constexpr double MIN_C2 = 1e-14;           // src/define.hpp:14
constexpr double RHO_MIN = 1e-10, RHO_MAX = 1e10, P_MIN = 1e-10, P_MAX = 1e10;  // define.hpp:16-19
constexpr int NGH = 1;  // ghost elements (define.hpp:21)
constexpr int nGH = 2;  // ghost FV cells (define.hpp:22)
constexpr int NVAR = 6; // rho, mx, my, mz, E, slot (define.hpp:35)
enum { D_ = 0, VX = 1, VY = 2, VZ = 3, P_ = 4, E_ = 4 };
enum { BC_PERIODIC = 0, BC_GRADFREE = 1, BC_REFLECTIVE = 2 };
enum { INT_ADER = 0, INT_RK = 1 };
enum { IC_SINE = 0, IC_SEDOV, IC_SPHERICAL_BLAST, IC_SQUARE, IC_SOD, IC_SHU_OSHER,
       IC_KH, IC_IMPLOSION, IC_RTI, IC_USER, IC_SEDOV_CELLS };

// std::max / std::min semantics (first argument returned on unordered compare)
static inline double smax(double a, double b) { return (a < b) ? b : a; }
static inline double smin(double a, double b) { return (b < a) ? b : a; }

struct Mat {
    int r = 0, c = 0;
    std::vector<double> a;
    void init(int r_, int c_) { r = r_; c = c_; a.assign((size_t)r * c, 0.0); }
    double& operator()(int i, int j) { return a[(size_t)i * c + j]; }
    double operator()(int i, int j) const { return a[(size_t)i * c + j]; }
};

struct Problem {  // src/global.hpp:22-37 plus the random-perturbation knobs
    double amp = 0.125, v1 = 1.0, v2 = 1.0, v3 = 0.0, d0 = 1.0, d1 = 0.125, p0 = 1.0, p1 = 0.1,
           radius = 0.1, sigma = 0.05;
    int dir = 0;
    double cx = 0.5, cy = 0.5, cz = 0.5;
    int perturb = 0;            // 0: reference sine perturbation, 1: seeded Gaussian noise
    uint64_t seed = 20251202ull;
    int noise_cells = 0;        // 0: one sample per FV sub-cell; M: per cell of an M^d lattice
    double hx = 0, hy = 0, hz = 0;  // element widths (for sedov_cells)
};

struct Config {  // src/global.hpp:41-58 and main.cpp:126-189
    int ndim = 3;
    bool act[3] = {true, true, true};
    int ngh[3] = {NGH, NGH, NGH}, fgh[3] = {nGH, nGH, nGH};
    int problem = 0;
    bool fallback = true;
    double gamma = 1.4, cfl = 0.8, g[3] = {0, 0, 0};
    double nad_tol = 1e-5;
    bool nad_delta = false, nad_moore = true, sed = true, blending = true;
    int bc[3] = {0, 0, 0};
    int integrator = INT_ADER, rk_order = 3;
    bool outputs = false;
    Problem pp;
    int p = 3, Nglob[3] = {8, 8, 8};
    double L[3] = {1, 1, 1};
    double tlim = 0.1, dt_output = 0.1;
    std::string system = "hydro";
    double nu = 0.0, beta = 0.0, eta = 0.0025;
    double vel[3] = {1.0, 1.0, 0.0};   // induction advection velocity (induction.cpp:28-31)
    double loop_c[2] = {0.5, 0.5};     // induction loop centre (initial_conditions.cpp:236-240)
    double loop_R = 0.3, loop_A0 = 0.001;
    int max_steps = -1;                // stop after this many steps (benchmark sampling)
};

// src/structs.hpp:19-76 (class dimension)
struct Dim {
    int dim = 0, Nglob = 1, N = 1, p = 0, Ntot = 1, nsp = 1, nfp = 1, fvn = 1, fvnf = 1, idL = 0, idR = 1;
    double L = 1.0, h = 1.0;
    Mat sdf, sdc;                 // (Ntot, nfp), (Ntot, nsp)
    std::vector<double> fvf, fvc; // fvnf, fvn
    void init(int dim_, int Nglob_, int N_, int p_, int start, double L_, const double* x_fp, bool active);
};

// SD array, C-order (t, var, k, j, i, kk, jj, ii): src/define.hpp:124, structs.hpp:104-129
struct SDA {
    int nt = 1, nv = 1, Nz = 1, Ny = 1, Nx = 1, nz = 1, ny = 1, nx = 1;
    std::string label;
    std::vector<double> a;
    void init(const char* name, int nt_, int nv_, const Dim& Z, const Dim& Y, const Dim& X, bool sz, bool sy, bool sx) {
        label = name; nt = nt_; nv = nv_;
        Nx = X.Ntot; Ny = Y.Ntot; Nz = Z.Ntot;
        nx = sx ? X.nfp : X.nsp; ny = sy ? Y.nfp : Y.nsp; nz = sz ? Z.nfp : Z.nsp;
        a.assign(size(), 0.0);
    }
    size_t size() const { return (size_t)nt * nv * Nz * Ny * Nx * nz * ny * nx; }
    size_t idx(int t, int v, int k, int j, int i, int kk, int jj, int ii) const {
        return ((((((((size_t)t * nv + v) * Nz + k) * Ny + j) * Nx + i) * nz + kk) * ny + jj) * nx + ii);
    }
    double& operator()(int t, int v, int k, int j, int i, int kk, int jj, int ii) { return a[idx(t, v, k, j, i, kk, jj, ii)]; }
    double operator()(int t, int v, int k, int j, int i, int kk, int jj, int ii) const { return a[idx(t, v, k, j, i, kk, jj, ii)]; }
};

// FV array, C-order (var, k, j, i): src/define.hpp:125, fv_structs.hpp:68-93
struct FVA {
    int nv = 1, Nz = 1, Ny = 1, Nx = 1;
    std::string label;
    std::vector<double> a;
    void init(const char* name, int nv_, const Dim& Z, const Dim& Y, const Dim& X, bool sz, bool sy, bool sx) {
        label = name; nv = nv_;
        Nx = X.fvn + (sx ? 1 : 0); Ny = Y.fvn + (sy ? 1 : 0); Nz = Z.fvn + (sz ? 1 : 0);
        a.assign(size(), 0.0);
    }
    size_t size() const { return (size_t)nv * Nz * Ny * Nx; }
    double& operator()(int v, int k, int j, int i) { return a[(((size_t)v * Nz + k) * Ny + j) * Nx + i]; }
    double operator()(int v, int k, int j, int i) const { return a[(((size_t)v * Nz + k) * Ny + j) * Nx + i]; }
};

// ---- polynomials (src/polynomials.cpp) -------------------------------------
double lagrange(const double* x, double y, int i, int n);
double lagrange_prime(const double* x, double y, int i, int n);
void gauss_legendre(double xi, double xf, int n, double* x, double* w);
void flux_points(double* x_fp, const double* x, int n);
void solution_points(double* x_sp, int n);
void lagrange_matrix(Mat& m, const double* xa, const double* xb, int na, int nb);
void lagrange_prime_matrix(Mat& m, const double* xa, const double* xb, int na, int nb);
void ader_matrix(Mat& m, const double* xt, const double* wt, int n);
void integral_matrix(Mat& m, const double* x_fp, const double* x_sp, int n_cv, int n_sp);
void inverse(const Mat& A, Mat& C, int n);

struct Ops {
    int p = 3, T = 4;
    std::vector<double> xgl, wgl, x_sp, x_fp, xx, wx, xt, wt;
    Mat sp_to_fp, fp_to_sp, dfp_to_sp, sp_to_cv, cv_to_sp, fp_to_cv, ader, invader;
    void build(int p, bool ader_int);
};

// ---- run entry (driver) -----------------------------------------------------
int parse_config(int argc, char** argv, Config& cfg, std::string& err);
int run_hydro(Config& cfg, const char* outdir, double* evol_seconds, int* steps, long long* dofs);
int run_induction(Config& cfg, const char* outdir, double* evol_seconds, int* steps, long long* dofs);

}  // namespace orc
