// Host driver: the counterpart of src/main.cpp and the integrators of
// src/hydro_ader.hpp / src/induction_ader.hpp. It parses the run, builds
// geometry, operators and initial conditions on the host, and sequences the
// device kernels through the C ABI of include/spdk.h.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <algorithm>
#include <memory>
#include <numeric>
#include <random>
#include <stdexcept>
#include <vector>
#include "../../../include/spdk.h"
#include "params.hpp"
#include "setup.hpp"
#include "driver.hpp"

namespace spdk {
void set_error(const std::string& msg);  // spdk_last_error() returns it
}

namespace spdk_host {

void check(int rc, const char* what) {
    if (rc != SPDK_OK) throw std::runtime_error(std::string(what) + ": " + spdk_last_error());
}
#define CK(call) check((call), #call)

double* DeviceArena::alloc(spdk_ctx* ctx, size_t count) {
    void* p = nullptr;
    CK(spdk_malloc(&p, count * sizeof(double)));
    ptrs.push_back(p);
    CK(spdk_memset0(ctx, p, count * sizeof(double)));
    return (double*)p;
}
void DeviceArena::release() {
    for (void* p : ptrs) spdk_free(p);
    ptrs.clear();
}
void fill_kernel_timing(spdk_ctx* ctx, spdk_run_result* res) {
    spdk_timing_read(ctx, &res->kernel_ms_total, &res->kernel_launches, res->kernel_name, sizeof(res->kernel_name));
}

size_t sd_count(const spdk_sd_shape& s) { return (size_t)s.nt * s.nv * s.Nz * s.Ny * s.Nx * s.nz * s.ny * s.nx; }
size_t fv_count(const spdk_fv_shape& s) { return (size_t)s.nv * s.Nz * s.Ny * s.Nx; }

spdk_ctx* make_context(const RunConfig& cfg, const Operators& op, const Axis* ax[3]) {
    spdk_params prm;
    std::memset(&prm, 0, sizeof(prm));
    prm.p = cfg.p;
    for (int d = 0; d < 3; d++) {
        prm.active[d] = cfg.act[d];
        prm.bc[d] = cfg.bc[d];
        prm.g[d] = cfg.g[d];
        prm.h[d] = ax[d]->h;
        prm.vel[d] = cfg.vel[d];
    }
    prm.n_ader = op.T;
    prm.rk_order = cfg.integrator == INT_RK ? cfg.rk_order : 0;
    prm.gamma = cfg.gamma;
    prm.cfl = cfg.cfl;
    prm.nu = cfg.nu;
    prm.beta = cfg.beta;
    prm.nad_tol = cfg.nad_tol;
    prm.nad_delta = cfg.nad_delta;
    prm.nad_moore = cfg.nad_moore;
    prm.sed = cfg.sed;
    prm.blending = cfg.blending;
    spdk_operators ops;
    std::memset(&ops, 0, sizeof(ops));
    ops.mat[SPDK_M_SP_TO_FP] = op.sp_to_fp.v.data();
    ops.mat[SPDK_M_FP_TO_SP] = op.fp_to_sp.v.data();
    ops.mat[SPDK_M_DFP_TO_SP] = op.dfp_to_sp.v.data();
    ops.mat[SPDK_M_SP_TO_CV] = op.sp_to_cv.v.data();
    ops.mat[SPDK_M_CV_TO_SP] = op.cv_to_sp.v.data();
    ops.mat[SPDK_M_FP_TO_CV] = op.fp_to_cv.v.data();
    ops.mat[SPDK_M_INVADER] = op.invader.v.data();
    ops.wt = op.wt.data();
    for (int d = 0; d < 3; d++) {
        ops.fv_faces[d] = ax[d]->fv_faces.data();
        ops.fv_centers[d] = ax[d]->fv_centers.data();
        ops.fv_nfaces[d] = (int)ax[d]->fv_faces.size();
        ops.fv_ncells[d] = (int)ax[d]->fv_centers.size();
    }
    spdk_ctx* ctx = nullptr;
    CK(spdk_ctx_create(&prm, &ops, nullptr, &ctx));
    return ctx;
}

// Weights file of the learned Picard guess: "SPDKMLP1", int32 dims[4]
// (in, h1, h2, out), int32 degree and dimensionality it was trained for,
// uint64 parameter count, FP32 parameters (the spdk_mlp_set_params layout).
struct MlpFile {
    int dims[4] = {0, 0, 0, 0};
    int p = 0, dim = 0;
    std::vector<float> params;
};
void save_mlp(const std::string& path, const MlpFile& f) {
    std::ofstream o(path, std::ios::binary);
    if (!o) throw std::runtime_error("cannot write " + path);
    o.write("SPDKMLP1", 8);
    o.write((const char*)f.dims, sizeof(f.dims));
    o.write((const char*)&f.p, 4);
    o.write((const char*)&f.dim, 4);
    uint64_t n = f.params.size();
    o.write((const char*)&n, 8);
    o.write((const char*)f.params.data(), n * 4);
}
MlpFile load_mlp(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot read " + path);
    char magic[8];
    MlpFile f;
    uint64_t n = 0;
    in.read(magic, 8);
    in.read((char*)f.dims, sizeof(f.dims));
    in.read((char*)&f.p, 4);
    in.read((char*)&f.dim, 4);
    in.read((char*)&n, 8);
    if (!in || std::memcmp(magic, "SPDKMLP1", 8) != 0 || n > (1ull << 28)) throw std::runtime_error(path + ": not an MLP weights file");
    f.params.resize(n);
    in.read((char*)f.params.data(), n * 4);
    if (!in) throw std::runtime_error(path + ": truncated");
    return f;
}

// ===========================================================================
// Hydro integrator: src/hydro_ader.hpp:76-586
struct HydroRun {
    RunConfig& cfg;
    std::string outdir;
    Axis X, Y, Z;
    Operators op;
    spdk_ctx* ctx = nullptr;
    DeviceArena mem;
    int T = 1, n_stages = 1, n_output = 0, n_step = 0;
    double rk_a[3] = {0, 0, 0};
    double t = 0, dt = 0, Dt = 0;
    bool fused = true;
    bool fused_pred = false;  // fused predictor and flux evaluation ahead of the stage FV update (fallback runs)
    spdk_sd_shape s_sp{}, s_ader{}, s_fp[3]{};
    spdk_fv_shape s_fv{}, s_th{}, s_fvd[3]{};
    double *U_sp = nullptr, *W_sp = nullptr, *W_cv = nullptr, *U_cv = nullptr, *T_sweep = nullptr, *U_ader = nullptr,
           *U0 = nullptr;
    double *U_fp[3] = {}, *F_fp[3] = {};
    double *U_new = nullptr, *W_new = nullptr, *W_old = nullptr, *troubles = nullptr, *theta = nullptr,
           *theta_tmp = nullptr;
    double *F_fv[3] = {}, *T_fp[3] = {};
    double* dU[3] = {};  // viscous gradients at the solution points (nu > 0), zeroed
    uint16_t* abits = nullptr;  // SED coefficient words of the trouble detection (zeroed)
    // sub-cell face fluxes of every ADER node, [t * 3 + d], when the flux
    // evaluation forms the face integrals itself (fv_direct)
    std::vector<double*> F_fvn;
    bool fv_direct = false;
    // the FV update keeps the cell values in the FV cell layout across the
    // ADER nodes (U_fva and U_new take turns), when the element-tile
    // transforms apply (leading active axes, at most 6 points per axis)
    bool fv_state = false;
    double* U_fva = nullptr;
    void* ws = nullptr;
    std::vector<double> host;  // staging for outputs
    bool timing = false;
    bool mixed = false;        // job/precision=mixed
    spdk_picard pc{};
    spdk_picard_stats pst{};
    spdk_mlp* net = nullptr;   // learned Picard guess (job/guess=mlp)

    HydroRun(RunConfig& c, const std::string& od) : cfg(c), outdir(od) {}
    ~HydroRun() {
        if (ctx) spdk_ctx_sync(ctx);
        if (ctx) spdk_mixed_attach_mlp(ctx, nullptr);
        if (net) spdk_mlp_destroy(net);
        mem.release();
        if (ctx) spdk_ctx_destroy(ctx);
    }

    void setup() {
        int p = cfg.p;
        op.build(p, cfg.integrator == INT_ADER);
        X.build(0, cfg.nx[0], cfg.act[0] ? cfg.nx[0] : 1, cfg.act[0] ? p : 0, 0, cfg.len[0], op.x_fp.data(), cfg.act[0]);
        Y.build(1, cfg.nx[1], cfg.act[1] ? cfg.nx[1] : 1, cfg.act[1] ? p : 0, 0, cfg.len[1], op.x_fp.data(), cfg.act[1]);
        Z.build(2, cfg.nx[2], cfg.act[2] ? cfg.nx[2] : 1, cfg.act[2] ? p : 0, 0, cfg.len[2], op.x_fp.data(), cfg.act[2]);
        cfg.pp.hx = X.h;
        cfg.pp.hy = Y.h;
        cfg.pp.hz = Z.h;
        if (cfg.outputs) write_dims(outdir, X, Y, Z);
        T = op.T;
        if (cfg.integrator == INT_RK) {
            n_stages = cfg.rk_order;
            if (cfg.rk_order == 2) rk_a[1] = 0.5;
            if (cfg.rk_order == 3) { rk_a[1] = 0.75; rk_a[2] = 1.0 / 3.0; }
        }
        const Axis* ax[3] = {&X, &Y, &Z};
        ctx = make_context(cfg, op, ax);
        fused = false;
        mixed = cfg.precision == "mixed";
        if (!mixed && cfg.path == "fused" && cfg.integrator == INT_ADER) {
            int Nq[3] = {X.n_total, Y.n_total, Z.n_total};
            fused = spdk_fused_workspace_bytes(ctx, Nq) > 0;  // else the stage sequence runs
        } else if (!mixed && cfg.path == "fused" && !cfg.fallback) {
            int Nq[3] = {X.n_total, Y.n_total, Z.n_total};
            fused = spdk_fused_workspace_bytes(ctx, Nq) > 0;
        }
        // with the fallback, the fused kernel runs the predictor and the last
        // flux evaluation, the stage kernels the FV update
        fused_pred = fused && cfg.fallback;
        // with the fallback, the fused FP64 sweep and the mixed predictor's
        // flux evaluation form the face integrals of every node themselves
        fv_direct = fused_pred || (mixed && cfg.fallback);
        spdk_timing_enable(ctx, timing ? 1 : 0);

        s_sp = spdk_sd_shape{1, kNVAR, Z.n_total, Y.n_total, X.n_total, Z.n_sp, Y.n_sp, X.n_sp};
        s_ader = s_sp;
        s_ader.nt = T;
        for (int d = 0; d < 3; d++) {
            s_fp[d] = s_ader;
            if (cfg.act[d]) (d == 0 ? s_fp[d].nx : d == 1 ? s_fp[d].ny : s_fp[d].nz) = op.p + 2;
        }
        size_t nsp = sd_count(s_sp);
        U_sp = mem.alloc(ctx, nsp);
        W_cv = mem.alloc(ctx, nsp);
        T_sweep = mem.alloc(ctx, nsp);
        W_sp = mem.alloc(ctx, nsp);
        F_fp[0] = mem.alloc(ctx, sd_count(s_fp[0]));  // output array in every path
        if (mixed) {
            int N[3] = {X.n_total, Y.n_total, Z.n_total};
            size_t b = spdk_mixed_workspace_bytes(ctx, N);
            if (b == 0) throw std::runtime_error(std::string("mixed workspace: ") + spdk_last_error());
            void* w = nullptr;
            CK(spdk_malloc(&w, b));
            mem.ptrs.push_back(w);
            CK(spdk_memset0(ctx, w, b));
            ws = w;
            pc.residual = cfg.picard == "residual";
            pc.guess = SPDK_GUESS_ZERO;
            pc.tol = cfg.picard_tol;
            pc.max_sweeps = cfg.picard_max;
            pc.guess_keep = cfg.guess_keep ? 1 : 0;
            if (cfg.guess == "mlp") {
                MlpFile f = load_mlp(cfg.mlp_weights);
                if (f.p != cfg.p || f.dim != cfg.ndim)
                    throw std::runtime_error(cfg.mlp_weights + ": trained for another degree or dimensionality");
                CK(spdk_mlp_create(f.dims, 65536, spdk_ctx_stream(ctx), &net));
                if (spdk_mlp_param_count(net) != f.params.size()) throw std::runtime_error(cfg.mlp_weights + ": size");
                CK(spdk_mlp_set_params(net, f.params.data()));
                CK(spdk_mixed_attach_mlp(ctx, net));
                pc.guess = SPDK_GUESS_MLP;
            }
        }
        if ((!fused && (!mixed || cfg.fallback)) || fused_pred) {  // the stage arrays (with fallback, also behind
                                                                     // the mixed or fused predictor)
            U_cv = mem.alloc(ctx, nsp);
            if (!fv_direct) U_ader = mem.alloc(ctx, sd_count(s_ader));
            if (cfg.nu > 0.0)  // every axis: an inactive axis' gradient stays zero
                for (int d = 0; d < 3; d++) dU[d] = mem.alloc(ctx, sd_count(s_ader));
            if (cfg.integrator == INT_RK) U0 = mem.alloc(ctx, nsp);
            const bool stage_fluxes = !fv_direct;
            for (int d = 0; d < 3; d++) {
                if (!cfg.act[d]) continue;
                if (stage_fluxes) U_fp[d] = mem.alloc(ctx, sd_count(s_fp[d]));
                if (d && !fv_direct) F_fp[d] = mem.alloc(ctx, sd_count(s_fp[d]));
            }
        }
        if (fused) {
            int N[3] = {X.n_total, Y.n_total, Z.n_total};
            size_t b = spdk_fused_workspace_bytes(ctx, N);
            if (b == 0) throw std::runtime_error(std::string("fused workspace: ") + spdk_last_error());
            void* w = nullptr;
            CK(spdk_malloc(&w, b));
            mem.ptrs.push_back(w);
            ws = w;
            if (cfg.integrator == INT_RK) U0 = nullptr;
        }
        if (cfg.fallback) {
            s_fv = spdk_fv_shape{kNVAR, Z.fv_ncells, Y.fv_ncells, X.fv_ncells};
            s_th = spdk_fv_shape{1, Z.fv_ncells, Y.fv_ncells, X.fv_ncells};
            size_t nfv = fv_count(s_fv);
            U_new = mem.alloc(ctx, nfv);
            W_new = mem.alloc(ctx, nfv);
            W_old = mem.alloc(ctx, nfv);
            troubles = mem.alloc(ctx, nfv);
            theta = mem.alloc(ctx, fv_count(s_th));
            theta_tmp = mem.alloc(ctx, fv_count(s_th));
            {
                const int na = cfg.act[0] + cfg.act[1] + cfg.act[2];
                fv_state = !cfg.fv_layout_sd && op.p + 1 <= 6 && cfg.act[0] && (na < 2 || cfg.act[1]) &&
                           (na < 3 || cfg.act[2]);
            }
            if (fv_state) U_fva = mem.alloc(ctx, nfv);
            if (fv_direct) F_fvn.assign((size_t)T * 3, nullptr);
            abits = (uint16_t*)mem.alloc(ctx, (nfv / kNVAR * sizeof(uint16_t) + sizeof(double) - 1) / sizeof(double));
            for (int d = 0; d < 3; d++) {
                s_fvd[d] = s_fv;
                if (cfg.act[d]) (d == 0 ? s_fvd[d].Nx : d == 1 ? s_fvd[d].Ny : s_fvd[d].Nz) += 1;
                if (!cfg.act[d]) continue;
                if (fv_direct) {  // one face array per node (zeroed: variable 5 and the ghost faces stay 0)
                    for (int a = 0; a < T; a++) F_fvn[(size_t)a * 3 + d] = mem.alloc(ctx, fv_count(s_fvd[d]));
                    continue;
                }
                F_fv[d] = mem.alloc(ctx, fv_count(s_fvd[d]));
                spdk_sd_shape st = s_fp[d];
                st.nt = 1;
                T_fp[d] = mem.alloc(ctx, sd_count(st));
            }
        }
        // initial conditions (src/hydro_ader.hpp:188-212)
        initialize_hydro(cfg, X, Y, Z, op, host);
        CK(spdk_memcpy_h2d(ctx, W_cv, host.data(), nsp * sizeof(double)));
        CK(spdk_transform(ctx, W_cv, W_sp, T_sweep, s_sp, SPDK_M_CV_TO_SP));
        CK(spdk_conservatives(ctx, W_sp, U_sp, s_sp));
        CK(spdk_min_dt(ctx, W_cv, s_sp, &Dt));
        std::printf("dx = %.17g dt = %.17g\n", X.h, Dt);
        if (!(Dt > 0.0) || !std::isfinite(Dt))
            throw std::runtime_error("the initial time step is " + std::to_string(Dt) +
                                     ": the initial state is outside the physical range");
        if (cfg.outputs) write_outputs();
    }

    // One '.' per step as the reference's loop prints (src/hydro_ader.hpp:246),
    // flushed at most every 2 s so that a long run shows its progress.
    std::chrono::steady_clock::time_point last_flush = std::chrono::steady_clock::now();
    void progress_dot() {
        std::fputc('.', stdout);
        auto now = std::chrono::steady_clock::now();
        if (std::chrono::duration<double>(now - last_flush).count() > 2.0) {
            std::fflush(stdout);
            last_flush = now;
        }
    }

    void write_outputs() {  // src/hydro_ader.hpp:419-426
        std::string f = output_folder(outdir);
        std::printf("\nOUTPUT %d\n", n_output);
        if (cfg.fallback && cfg.out_troubles) {
            host.resize(fv_count(s_fv));
            CK(spdk_memcpy_d2h(ctx, host.data(), troubles, host.size() * sizeof(double)));
            write_raw(f + fv_name("troubles", s_fv.Nx, n_output), host.data(), host.size());
        }
        if (cfg.out_flux) {
            host.resize(sd_count(s_fp[0]));
            CK(spdk_memcpy_d2h(ctx, host.data(), F_fp[0], host.size() * sizeof(double)));
            write_raw(f + sd_name("F_ader_fp_x", s_fp[0].Nx, s_fp[0].nx, n_output), host.data(), host.size());
        }
        if (cfg.out_wcv) {
            host.resize(sd_count(s_sp));
            CK(spdk_memcpy_d2h(ctx, host.data(), W_cv, host.size() * sizeof(double)));
            write_raw(f + sd_name("W_cv", s_sp.Nx, s_sp.nx, n_output), host.data(), host.size());
        }
        n_output++;
    }

    void solve_fluxes() {  // src/hydro_ader.hpp:278-289, with Viscosity (:380-405) when nu > 0
        const bool visc = cfg.nu > 0.0;
        for (int d = 0; d < 3; d++)
            if (cfg.act[d]) CK(spdk_transform_1d(ctx, U_ader, s_ader, U_fp[d], s_fp[d], SPDK_M_SP_TO_FP, d, -1));
        for (int d = 0; d < 3; d++)
            if (cfg.act[d]) CK(spdk_hydro_flux(ctx, U_fp[d], F_fp[d], s_fp[d], d));
        for (int d = 0; d < 3; d++)
            if (cfg.act[d]) CK(spdk_fill_sd_traces(ctx, U_fp[d], s_fp[d], d));
        for (int d = 0; d < 3; d++)
            if (cfg.act[d]) CK(spdk_sd_riemann(ctx, U_fp[d], F_fp[d], s_fp[d], d, visc ? 1 : 0));
        if (!visc) return;
        for (int d = 0; d < 3; d++)
            if (cfg.act[d]) CK(spdk_sd_gradient(ctx, U_fp[d], s_fp[d], dU[d], s_ader, d));
        for (int d = 0; d < 3; d++)
            if (cfg.act[d])
                CK(spdk_viscous_flux(ctx, U_fp[d], s_fp[d], dU[d], dU[(d + 1) % 3], dU[(d + 2) % 3], s_ader, d));
        for (int d = 0; d < 3; d++)
            if (cfg.act[d]) CK(spdk_fill_sd_traces(ctx, U_fp[d], s_fp[d], d));
        for (int d = 0; d < 3; d++)
            if (cfg.act[d]) CK(spdk_sd_rusanov(ctx, U_fp[d], F_fp[d], s_fp[d], d));
    }
    void correct() {
        if (cfg.fallback) fv_update();
        else CK(spdk_ader_correct(ctx, U_sp, U_ader, F_fp[0], F_fp[1], F_fp[2], s_ader, op.p + 2, dt));
    }
    void ader_step() {  // src/hydro_ader.hpp:291-305
        CK(spdk_copy_ader(ctx, U_sp, U_ader, s_ader));
        for (int a = 0; a < T; a++) {
            solve_fluxes();
            if (a < T - 1) CK(spdk_ader_predict(ctx, U_sp, U_ader, F_fp[0], F_fp[1], F_fp[2], s_ader, op.p + 2, dt));
        }
        correct();
    }
    void rk_step() {  // src/hydro_ader.hpp:313-325
        CK(spdk_memcpy_d2d(ctx, U0, U_sp, sd_count(s_sp) * sizeof(double)));
        for (int s = 0; s < n_stages; s++) {
            CK(spdk_copy_ader(ctx, U_sp, U_ader, s_ader));
            solve_fluxes();
            correct();
            if (rk_a[s] > 0) CK(spdk_rk_combine(ctx, U_sp, U0, s_sp, rk_a[s]));
        }
    }
    // The FV update with the cell values in the FV cell layout: U_fva and
    // U_new take turns as the state (the blended update rewrites only the
    // cells whose fluxes the blend changed); at the end U_sp, and when given
    // W_cv and the next dt, in the same pass.
    void fv_update_state(double* W_out, double* dt_next) {  // src/hydro_ader.hpp:500-575
        const int mask = (1 << 0) | (1 << 4);
        CK(spdk_fv_state_from_sd(ctx, U_sp, s_sp, U_fva, s_fv));
        double *A = U_fva, *B = U_new;
        for (int a = 0; a < T; a++) {
            double* Fv[3] = {F_fv[0], F_fv[1], F_fv[2]};
            if (fv_direct) {
                for (int d = 0; d < 3; d++) Fv[d] = F_fvn[(size_t)a * 3 + d];
            } else {
                for (int d = 0; d < 3; d++)
                    if (cfg.act[d]) CK(spdk_fv_face_integral(ctx, F_fp[d], s_fp[d], F_fv[d], s_fvd[d], T_fp[d], a, d));
            }
            // W_new: the detected variables and density and pressure (PAD)
            CK(spdk_fv_state_update_prims(ctx, B, W_new, W_old, A, Fv[0], Fv[1], Fv[2], s_fv, a, dt, mask | 0x11));
            CK(spdk_fill_fv_ghosts(ctx, W_old, s_fv));
            CK(spdk_fill_fv_ghosts(ctx, W_new, s_fv));
            CK(spdk_detect_troubles_bits(ctx, W_new, W_old, troubles, abits, s_fv, 1, mask));
            CK(spdk_fill_fv_ghosts(ctx, troubles, s_fv));
            CK(spdk_blend_theta(ctx, troubles, theta_tmp, theta, s_fv));
            CK(spdk_fill_fv_ghosts(ctx, theta, s_th));
            CK(spdk_muscl_hllc_blend(ctx, W_old, theta, Fv[0], Fv[1], Fv[2], s_fv, op.wt[a] * dt));
            CK(spdk_fv_state_update_blended(ctx, B, A, Fv[0], Fv[1], Fv[2], theta, s_fv, a, dt));
            std::swap(A, B);
        }
        CK(spdk_fv_state_to_sd(ctx, A, s_fv, U_sp, s_sp, W_out, dt_next));
    }
    void fv_update() {  // src/hydro_ader.hpp:500-575
        if (fv_state) {
            fv_update_state(nullptr, nullptr);
            return;
        }
        const int mask = (1 << 0) | (1 << 4);
        CK(spdk_transform(ctx, U_sp, U_cv, T_sweep, s_sp, SPDK_M_SP_TO_CV));
        for (int a = 0; a < T; a++) {
            double* Fv[3] = {F_fv[0], F_fv[1], F_fv[2]};
            if (fv_direct) {  // the flux evaluation formed this node's face integrals
                for (int d = 0; d < 3; d++) Fv[d] = F_fvn[(size_t)a * 3 + d];
            } else {
                for (int d = 0; d < 3; d++)
                    if (cfg.act[d]) CK(spdk_fv_face_integral(ctx, F_fp[d], s_fp[d], F_fv[d], s_fvd[d], T_fp[d], a, d));
            }
            // update = 0 and the primitives of U_new and U_old in one pass, ghosts filled on W
            CK(spdk_fv_update_prims(ctx, U_new, W_new, W_old, U_cv, s_sp, Fv[0], Fv[1], Fv[2], s_fv, a, dt));
            CK(spdk_fill_fv_ghosts(ctx, W_old, s_fv));
            CK(spdk_fill_fv_ghosts(ctx, W_new, s_fv));
            CK(spdk_detect_troubles_bits(ctx, W_new, W_old, troubles, abits, s_fv, 1, mask));
            CK(spdk_fill_fv_ghosts(ctx, troubles, s_fv));
            CK(spdk_blend_theta(ctx, troubles, theta_tmp, theta, s_fv));
            CK(spdk_fill_fv_ghosts(ctx, theta, s_th));
            CK(spdk_muscl_hllc_blend(ctx, W_old, theta, Fv[0], Fv[1], Fv[2], s_fv, op.wt[a] * dt));
            CK(spdk_fv_update_blended(ctx, U_new, U_cv, s_sp, Fv[0], Fv[1], Fv[2], theta, s_fv, a, dt));
        }
        CK(spdk_transform(ctx, U_cv, U_sp, T_sweep, s_sp, SPDK_M_CV_TO_SP));
    }

    // One step and the dt that follows it (src/hydro_ader.hpp:235-243); a dt
    // that is not a positive finite number stops the run with an error (the
    // state has left the physical range) instead of stalling the time loop.
    double step(bool output_due) {
        double next = step_raw(output_due);
        if (!(next > 0.0) || !std::isfinite(next))
            throw std::runtime_error("the next time step is " + std::to_string(next) +
                                     ": the state has left the physical range (t = " + std::to_string(t) + ")");
        return next;
    }
    // the FV update, then W_cv (for a step whose output shows it) and the next
    // dt from the same pass
    double fv_update_finish(bool output_due) {
        double next = 0;
        fv_update_state(output_due ? W_cv : nullptr, &next);
        return next;
    }
    // the x flux-point fluxes are formed only for a step whose output shows them
    double* fx_out(bool output_due) { return cfg.out_flux && (output_due || cfg.final_output) ? F_fp[0] : nullptr; }
    double step_raw(bool output_due) {
        if (mixed && cfg.fallback) {
            // the predictor in mixed precision (fixed count or residual
            // test, zero-order or learned guess), then the last flux
            // evaluation and the FV update with trouble detection and the
            // MUSCL fallback in FP64 (src/hydro_ader.hpp:291-305, :500-586)
            CK(spdk_mixed_predict_fv_fluxes(ctx, U_sp, s_sp, ws, dt, &pc, &pst, F_fvn.data(), s_fv, fx_out(output_due),
                                            cfg.fv_valu ? 1 : 0));
            if (fv_state) return fv_update_finish(output_due);
            fv_update();
            CK(spdk_primitives(ctx, U_sp, W_sp, s_sp));
            CK(spdk_transform(ctx, W_sp, W_cv, T_sweep, s_sp, SPDK_M_SP_TO_CV));
            double next = 0;
            CK(spdk_min_dt(ctx, W_cv, s_sp, &next));
            return next;
        }
        if (fused_pred) {
            // the x flux-point fluxes only when an output will show them
            CK(spdk_fused_predict_fv_fluxes(ctx, U_sp, s_sp, ws, dt, F_fvn.data(), s_fv,
                                            fx_out(output_due), s_ader));
            if (fv_state) return fv_update_finish(output_due);
            fv_update();
            CK(spdk_primitives(ctx, U_sp, W_sp, s_sp));
            CK(spdk_transform(ctx, W_sp, W_cv, T_sweep, s_sp, SPDK_M_SP_TO_CV));
            double next = 0;
            CK(spdk_min_dt(ctx, W_cv, s_sp, &next));
            return next;
        }
        if (mixed) {
            double next = 0;
            CK(spdk_hydro_step_mixed(ctx, U_sp, s_sp, ws, dt, output_due ? W_cv : nullptr, fx_out(output_due), &pc,
                                     &pst, &next));
            return next;
        }
        if (fused) {
            double next = 0;
            CK(spdk_hydro_step_fused(ctx, U_sp, s_sp, ws, dt, output_due ? W_cv : nullptr, fx_out(output_due), &next));
            return next;
        }
        if (cfg.integrator == INT_RK) rk_step();
        else ader_step();
        CK(spdk_primitives(ctx, U_sp, W_sp, s_sp));
        CK(spdk_transform(ctx, W_sp, W_cv, T_sweep, s_sp, SPDK_M_SP_TO_CV));
        double next = 0;
        CK(spdk_min_dt(ctx, W_cv, s_sp, &next));
        return next;
    }

    // Learned-guess training (job/mode=train_guess): snapshots of this run, each
    // paired with its Picard fixed point (mlp/sweeps sweeps from the zero-order
    // guess), then Adam on the device.
    void train_guess() {
        int din = 0, dout = 0;
        CK(spdk_mixed_guess_dims(ctx, &din, &dout));
        const int N[3] = {X.n_total, Y.n_total, Z.n_total};
        const size_t E = (size_t)N[0] * N[1] * N[2];
        const int S = cfg.mlp_snapshots;
        float *Xs = nullptr, *Ys = nullptr;
        void* p = nullptr;
        CK(spdk_malloc(&p, (size_t)S * E * din * 4));
        mem.ptrs.push_back(p);
        Xs = (float*)p;
        CK(spdk_malloc(&p, (size_t)S * E * dout * 4));
        mem.ptrs.push_back(p);
        Ys = (float*)p;
        dt = Dt;
        auto t0 = std::chrono::steady_clock::now();
        for (int k = 0; k < S; k++) {
            for (int st = 0; st < cfg.mlp_every; st++) {
                double next = step(false);
                t += dt;
                n_step++;
                dt = next;
            }
            CK(spdk_mixed_iterate(ctx, U_sp, s_sp, ws, dt, cfg.mlp_sweeps));
            CK(spdk_mixed_guess_data(ctx, ws, N, dt, 0, (int)E, Xs + (size_t)k * E * din, din,
                                     Ys + (size_t)k * E * dout, dout));
            CK(spdk_ctx_sync(ctx));
            std::printf("snapshot %d: t = %.6g, dt = %.6g\n", k, t, dt);
        }
        const double gen_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        const int dims[4] = {din, cfg.mlp_hidden, cfg.mlp_hidden, dout};
        const int B = cfg.mlp_batch;
        CK(spdk_mlp_create(dims, B, spdk_ctx_stream(ctx), &net));
        // initial parameters: N(0, 2/fan_in) weights, zero biases
        std::mt19937_64 rng(cfg.mlp_seed);
        std::normal_distribution<float> nd(0.f, 1.f);
        MlpFile f;
        for (int i = 0; i < 4; i++) f.dims[i] = dims[i];
        f.p = cfg.p;
        f.dim = cfg.ndim;
        f.params.resize(spdk_mlp_param_count(net));
        {
            size_t o = 0;
            for (int l = 0; l < 3; l++) {
                const float sc = std::sqrt(2.0f / dims[l]) * (l == 2 ? 0.1f : 1.0f);
                for (size_t i = 0; i < (size_t)dims[l + 1] * dims[l]; i++) f.params[o++] = nd(rng) * sc;
                for (int i = 0; i < dims[l + 1]; i++) f.params[o++] = 0.f;
            }
        }
        CK(spdk_mlp_set_params(net, f.params.data()));
        // rows whose targets are not finite (Picard diverged there within
        // mlp/sweeps sweeps) or exceed mlp/max_target (elements nearly uniform
        // inside whose fixed point is driven by their neighbours: the target
        // is normalised by their tiny amplitude) are left out. Of the rest, a
        // seeded random eighth of the earlier snapshots' rows is held out
        // (the loss on unseen rows of the training distribution), and the
        // last snapshot is never trained on (the loss at a later time).
        void* strm = spdk_ctx_stream(ctx);
        std::vector<int> train_rows, held_rows, last_rows;
        {
            const size_t rows = (size_t)S * E;
            void* fp = nullptr;
            CK(spdk_malloc(&fp, 2 * rows));
            mem.ptrs.push_back(fp);
            std::vector<unsigned char> ok(2 * rows);
            CK(spdk_rows_finite(Ys, dout, (int)rows, dout, (float)cfg.mlp_max_target, (unsigned char*)fp, strm));
            CK(spdk_rows_finite(Xs, din, (int)rows, din, 1e6f, (unsigned char*)fp + rows, strm));
            CK(spdk_memcpy_d2h(ctx, ok.data(), fp, 2 * rows));
            // ghost elements are not part of the solution: left out as well
            auto interior = [&](size_t e) {
                const int i = (int)(e % N[0]), j = (int)(e / N[0] % N[1]), k = (int)(e / ((size_t)N[0] * N[1]));
                const int ijk[3] = {i, j, k};
                for (int d = 0; d < 3; d++)
                    if (N[d] > 1 && (ijk[d] < 1 || ijk[d] > N[d] - 2)) return false;
                return true;
            };
            size_t ghosts = 0;
            std::uniform_int_distribution<int> pick8(0, 7);
            for (size_t r = 0; r < rows; r++) {
                if (!interior(r % E)) { ghosts++; continue; }
                if (!(ok[r] && ok[rows + r])) continue;
                if (r >= (size_t)(S - 1) * E) last_rows.push_back((int)r);
                else (pick8(rng) == 0 ? held_rows : train_rows).push_back((int)r);
            }
            std::printf("training rows: %zu, held out: %zu random + %zu of the last snapshot, dropped (non-finite "
                        "or |target| >= %g): %zu\n", train_rows.size(), held_rows.size(), last_rows.size(),
                        cfg.mlp_max_target, rows - ghosts - train_rows.size() - held_rows.size() - last_rows.size());
        }
        const size_t ntrain = train_rows.size();
        float *Xb = nullptr, *Yb = nullptr;
        int *idx = nullptr, *hidx = nullptr, *lidx = nullptr;
        CK(spdk_malloc(&p, (size_t)B * din * 4)); mem.ptrs.push_back(p); Xb = (float*)p;
        CK(spdk_malloc(&p, (size_t)B * dout * 4)); mem.ptrs.push_back(p); Yb = (float*)p;
        CK(spdk_malloc(&p, ntrain * 4 + 4)); mem.ptrs.push_back(p); idx = (int*)p;
        CK(spdk_malloc(&p, held_rows.size() * 4 + 4)); mem.ptrs.push_back(p); hidx = (int*)p;
        CK(spdk_malloc(&p, last_rows.size() * 4 + 4)); mem.ptrs.push_back(p); lidx = (int*)p;
        if (!held_rows.empty()) CK(spdk_memcpy_h2d(ctx, hidx, held_rows.data(), held_rows.size() * 4));
        if (!last_rows.empty()) CK(spdk_memcpy_h2d(ctx, lidx, last_rows.data(), last_rows.size() * 4));
        auto t1 = std::chrono::steady_clock::now();
        // mean loss over a row set (device index array, n rows), batches of up to B
        auto eval = [&](const int* dev_idx, size_t n) -> double {
            double acc = 0;
            size_t cnt = 0;
            for (size_t r = 0; r < n; r += B) {
                const int b = (int)std::min((size_t)B, n - r);
                float loss = 0;
                CK(spdk_gather_rows(Xs, din, dev_idx + r, b, din, Xb, strm));
                CK(spdk_gather_rows(Ys, dout, dev_idx + r, b, dout, Yb, strm));
                CK(spdk_mlp_gradients(net, Xb, din, Yb, dout, b, nullptr, &loss));
                acc += (double)loss * b;
                cnt += b;
            }
            return cnt ? acc / cnt : 0.0;
        };
        std::printf("training: %zu rows (%d snapshots), in %d hidden %d out %d\n", ntrain, S - 1, din,
                    cfg.mlp_hidden, dout);
        std::printf("held-out loss before training %.4e, last snapshot %.4e\n", eval(hidx, held_rows.size()),
                    eval(lidx, last_rows.size()));
        for (int ep = 0; ep < cfg.mlp_epochs; ep++) {
            std::shuffle(train_rows.begin(), train_rows.end(), rng);
            CK(spdk_memcpy_h2d(ctx, idx, train_rows.data(), ntrain * 4));
            double acc = 0;
            size_t nb = 0, nrd = 0;
            const float lr = (float)(cfg.mlp_lr * (ep + 1 == cfg.mlp_epochs ? 0.3 : 1.0));
            for (size_t b = 0; b + B <= ntrain; b += B) {
                CK(spdk_gather_rows(Xs, din, idx + b, B, din, Xb, strm));
                CK(spdk_gather_rows(Ys, dout, idx + b, B, dout, Yb, strm));
                float loss = 0;
                const bool rd = (nb % 64) == 0;
                CK(spdk_mlp_train_step(net, Xb, din, Yb, dout, B, lr, rd ? &loss : nullptr));
                if (rd) { acc += loss; nrd++; }
                nb++;
            }
            std::printf("epoch %d: train loss %.4e (sampled), held-out loss %.4e, last snapshot %.4e\n", ep,
                        nrd ? acc / nrd : 0.0, eval(hidx, held_rows.size()), eval(lidx, last_rows.size()));
        }
        CK(spdk_ctx_sync(ctx));
        const double train_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
        CK(spdk_mlp_get_params(net, f.params.data()));
        save_mlp(cfg.mlp_out, f);
        std::printf("data %.2f s, training %.2f s; wrote %s\n", gen_s, train_s, cfg.mlp_out.c_str());
    }

    void evolve(int warmup, spdk_run_result* res) {  // src/hydro_ader.hpp:215-273
        dt = Dt;
        double t_output = cfg.dt_output;
        if (warmup > 0) {  // untimed steps; the state is restored afterwards
            double* save = mem.alloc(ctx, sd_count(s_sp));
            CK(spdk_memcpy_d2d(ctx, save, U_sp, sd_count(s_sp) * sizeof(double)));
            for (int w = 0; w < warmup; w++) step(false);
            CK(spdk_memcpy_d2d(ctx, U_sp, save, sd_count(s_sp) * sizeof(double)));
            pst = spdk_picard_stats{};
            if (mixed && pc.residual) {
                int N[3] = {X.n_total, Y.n_total, Z.n_total};
                CK(spdk_mixed_histogram(ctx, ws, N, nullptr, 0, 1));
            }
        }
        CK(spdk_ctx_sync(ctx));
        spdk_timing_reset(ctx);
        auto t0 = std::chrono::steady_clock::now();
        double t_io = 0;
        int step0 = n_step;
        while (t < cfg.tlim) {
            if (cfg.max_steps >= 0 && n_step - step0 >= cfg.max_steps) break;
            bool out_due = cfg.outputs && (t + dt >= t_output);
            double next = step(out_due);
            t += dt;
            n_step++;
            dt = next;
            progress_dot();
            if (cfg.outputs) {
                if (t >= t_output) {
                    t_output = t + cfg.dt_output;
                    if ((fused || mixed) && !out_due) {  // keep W_cv current for the writer
                        CK(spdk_primitives(ctx, U_sp, W_sp, s_sp));
                        CK(spdk_transform(ctx, W_sp, W_cv, T_sweep, s_sp, SPDK_M_SP_TO_CV));
                    }
                    CK(spdk_ctx_sync(ctx));
                    auto a = std::chrono::steady_clock::now();
                    write_outputs();
                    t_io += std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
                }
                if (t + dt > t_output) dt = t_output - t;
            }
        }
        CK(spdk_ctx_sync(ctx));
        double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() - t_io;
        if (cfg.final_output) {  // the state after the last step (outside the timed loop)
            if (fused || mixed || cfg.fallback) {
                CK(spdk_primitives(ctx, U_sp, W_sp, s_sp));
                CK(spdk_transform(ctx, W_sp, W_cv, T_sweep, s_sp, SPDK_M_SP_TO_CV));
            }
            CK(spdk_ctx_sync(ctx));
            write_dims(outdir, X, Y, Z);
            write_outputs();
        }
        res->steps = n_step - step0;
        res->dofs = (long long)(X.n * X.n_sp) * (Y.n * Y.n_sp) * (Z.n * Z.n_sp);
        res->evolution_seconds = sec;
        res->dof_updates_per_s = sec > 0 ? (double)res->dofs * res->steps / sec : 0;
        res->t_final = t;
        std::printf("\nevolution: %d steps, %g s (%g ms/step), %g zone-cycles/s\n", res->steps, sec,
                    res->steps ? sec / res->steps * 1e3 : 0.0, res->dof_updates_per_s);
        if (mixed) {
            res->picard_steps = pst.steps;
            res->picard_sweeps = pst.sweeps;
            res->picard_fallbacks = pst.fallbacks;
            std::printf("picard: %lld steps, %.3f sweeps/step, %lld steps with capped elements, last residual %.3e\n",
                        pst.steps,
                        pst.steps ? pst.sweeps / pst.steps : 0.0, pst.fallbacks, pst.last_residual);
            if (pc.residual) {
                int N[3] = {X.n_total, Y.n_total, Z.n_total};
                unsigned long long hist[32];
                CK(spdk_mixed_histogram(ctx, ws, N, hist, 32, 0));
                std::printf("picard histogram (elements by predictor sweeps to tolerance):");
                for (int b = 0; b < 32; b++)
                    if (hist[b]) {
                        if (b == 31) std::printf(" capped:%llu", hist[b]);
                        else std::printf(" %d:%llu", b, hist[b]);
                    }
                std::printf("\n");
                for (int b = 0; b < 32; b++) res->picard_hist[b] = (long long)hist[b];
            }
        }
    }
};

int run_hydro(RunConfig& cfg, const std::string& outdir, int warmup, spdk_run_result* res) {
    HydroRun h(cfg, outdir);
    h.timing = cfg.timing;
    h.setup();
    if (cfg.mode == "train_guess") {
        h.train_guess();
        return 0;
    }
    h.evolve(warmup, res);
    fill_kernel_timing(h.ctx, res);
    return 0;
}

}  // namespace spdk_host

using namespace spdk_host;

extern "C" int spdk_run_argv(int argc, const char* const* argv, const char* outdir, spdk_run_result* res) {
    spdk_run_result local;
    if (!res) res = &local;
    std::memset(res, 0, sizeof(*res));
    try {
        Params pin;
        std::string input;
        std::vector<std::string> overrides;
        for (int n = 0; n < argc; n++) {
            std::string a = argv[n];
            if (a == "-i" && n + 1 < argc) input = argv[++n];
            else if (a.find('/') != std::string::npos && a.find('=') != std::string::npos) overrides.push_back(a);
            else {
                std::fprintf(stderr, "ERROR: unrecognized argument '%s'\n", a.c_str());
                return SPDK_E_CONFIG;
            }
        }
        if (!input.empty()) pin.load_file(input);
        for (auto& o : overrides) pin.apply_override(o);
        RunConfig cfg;
        std::string err = configure(pin, cfg);
        if (!err.empty()) {
            std::fprintf(stderr, "ERROR: %s\n", err.c_str());
            return SPDK_E_CONFIG;
        }
        std::string od = outdir ? std::string(outdir) : std::string("output");
        std::printf("system = %s, ndim = %d, p = %d, N = (%d,%d,%d), integrator = %s, outputs = %s\n", cfg.system.c_str(),
                    cfg.ndim, cfg.p, cfg.act[0] ? cfg.nx[0] : 1, cfg.act[1] ? cfg.nx[1] : 1, cfg.act[2] ? cfg.nx[2] : 1,
                    cfg.integrator == INT_ADER ? "ader" : ("rk" + std::to_string(cfg.rk_order)).c_str(),
                    cfg.outputs ? "on" : "off");
        if (cfg.outputs) {
            std::ofstream dump(output_folder(od) + "parameters.txt");
            pin.dump(dump);
        }
        int warmup = pin.has("job/warmup") ? pin.get<int>("job/warmup") : 0;
        if (cfg.system == "hydro") return run_hydro(cfg, od, warmup, res);
        if (cfg.system == "induction") return run_induction(cfg, od, warmup, res);
        std::fprintf(stderr, "ERROR: unknown system '%s'\n", cfg.system.c_str());
        return SPDK_E_CONFIG;
    } catch (std::exception& e) {
        std::fprintf(stderr, "ERROR: %s\n", e.what());
        spdk::set_error(e.what());
        return SPDK_E_CONFIG;
    }
}
