// Induction integrator: the counterpart of src/induction_ader.hpp (3-D,
// staggered face B, edge-centred E, ADER or SSP-RK, optional FV update with
// trouble detection), extended to runs with z inactive (the SD update of the
// 2-D field loop). Kernels run through the C ABI.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include "driver.hpp"

namespace spdk_host {

#define CK(call) check((call), #call)

namespace {

spdk_sd_shape sd_of(int nt, int nv, const Axis& X, const Axis& Y, const Axis& Z, bool sz, bool sy, bool sx) {
    return spdk_sd_shape{nt, nv, Z.n_total, Y.n_total, X.n_total, sz ? Z.n_fp : Z.n_sp, sy ? Y.n_fp : Y.n_sp,
                         sx ? X.n_fp : X.n_sp};
}
spdk_fv_shape fv_of(int nv, const Axis& X, const Axis& Y, const Axis& Z, bool sz, bool sy, bool sx) {
    return spdk_fv_shape{nv, Z.fv_ncells + (sz ? 1 : 0), Y.fv_ncells + (sy ? 1 : 0), X.fv_ncells + (sx ? 1 : 0)};
}

// A_z of the field loop at a point (src/initial_conditions.cpp:234-249, with
// the loop centre, radius and amplitude as parameters)
double loop_potential(const RunConfig& c, double x, double y) {
    x = x - c.loop_c[0];
    y = y - c.loop_c[1];
    double r = std::sqrt(x * x + y * y);
    return (r <= c.loop_R) ? c.loop_A0 * (c.loop_R - r) : 0;
}

}  // namespace

struct InductionRun {
    RunConfig& cfg;
    std::string outdir;
    Axis X, Y, Z;
    Operators op;
    spdk_ctx* ctx = nullptr;
    DeviceArena mem;
    int T = 1, n_stages = 1, n_output = 0, n_step = 0;
    double rk_a[3] = {0, 0, 0};
    double t = 0, dt = 0, Dt = 0;
    spdk_sd_shape sB[3]{}, sBa[3]{}, sE[3]{}, sB2{};
    spdk_fv_shape sBo[3]{}, sC{}, sEf[3]{};
    double *Bf[3] = {}, *Ba[3] = {}, *B0[3] = {}, *Eep[3] = {}, *Bcf[3] = {}, *Tfp[3] = {};
    double *Bold[3] = {}, *Bnew[3] = {}, *Efv[3] = {}, *alpha[3] = {};
    double *B_old = nullptr, *B_new = nullptr, *troubles = nullptr, *B2 = nullptr;
    bool fused = false;    // fused ADER sweeps (spdk_induction_step_fused) instead of the stages
    bool mixed = false;    // their FP32 form (spdk_induction_step_mixed)
    double* ws = nullptr;  // their workspace
    std::vector<double> host;

    InductionRun(RunConfig& c, const std::string& od) : cfg(c), outdir(od), mixed(c.precision == "mixed") {}
    ~InductionRun() {
        if (ctx) spdk_ctx_sync(ctx);
        mem.release();
        if (ctx) spdk_ctx_destroy(ctx);
    }

    double* zeros(size_t n) { return mem.alloc(ctx, n); }

    void setup() {  // src/induction_ader.hpp:94-272
        int p = cfg.p;
        op.build(p, cfg.integrator == INT_ADER);
        X.build(0, cfg.nx[0], cfg.nx[0], p, 0, cfg.len[0], op.x_fp.data(), cfg.act[0]);
        Y.build(1, cfg.nx[1], cfg.nx[1], p, 0, cfg.len[1], op.x_fp.data(), cfg.act[1]);
        Z.build(2, cfg.nx[2], cfg.nx[2], p, 0, cfg.len[2], op.x_fp.data(), cfg.act[2]);
        if (cfg.outputs) write_dims(outdir, X, Y, Z);
        T = op.T;
        if (cfg.integrator == INT_RK) {
            n_stages = cfg.rk_order;
            if (cfg.rk_order == 2) rk_a[1] = 0.5;
            if (cfg.rk_order == 3) { rk_a[1] = 0.75; rk_a[2] = 1.0 / 3.0; }
        }
        const Axis* ax[3] = {&X, &Y, &Z};
        ctx = make_context(cfg, op, ax);
        spdk_timing_enable(ctx, cfg.timing ? 1 : 0);
        if ((cfg.path == "fused" || mixed) && !cfg.fallback && cfg.integrator == INT_ADER) {
            int Ng[3] = {X.n_total, Y.n_total, Z.n_total};
            size_t b = mixed ? spdk_induction_mixed_workspace_bytes(ctx, Ng) : spdk_induction_fused_workspace_bytes(ctx, Ng);
            if (b > 0) {
                fused = true;
                ws = zeros((b + sizeof(double) - 1) / sizeof(double));
            }
        }
        if (mixed && !fused) throw std::runtime_error(std::string("mixed induction step: ") + spdk_last_error());
        for (int d = 0; d < 3; d++) {
            sB[d] = sd_of(1, 1, X, Y, Z, d == 2, d == 1, d == 0);
            sBa[d] = sB[d];
            sBa[d].nt = T;
            sE[d] = sd_of(T, 8, X, Y, Z, d != 2, d != 1, d != 0);
            Bf[d] = zeros(sd_count(sB[d]));
            if (!fused) {  // the stage sequence's slices and edge arrays
                Ba[d] = zeros(sd_count(sBa[d]));
                Eep[d] = zeros(sd_count(sE[d]));
            }
            if (cfg.integrator == INT_RK) B0[d] = zeros(sd_count(sB[d]));
        }
        sB2 = sd_of(1, 4, X, Y, Z, false, false, false);
        B2 = zeros(sd_count(sB2));
        // vector potential at edge points (src/induction_ader.hpp:176-178): only
        // A_z (z solution points, x and y flux points) is non-zero
        spdk_sd_shape sA[3];
        double* A[3];
        for (int d = 0; d < 3; d++) {
            sA[d] = sd_of(1, 1, X, Y, Z, d != 2, d != 1, d != 0);
            A[d] = zeros(sd_count(sA[d]));
        }
        {
            const spdk_sd_shape& s = sA[2];
            host.assign(sd_count(s), 0.0);
            for (int k = 0; k < s.Nz; k++)
                for (int j = 0; j < s.Ny; j++)
                    for (int i = 0; i < s.Nx; i++)
                        for (int kk = 0; kk < s.nz; kk++)
                            for (int jj = 0; jj < s.ny; jj++)
                                for (int ii = 0; ii < s.nx; ii++) {
                                    size_t o = (((((size_t)k * s.Ny + j) * s.Nx + i) * s.nz + kk) * s.ny + jj) * s.nx + ii;
                                    host[o] = loop_potential(cfg, X.sd_faces.at(i, ii), Y.sd_faces.at(j, jj));
                                }
            CK(spdk_memcpy_h2d(ctx, A[2], host.data(), host.size() * sizeof(double)));
        }
        // B = curl A (src/induction_ader.hpp:185-187)
        CK(spdk_B_curl(ctx, A[1], sA[1], A[2], sA[2], Bf[0], sB[0], Y.h, Z.h, 0));
        CK(spdk_B_curl(ctx, A[2], sA[2], A[0], sA[0], Bf[1], sB[1], Z.h, X.h, 1));
        CK(spdk_B_curl(ctx, A[0], sA[0], A[1], sA[1], Bf[2], sB[2], X.h, Y.h, 2));
        if (cfg.fallback) {
            sC = fv_of(4, X, Y, Z, false, false, false);
            B_old = zeros(fv_count(sC));
            B_new = zeros(fv_count(sC));
            troubles = zeros(fv_count(sC));
            for (int d = 0; d < 3; d++) {
                Bcf[d] = zeros(sd_count(sB[d]));
                Tfp[d] = zeros(sd_count(sB[d]));
                sBo[d] = fv_of(1, X, Y, Z, d == 2, d == 1, d == 0);
                Bold[d] = zeros(fv_count(sBo[d]));
                Bnew[d] = zeros(fv_count(sBo[d]));
                sEf[d] = fv_of(1, X, Y, Z, d != 2, d != 1, d != 0);
                Efv[d] = zeros(fv_count(sEf[d]));
                alpha[d] = zeros(fv_count(sC));
            }
        }
        // src/induction.cpp:52-92 and src/induction_ader.hpp:254-268 (sums and
        // minima over the active axes; all three in the reference's 3-D)
        double vmax = 0, dx_min = 1;
        for (int d = 0; d < 3; d++)
            if (cfg.act[d]) {
                vmax += std::fabs(cfg.vel[d]);
                dx_min = std::min(dx_min, ax[d]->h);
            }
        dt = cfg.cfl * dx_min / vmax / sB2.nx;
        if (cfg.eta > 0.0) Dt = 0.5 * cfg.cfl * (X.h * X.h / cfg.eta) / (cfg.ndim) / (cfg.ndim) / (p + 1) / (p + 1);
        else Dt = 1E6;
        std::printf("dx = %.17g dt_nu = %.17g dt_v = %.17g\n", X.h, Dt, dt);
        Dt = dt < Dt ? dt : Dt;
        if (cfg.outputs) write_outputs();
        else diagnostics();
    }

    // B2_cv, then max |div B| and the magnetic energy (not in the reference),
    // printed; also at the start and end of a run without outputs
    void diagnostics() {
        CK(spdk_B2_cv(ctx, B2, sB2, Bf[0], sB[0], Bf[1], sB[1], Bf[2], sB[2]));
        double diag[2];
        CK(spdk_induction_diagnostics(ctx, Bf[0], sB[0], Bf[1], sB[1], Bf[2], sB[2], B2, sB2, diag));
        std::printf("diagnostics: t = %.17g max|div B| = %.17g magnetic energy = %.17g\n", t, diag[0], diag[1]);
    }

    void write_outputs() {  // src/induction_ader.hpp:487-494
        std::string f = output_folder(outdir);
        std::printf("\nOUTPUT %d\n", n_output);
        diagnostics();
        if (cfg.fallback) {
            host.resize(fv_count(sC));
            CK(spdk_memcpy_d2h(ctx, host.data(), troubles, host.size() * sizeof(double)));
            write_raw(f + fv_name("troubles", sC.Nx, n_output), host.data(), host.size());
        }
        host.resize(sd_count(sB2));
        CK(spdk_memcpy_d2h(ctx, host.data(), B2, host.size() * sizeof(double)));
        write_raw(f + sd_name("B2_cv", sB2.Nx, sB2.nx, n_output), host.data(), host.size());
        n_output++;
    }

    void solve_E() {  // src/induction_ader.hpp:332-341, :386-448
        CK(spdk_induction_E(ctx, Eep[2], sE[2], Ba[0], sBa[0], Ba[1], sBa[1], 2));
        CK(spdk_induction_E(ctx, Eep[1], sE[1], Ba[2], sBa[2], Ba[0], sBa[0], 1));
        CK(spdk_induction_E(ctx, Eep[0], sE[0], Ba[1], sBa[1], Ba[2], sBa[2], 0));
        // trace fills and face solves along the active axes only
        static const int fill[3][2] = {{1, 2}, {0, 2}, {0, 1}};
        for (int d = 0; d < 3; d++)
            if (cfg.act[d]) {
                CK(spdk_fill_sd_traces(ctx, Eep[fill[d][0]], sE[fill[d][0]], d));
                CK(spdk_fill_sd_traces(ctx, Eep[fill[d][1]], sE[fill[d][1]], d));
            }
        // along axis d: E_{d+1} upwinded by variable 4, E_{d+2} by variable 3
        for (int d = 0; d < 3; d++)
            if (cfg.act[d]) {
                int e1 = (d + 1) % 3, e2 = (d + 2) % 3;
                CK(spdk_E_riemann(ctx, Eep[e1], sE[e1], d, 4));
                CK(spdk_E_riemann(ctx, Eep[e2], sE[e2], d, 3));
            }
    }
    void update_prediction() {  // src/induction_ader.hpp:471-479
        CK(spdk_B_predict(ctx, Bf[0], Ba[0], sBa[0], Eep[1], sE[1], Eep[2], sE[2], Y.h, Z.h, dt, 0));
        CK(spdk_B_predict(ctx, Bf[1], Ba[1], sBa[1], Eep[2], sE[2], Eep[0], sE[0], Z.h, X.h, dt, 1));
        CK(spdk_B_predict(ctx, Bf[2], Ba[2], sBa[2], Eep[0], sE[0], Eep[1], sE[1], X.h, Y.h, dt, 2));
    }
    void update_solution() {  // src/induction_ader.hpp:481-485
        CK(spdk_B_correct(ctx, Bf[0], sB[0], Eep[1], sE[1], Eep[2], sE[2], Y.h, Z.h, dt, 0));
        CK(spdk_B_correct(ctx, Bf[1], sB[1], Eep[2], sE[2], Eep[0], sE[0], Z.h, X.h, dt, 1));
        CK(spdk_B_correct(ctx, Bf[2], sB[2], Eep[0], sE[0], Eep[1], sE[1], X.h, Y.h, dt, 2));
    }
    void transform_2d(const double* A, double* B, double* Tm, const spdk_sd_shape& s, int mat, int dim) {
        // src/transforms.cpp:226-249: transverse sweeps (both active in 3-D)
        int dim1 = dim == 0 ? 1 : (dim == 1 ? 2 : 0), dim2 = dim1 == 0 ? 1 : (dim1 == 1 ? 2 : 0);
        CK(spdk_transform_1d(ctx, A, s, Tm, s, mat, dim1, -1));
        CK(spdk_transform_1d(ctx, Tm, s, B, s, mat, dim2, -1));
    }
    void fv_update() {  // src/induction_ader.hpp:534-603
        for (int d = 0; d < 3; d++) transform_2d(Bf[d], Bcf[d], Tfp[d], sB[d], SPDK_M_SP_TO_CV, d);
        CK(spdk_B_cv_from_cf(ctx, B_old, sC, Bcf[0], sB[0], Bcf[1], sB[1], Bcf[2], sB[2]));
        for (int a = 0; a < T; a++) {
            for (int d = 0; d < 3; d++) CK(spdk_edge_integral(ctx, Eep[d], sE[d], Efv[d], sEf[d], a, d));
            CK(spdk_fv_update_B(ctx, Bnew[0], Bold[0], sBo[0], Bcf[0], sB[0], Efv[1], sEf[1], Efv[2], sEf[2], a, dt, 0, 1));
            CK(spdk_fv_update_B(ctx, Bnew[1], Bold[1], sBo[1], Bcf[1], sB[1], Efv[2], sEf[2], Efv[0], sEf[0], a, dt, 1, 1));
            CK(spdk_fv_update_B(ctx, Bnew[2], Bold[2], sBo[2], Bcf[2], sB[2], Efv[0], sEf[0], Efv[1], sEf[1], a, dt, 2, 1));
            CK(spdk_B_cv_from_cf(ctx, B_new, sC, Bcf[0], sB[0], Bcf[1], sB[1], Bcf[2], sB[2]));
            CK(spdk_fill_fv_ghosts_aligned(ctx, B_old, sC, 0, 0));
            for (int d = 0; d < 3; d++) CK(spdk_fill_fv_ghosts_aligned(ctx, Bold[d], sBo[d], 1, d));
            CK(spdk_detect_troubles(ctx, B_new, B_old, troubles, alpha[0], alpha[1], alpha[2], sC, 0, 0b0111));
        }
        for (int d = 0; d < 3; d++) transform_2d(Bcf[d], Bf[d], Tfp[d], sB[d], SPDK_M_CV_TO_SP, d);
    }
    void ader_step() {  // src/induction_ader.hpp:343-359
        if (fused) {
            if (mixed) CK(spdk_induction_step_mixed(ctx, Bf[0], sB[0], Bf[1], sB[1], Bf[2], sB[2], ws, dt));
            else CK(spdk_induction_step_fused(ctx, Bf[0], sB[0], Bf[1], sB[1], Bf[2], sB[2], ws, dt));
            return;
        }
        for (int d = 0; d < 3; d++) CK(spdk_copy_ader(ctx, Bf[d], Ba[d], sBa[d]));
        for (int a = 0; a < T; a++) {
            solve_E();
            if (a < T - 1) update_prediction();
        }
        if (cfg.fallback) fv_update();
        else update_solution();
    }
    void rk_step() {  // src/induction_ader.hpp:365-384
        for (int d = 0; d < 3; d++) CK(spdk_memcpy_d2d(ctx, B0[d], Bf[d], sd_count(sB[d]) * sizeof(double)));
        for (int s = 0; s < n_stages; s++) {
            for (int d = 0; d < 3; d++) CK(spdk_copy_ader(ctx, Bf[d], Ba[d], sBa[d]));
            solve_E();
            if (cfg.fallback) fv_update();
            else update_solution();
            if (rk_a[s] > 0)
                for (int d = 0; d < 3; d++) CK(spdk_rk_combine(ctx, Bf[d], B0[d], sB[d], rk_a[s]));
        }
    }
    void evolve(int warmup, spdk_run_result* res) {  // src/induction_ader.hpp:274-327
        dt = Dt;
        double t_output = cfg.dt_output;
        if (warmup > 0) {  // untimed steps; B is restored afterwards
            std::vector<double*> save;
            for (int d = 0; d < 3; d++) {
                save.push_back(zeros(sd_count(sB[d])));
                CK(spdk_memcpy_d2d(ctx, save[d], Bf[d], sd_count(sB[d]) * sizeof(double)));
            }
            for (int w = 0; w < warmup; w++) {
                if (cfg.integrator == INT_RK) rk_step();
                else ader_step();
            }
            for (int d = 0; d < 3; d++) CK(spdk_memcpy_d2d(ctx, Bf[d], save[d], sd_count(sB[d]) * sizeof(double)));
            CK(spdk_ctx_sync(ctx));
            spdk_timing_reset(ctx);
        }
        CK(spdk_ctx_sync(ctx));
        auto t0 = std::chrono::steady_clock::now();
        double t_io = 0;
        int step0 = n_step;
        while (t < cfg.tlim) {
            if (cfg.max_steps >= 0 && n_step - step0 >= cfg.max_steps) break;
            if (cfg.integrator == INT_RK) rk_step();
            else ader_step();
            t += dt;
            n_step++;
            dt = Dt;
            if (cfg.outputs) {
                if (t >= t_output) {
                    t_output = t + cfg.dt_output;
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
            write_dims(outdir, X, Y, Z);
            write_outputs();
        } else if (!cfg.outputs) {
            diagnostics();
        }
        res->steps = n_step - step0;
        res->dofs = (long long)(X.n * X.n_sp) * (Y.n * Y.n_sp) * (Z.n * Z.n_sp);
        res->evolution_seconds = sec;
        res->dof_updates_per_s = sec > 0 ? (double)res->dofs * res->steps / sec : 0;
        res->t_final = t;
        std::printf("\nevolution: %d steps, %g s (%g ms/step), %g zone-cycles/s\n", res->steps, sec,
                    res->steps ? sec / res->steps * 1e3 : 0.0, res->dof_updates_per_s);
    }
};

int run_induction(RunConfig& cfg, const std::string& outdir, int warmup, spdk_run_result* res) {
    // the reference requires 3 active dimensions; runs with x and y active
    // and z inactive are this build's extension (SD update only)
    if (!(cfg.ndim == 3 || (cfg.ndim == 2 && !cfg.act[2]))) {
        std::fprintf(stderr, "ERROR: the induction solver requires 3 active dimensions, or x and y active\n");
        return SPDK_E_CONFIG;
    }
    if (cfg.ndim == 2 && cfg.fallback) {
        std::fprintf(stderr, "ERROR: the induction FV update requires 3 active dimensions\n");
        return SPDK_E_CONFIG;
    }
    InductionRun r(cfg, outdir);
    r.setup();
    std::printf("induction path: %s\n", r.fused ? (r.mixed ? "fused mixed" : "fused") : "stages");
    r.evolve(warmup, res);
    fill_kernel_timing(r.ctx, res);
    return 0;
}

}  // namespace spdk_host
