// Fused element-tile path for the hydro SD step (plain FP64).
//
// One Picard sweep of ADER_step / one RK stage (src/hydro_ader.hpp:291-325)
// is one kernel over element tiles. A workgroup owns EPB whole elements and
// keeps their state, flux-point values and fluxes in LDS; the only data that
// crosses elements are the face traces (the solution interpolated to the
// first and last flux point of every pencil), which each sweep writes for the
// next one. The Riemann flux of a face is evaluated by both elements that
// share it, from the same two traces, so both see identical bits.
//
// The arithmetic of every value (interpolation, Euler flux, LLF, the
// x/y/z-interleaved flux divergence, predictor and corrector sums, cv
// transform, dt) is performed in the same order as the stage kernels and the
// reference, so this path is bit-identical to the stage path.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include "ctx.hpp"
#include "physics.hpp"

using namespace spdk;

namespace {

constexpr int kMaxT = 8;

template <int N, int DIM>
struct Tile {
    static constexpr int M = N + 1;                                // flux points per axis
    static constexpr int NT = DIM == 1 ? 1 : (DIM == 2 ? N : N * N);  // points per element face
    static constexpr int TPE = N * NT;                             // solution points per element
    static constexpr int EPB = TPE >= 256 ? 1 : 256 / TPE;         // elements per workgroup
    static constexpr int BLOCK = ((EPB * TPE + 63) / 64) * 64;
    static constexpr int FSZ = M * NT;                             // flux points per axis per element
    // LDS per element (doubles): state + fluxes of every axis
    static constexpr int LDS_U = NVP * TPE;
    static constexpr int LDS_F = DIM * NVP * FSZ;
    // + 4 doubles: the element areas of a block start 8 banks apart (reads
    // across elements at one in-element offset do not share a bank)
    static constexpr int LDS_ELEM = LDS_U + LDS_F + 4;
    static_assert(BLOCK == 256, "tile workgroups are 256 threads");
    static_assert(LDS_F >= LDS_U, "cv ping-pong reuses the flux area");
};

struct SweepArgs {
    const double* U_sp;   // (1,6,E,TPE) solution (reference layout)
    double* U_sp_out;     // corrector output (same array)
    const double* U_in;   // (T,5,E,TPE) predictor state in (IN_ADER)
    double* U_out;        // (T,5,E,TPE) predictor state out (OUT_PRED)
    const double* tr_in[3];  // per axis (Tin,5,E,2,NT); Tin = 1 when traces are slice-independent
    double* tr_out[3];       // per axis (T,5,E,2,NT)
    double* U0;           // (5,E,TPE) RK step-start state
    double* W_cv;         // optional (1,6,E,TPE)
    double* F_x;          // optional (T,6,E,NT,M) x flux-point fluxes
    double* F_out[3];     // flux-only sweep: per axis (T,out_nv,E,FSZ) flux-point fluxes
    double* part;         // per-block dt minima
    const double* Dm;     // dfp_to_sp (N x M)
    const double* S2F;    // sp_to_fp  (M x N)
    const double* S2C;    // sp_to_cv  (N x N)
    const double* inv;    // invader (T x T)
    const double* w;      // time weights (T)
    int Ne[3];            // elements per axis incl. ghosts (x,y,z)
    int act[3], bc[3];
    int E;
    double h[3];
    double dt, gm, cfl, rk_a;
    int tin_slices;       // slices present in tr_in (1 or T)
    int in_nv;            // variables per slice of U_in
    int out_nv;           // variables per slice of F_out (slots past the physical ones get zero)
    double* Ffv[kMaxT][3];  // flux-only sweep: sub-cell face fluxes of slice t along axis d (FV layout)
    int fvn[3];           // FV cells per axis incl. ghost layers (an axis' face array has one more along it)
    int fgh[3];           // FV ghost layers per axis
    int flux_only;        // write the flux-point fluxes of every slice and axis; no divergence
    int in_ader;          // state from U_in (else U_sp broadcast to every slice)
    int out_pred;         // write predictor U_out + tr_out, else corrector
    int rk_first;         // RK: write U0 = U_sp
    int want_dt;          // corrector: compute cv primitives and dt partials
    double g[3];          // constant gravity (hydro/g1..g3)
    int grav;             // any component nonzero
};

__device__ __forceinline__ int fidx(int d, int ll, int q, int N, int NT, int M) {
    // LDS index of flux point ll of pencil q along axis d (see header comment)
    if (d == 0) return q * M + ll;
    if (d == 1) return ((q / N) * M + ll) * N + (q % N);
    return ll * NT + q;
}

// Solution-point index of position ll along axis d of pencil q.
template <int N, int DIM>
__device__ __forceinline__ int spidx(int d, int ll, int q) {
    if (DIM == 1) return ll;
    if (DIM == 2) return d == 0 ? q * N + ll : ll * N + q;
    if (d == 0) return q * N + ll;                          // q = kk*N + jj
    if (d == 1) return ((q / N) * N + ll) * N + (q % N);   // q = kk*N + ii
    return ll * N * N + q;                                  // q = jj*N + ii
}

// face_integral of slice t for every active axis (src/transforms.cpp:378-481)
// from the element's flux-point fluxes in LDS (Fs, the fidx layout of each
// axis): each element writes the sub-cell faces of its cells' lower sides
// along d (the right ghost along d writes the domain's last face), so each
// FV face is written once, by the element whose copy face_integral reads.
// With both transverse axes active, the dim1 sweep u(f, o1, nn) = sum_ll
// F(f, ll, nn) S(o1, ll) goes through `scr` (the element's state area, free
// after the pencils) and the dim2 sum over nn of u S(o2, nn) is written;
// otherwise the single contraction of face_integral_ref. Sums in the
// reference's order, so the faces equal the stage kernels' bit for bit.
__device__ __forceinline__ int el_of(int tid, int tpe) { return tid / tpe; }

// Element block of a face-integral sweep workgroup: workgroups go to the 8
// XCDs in turn and XCD x takes the x-th contiguous eighth of the blocks, so
// the face rows that neighbouring blocks share (rows are not line-aligned)
// are written through one L2. The grid is padded to a multiple of 8.
__device__ __forceinline__ int face_sweep_block() {
    return (int)((blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8);
}

template <int N, int DIM>
__device__ __forceinline__ void fv_faces(const SweepArgs& a, double* scr, const double* Fs, int t, int s, bool live,
                                         int ei, int ej, int ek) {
    using G = Tile<N, DIM>;
    constexpr int M = G::M, NT = G::NT, TPE = G::TPE, FSZ = G::FSZ;
    const double* S = a.S2C;
    const int ec[3] = {ei, ej, ek};
    const int o[3] = {s % N, DIM >= 2 ? (s / N) % N : 0, DIM == 3 ? s / (N * N) : 0};  // x fastest
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        bool own = live;
#pragma unroll
        for (int x = 0; x < DIM; x++) own = own && ec[x] >= 1 && ec[x] <= (x == d ? a.Ne[x] - 1 : a.Ne[x] - 2);
        const int f = o[d];
        const bool put = own && (ec[d] < a.Ne[d] - 1 || f == 0);
        const int dim1 = d == 0 ? 1 : (d == 1 ? 2 : 0), dim2 = dim1 == 0 ? 1 : (dim1 == 1 ? 2 : 0);
        const int nx = a.fvn[0] + (d == 0), ny = a.fvn[1] + (d == 1), nz = a.fvn[2] + (d == 2);
        const int I = o[0] + a.fgh[0] + (ei - 1) * N;
        const int J = DIM >= 2 ? o[1] + a.fgh[1] + (ej - 1) * N : 0;
        const int K = DIM == 3 ? o[2] + a.fgh[2] + (ek - 1) * N : 0;
        double* out = a.Ffv[t][d] + ((size_t)K * ny + J) * nx + I;
        const size_t vstride = (size_t)nz * ny * nx;
        if (DIM == 3) {
            {  // dim1 sweep: thread s -> (f, o1, nn)
                const int f1 = s / (N * N), o1 = (s / N) % N, nn = s % N;
                if (own)
#pragma unroll
                    for (int v = 0; v < NVP; v++) {
                        double u = 0;
#pragma unroll
                        for (int ll = 0; ll < N; ll++) {
                            const int q = d == 1 ? ll * N + nn : nn * N + ll;
                            u += Fs[(d * NVP + v) * FSZ + fidx(d, f1, q, N, NT, M)] * S[o1 * N + ll];
                        }
                        scr[v * TPE + s] = u;
                    }
            }
            __syncthreads();
            double accv[NVP];
            if (own) {
                const int o1 = o[dim1], o2 = o[dim2];
#pragma unroll
                for (int v = 0; v < NVP; v++) {
                    double acc = 0;
#pragma unroll
                    for (int nn = 0; nn < N; nn++) acc += scr[v * TPE + (f * N + o1) * N + nn] * S[o2 * N + nn];
                    accv[v] = acc;
                }
            }
            if (G::EPB == 4 && N == 4) {
                // staged: the block's four elements (consecutive along x) write
                // 16 consecutive faces per row, whole 128-byte rows
                __syncthreads();
                if (own)
#pragma unroll
                    for (int v = 0; v < NVP; v++) scr[v * TPE + s] = accv[v];
                __syncthreads();
                const int tid = threadIdx.x, xx = tid % 16, row = tid / 16;
                const int e2l = xx / N, ox = xx % N, oy = row % N, oz = row / N;
                const int e2 = face_sweep_block() * G::EPB + e2l;
                if (e2 < a.E) {
                    const int ec2[3] = {e2 % a.Ne[0], (e2 / a.Ne[0]) % a.Ne[1], e2 / (a.Ne[0] * a.Ne[1])};
                    bool own2 = true;
#pragma unroll
                    for (int x = 0; x < DIM; x++)
                        own2 = own2 && ec2[x] >= 1 && ec2[x] <= (x == d ? a.Ne[x] - 1 : a.Ne[x] - 2);
                    const int o2v[3] = {ox, oy, oz};
                    if (own2 && (ec2[d] < a.Ne[d] - 1 || o2v[d] == 0)) {
                        const int I2 = ox + a.fgh[0] + (ec2[0] - 1) * N, J2 = oy + a.fgh[1] + (ec2[1] - 1) * N;
                        const int K2 = oz + a.fgh[2] + (ec2[2] - 1) * N;
                        double* out2 = a.Ffv[t][d] + ((size_t)K2 * ny + J2) * nx + I2;
                        const double* sc2 = scr + (long)(e2l - el_of(threadIdx.x, TPE)) * G::LDS_ELEM;
#pragma unroll
                        for (int v = 0; v < NVP; v++) out2[(size_t)v * vstride] = sc2[v * TPE + ox + N * (oy + N * oz)];
                    }
                }
            } else if (put) {
#pragma unroll
                for (int v = 0; v < NVP; v++) out[(size_t)v * vstride] = accv[v];
            }
            __syncthreads();
        } else if (put) {  // at most one transverse axis active: face_integral_ref
#pragma unroll
            for (int v = 0; v < NVP; v++) {
                double acc = 0;
                if (DIM == 1) {
                    acc += Fs[v * FSZ + fidx(0, f, 0, N, NT, M)];
                } else if (d == 0) {  // dim1 = y active, dim2 = z inactive
#pragma unroll
                    for (int ll = 0; ll < N; ll++) {
                        double x = Fs[v * FSZ + fidx(0, f, ll, N, NT, M)];
                        x *= S[o[1] * N + ll];
                        acc += x;
                    }
                } else {  // d = 1: dim1 = z inactive, dim2 = x active
#pragma unroll
                    for (int nn = 0; nn < N; nn++) {
                        double x = Fs[(NVP + v) * FSZ + fidx(1, f, nn, N, NT, M)];
                        x *= S[o[0] * N + nn];
                        acc += x;
                    }
                }
                out[(size_t)v * vstride] = acc;
            }
        }
    }
}

// FVF: the flux-only sweep that also forms the face integrals (fv_faces),
// compiled apart so that the predictor and corrector sweeps keep their
// register budget (2 waves per SIMD at 3-D p=3).
template <int N, int DIM, int T, bool FVF>
__global__ void __launch_bounds__(256) k_sweep(SweepArgs a) {
    using G = Tile<N, DIM>;
    constexpr int M = G::M, NT = G::NT, TPE = G::TPE, EPB = G::EPB;
    extern __shared__ double lds[];
    const int tid = threadIdx.x;
    const int el = tid / TPE;      // element slot in the tile
    const int s = tid % TPE;       // solution point within the element
    const int bx = FVF ? face_sweep_block() : (int)blockIdx.x;
    if (FVF && bx * EPB >= a.E) return;  // padding of the XCD-ordered grid (uniform per block)
    const int e = bx * EPB + el;
    const bool live = el < EPB && e < a.E;
    double* us = lds + (el < EPB ? el : 0) * G::LDS_ELEM;
    double* Fs = us + G::LDS_U;
    const size_t E = (size_t)a.E;
    const size_t slabE = E * TPE;  // one (t,v) plane of a state array

    // element coordinates
    int ei = 0, ej = 0, ek = 0;
    if (live) {
        ei = e % a.Ne[0];
        ej = (e / a.Ne[0]) % a.Ne[1];
        ek = e / (a.Ne[0] * a.Ne[1]);
    }
    const int ii = s % N, jj = DIM >= 2 ? (s / N) % N : 0, kk = DIM == 3 ? s / (N * N) : 0;

    // derivative rows of this solution point
    double Dx[M], Dy[M], Dz[M];
#pragma unroll
    for (int l = 0; l < M; l++) {
        Dx[l] = a.Dm[ii * M + l];
        Dy[l] = DIM >= 2 ? a.Dm[jj * M + l] : 0.0;
        Dz[l] = DIM == 3 ? a.Dm[kk * M + l] : 0.0;
    }
    double u_old[NVP];
#pragma unroll
    for (int v = 0; v < NVP; v++) u_old[v] = live && !a.flux_only ? a.U_sp[(size_t)v * slabE + (size_t)e * TPE + s] : 0.0;

    double dudt[T][NVP];
    const int nslices = a.in_ader ? T : 1;  // a broadcast state has identical slices

#pragma unroll
    for (int t = 0; t < T; t++) {
        if (t >= nslices) break;
        // ---- stage the slice in LDS
        if (live) {
#pragma unroll
            for (int v = 0; v < NVP; v++)
                us[v * TPE + s] = a.in_ader ? a.U_in[((size_t)t * a.in_nv + v) * slabE + (size_t)e * TPE + s] : u_old[v];
        }
        __syncthreads();
        // ---- pencils: interpolation to flux points, fluxes, face Riemann solves
        // (DIM * NT pencils; more than the element's points only for p = 1 in 3-D)
        for (int task = s; live && task < DIM * NT; task += TPE) {
            const int d = task / NT, q = task % NT;
            const int ed = d == 0 ? ei : (d == 1 ? ej : ek);
            const int Nd = a.Ne[d];
            const size_t stride = d == 0 ? 1 : (d == 1 ? (size_t)a.Ne[0] : (size_t)a.Ne[0] * a.Ne[1]);
            const size_t ebase = (size_t)e - (size_t)ed * stride;  // element with ed = 0 on this line
            const int ts = a.tin_slices > 1 ? t : 0;
            const double* tr = a.tr_in[d];
            auto trace = [&](int ed_src, int side, int v) -> double {
                size_t src = ebase + (size_t)ed_src * stride;
                return tr[((((size_t)ts * NVP + v) * E + src) * 2 + side) * NT + q];
            };
            double uf[M][NVP];
#pragma unroll
            for (int v = 0; v < NVP; v++) {
                double p[N];
#pragma unroll
                for (int l = 0; l < N; l++) p[l] = us[v * TPE + spidx<N, DIM>(d, l, q)];
#pragma unroll
                for (int f = 0; f < M; f++) {
                    double x = 0;
#pragma unroll
                    for (int l = 0; l < N; l++) x += p[l] * a.S2F[f * N + l];
                    uf[f][v] = x;
                }
            }
            double* Fd = Fs + d * NVP * G::FSZ;
            // interior flux points
#pragma unroll
            for (int f = 1; f < M - 1; f++) {
                double fl[NVP];
                if (d == 0) state_flux<0>(uf[f], fl, a.gm);
                else if (d == 1) state_flux<1>(uf[f], fl, a.gm);
                else state_flux<2>(uf[f], fl, a.gm);
#pragma unroll
                for (int v = 0; v < NVP; v++) Fd[v * G::FSZ + fidx(d, f, q, N, NT, M)] = fl[v];
            }
            // the two faces: ghost values follow src/boundary.cpp:48-113
            const int type = a.bc[d];
            double bl[NVP], br[NVP];  // values the trace fill writes at (0, M-1) and (Nd-1, 0)
#pragma unroll
            for (int v = 0; v < NVP; v++) {
                double sg = (type == SPDK_BC_REFLECTIVE && v == 1 + d) ? -1.0 : 1.0;
                if (type == SPDK_BC_PERIODIC) {
                    bl[v] = (ed == 1 || ed == 0) ? trace(Nd - 2, 1, v) : 0.0;
                    br[v] = (ed == Nd - 2 || ed == Nd - 1) ? trace(1, 0, v) : 0.0;
                } else {
                    bl[v] = (ed == 1 || ed == 0) ? sg * trace(1, 0, v) : 0.0;
                    br[v] = (ed == Nd - 2 || ed == Nd - 1) ? sg * trace(Nd - 2, 1, v) : 0.0;
                }
            }
            // the face integrals read faces 0..M-2 only: the upper face's
            // flux is formed when the flux-point array is written as well
            const int nsides = (FVF && !a.F_out[d]) ? 1 : 2;
            for (int side = 0; side < nsides; side++) {
                double fl[NVP];
                const int f = side ? M - 1 : 0;
                bool raw = side ? (ed == Nd - 1) : (ed == 0);
                if (raw) {
                    if (d == 0) state_flux<0>(uf[f], fl, a.gm);
                    else if (d == 1) state_flux<1>(uf[f], fl, a.gm);
                    else state_flux<2>(uf[f], fl, a.gm);
                } else {
                    double uL[NVP], uR[NVP];
#pragma unroll
                    for (int v = 0; v < NVP; v++) {
                        if (side == 0) {
                            uL[v] = (ed - 1 == 0) ? bl[v] : trace(ed - 1, 1, v);
                            uR[v] = (ed == Nd - 1) ? br[v] : uf[0][v];
                        } else {
                            uL[v] = (ed == 0) ? bl[v] : uf[M - 1][v];
                            uR[v] = (ed + 1 == Nd - 1) ? br[v] : trace(ed + 1, 0, v);
                        }
                    }
                    if (d == 0) riemann_llf<0>(uL, uR, fl, a.gm);
                    else if (d == 1) riemann_llf<1>(uL, uR, fl, a.gm);
                    else riemann_llf<2>(uL, uR, fl, a.gm);
                }
#pragma unroll
                for (int v = 0; v < NVP; v++) Fd[v * G::FSZ + fidx(d, f, q, N, NT, M)] = fl[v];
            }
        }
        __syncthreads();
        if (FVF || a.flux_only) {  // the fluxes of slice t in the stage flux-point layout (src/hydro_ader.hpp:278-289)
            if (FVF) fv_faces<N, DIM>(a, us, Fs, t, s, live, ei, ej, ek);
            if (live) {
#pragma unroll
                for (int d = 0; d < DIM; d++)
                    for (int idx = s; a.F_out[d] && idx < G::FSZ; idx += TPE) {
                        double* o = a.F_out[d] + ((size_t)t * a.out_nv * E + e) * G::FSZ + idx;
#pragma unroll
                        for (int v = 0; v < NVP; v++) o[(size_t)v * E * G::FSZ] = Fs[(d * NVP + v) * G::FSZ + idx];
                        for (int v = NVP; v < a.out_nv; v++) o[(size_t)v * E * G::FSZ] = 0.0;
                    }
            }
            __syncthreads();
            continue;
        }
        // ---- flux divergence at the solution points (src/transforms.cpp:301-307)
        if (live) {
            const int qx = DIM == 3 ? kk * N + jj : jj;
            const int qy = DIM == 3 ? kk * N + ii : ii;
            const int qz = jj * N + ii;
#pragma unroll
            for (int v = 0; v < NVP; v++) {
                double acc = 0;
                const double* Fx = Fs + v * G::FSZ;
                const double* Fy = Fs + (NVP + v) * G::FSZ;
                const double* Fz = Fs + (2 * NVP + v) * G::FSZ;
#pragma unroll
                for (int l = 0; l < M; l++) {
                    acc += Fx[fidx(0, l, qx, N, NT, M)] * Dx[l] / a.h[0];
                    if (DIM >= 2) acc += Fy[fidx(1, l, qy, N, NT, M)] * Dy[l] / a.h[1];
                    if (DIM == 3) acc += Fz[fidx(2, l, qz, N, NT, M)] * Dz[l] / a.h[2];
                }
                dudt[t][v] = acc;
            }
            if (a.grav && !a.out_pred) {
                // update_solution's gravity source (src/transforms.cpp:17-27,
                // :364-367) from the ADER slice staged in LDS: rho g and m . g
                const double rho = us[s], m1 = us[TPE + s], m2 = us[2 * TPE + s], m3 = us[3 * TPE + s];
                dudt[t][1] -= rho * a.g[0];
                dudt[t][2] -= rho * a.g[1];
                dudt[t][3] -= rho * a.g[2];
                dudt[t][4] -= m1 * a.g[0] + m2 * a.g[1] + m3 * a.g[2];
            }
            if (a.F_x && !a.out_pred) {  // x flux-point fluxes of the final sweep (outputs)
                for (int idx = s; idx < G::FSZ; idx += TPE) {
                    size_t o = (((size_t)t * NV) * E + e) * G::FSZ + idx;
#pragma unroll
                    for (int v = 0; v < NVP; v++) a.F_x[o + (size_t)v * E * G::FSZ] = Fs[v * G::FSZ + idx];
                    a.F_x[o + (size_t)NVP * E * G::FSZ] = 0.0;
                }
            }
        }
        __syncthreads();
    }
    if (FVF || a.flux_only) return;
    if (!a.in_ader) {
#pragma unroll
        for (int t = 1; t < T; t++)
#pragma unroll
            for (int v = 0; v < NVP; v++) dudt[t][v] = dudt[0][v];
    }

    if (a.out_pred && a.grav && live) {
        // update_prediction's gravity source (src/transforms.cpp:17-27,
        // :309-313). That kernel writes U_ader variable by variable, and the
        // source reads U_ader back: the momenta's source takes the density
        // just written, the energy's the momenta just written. Those values
        // are formed here exactly as the predictor below forms them.
        double rn[T];
#pragma unroll
        for (int t = 0; t < T; t++) {
            double du = 0;
#pragma unroll
            for (int l = 0; l < T; l++) du += dudt[l][0] * a.inv[t * T + l] * a.w[l] * a.dt;
            rn[t] = u_old[0] - du;
        }
        double mn[T][3];
#pragma unroll
        for (int v = 1; v <= 3; v++) {
#pragma unroll
            for (int t = 0; t < T; t++) dudt[t][v] -= rn[t] * a.g[v - 1];
#pragma unroll
            for (int t = 0; t < T; t++) {
                double du = 0;
#pragma unroll
                for (int l = 0; l < T; l++) du += dudt[l][v] * a.inv[t * T + l] * a.w[l] * a.dt;
                mn[t][v - 1] = u_old[v] - du;
            }
        }
#pragma unroll
        for (int t = 0; t < T; t++) dudt[t][4] -= mn[t][0] * a.g[0] + mn[t][1] * a.g[1] + mn[t][2] * a.g[2];
    }
    if (a.out_pred) {
        // ---- predictor (src/transforms.cpp:316-322) and traces of the new slices
#pragma unroll
        for (int t = 0; t < T; t++) {
            if (live) {
#pragma unroll
                for (int v = 0; v < NVP; v++) {
                    double du = 0;
#pragma unroll
                    for (int l = 0; l < T; l++) du += dudt[l][v] * a.inv[t * T + l] * a.w[l] * a.dt;
                    double un = u_old[v] - du;
                    a.U_out[((size_t)t * NVP + v) * slabE + (size_t)e * TPE + s] = un;
                    us[v * TPE + s] = un;
                }
            }
            __syncthreads();
            for (int task = s; live && task < DIM * NT; task += TPE) {
                const int d = task / NT, q = task % NT;
#pragma unroll
                for (int v = 0; v < NVP; v++) {
                    double p[N];
#pragma unroll
                    for (int l = 0; l < N; l++) p[l] = us[v * TPE + spidx<N, DIM>(d, l, q)];
                    double x0 = 0, x1 = 0;
#pragma unroll
                    for (int l = 0; l < N; l++) x0 += p[l] * a.S2F[l];
#pragma unroll
                    for (int l = 0; l < N; l++) x1 += p[l] * a.S2F[(M - 1) * N + l];
                    double* to = a.tr_out[d] + ((((size_t)t * NVP + v) * E + e) * 2) * NT + q;
                    to[0] = x0;
                    to[NT] = x1;
                }
            }
            __syncthreads();
        }
        return;
    }

    // ---- corrector (src/transforms.cpp:354-372), optional SSP combine
    double un[NVP];
#pragma unroll
    for (int v = 0; v < NVP; v++) {
        double du = 0;
#pragma unroll
        for (int t = 0; t < T; t++) du += dudt[t][v] * a.w[t] * a.dt;
        un[v] = u_old[v] - du;
    }
    if (a.rk_a != 0.0 && live) {  // combine_solution (src/transforms.cpp:252-261)
#pragma unroll
        for (int v = 0; v < NVP; v++) {
            double u0 = a.U0[(size_t)v * slabE + (size_t)e * TPE + s];
            un[v] = a.rk_a * u0 + (1.0 - a.rk_a) * un[v];
        }
    }
    if (live) {
#pragma unroll
        for (int v = 0; v < NVP; v++) {
            if (a.rk_first) a.U0[(size_t)v * slabE + (size_t)e * TPE + s] = u_old[v];
            a.U_sp_out[(size_t)v * slabE + (size_t)e * TPE + s] = un[v];
        }
    }
    if (!a.want_dt) return;
    // ---- cv primitives: compute_primitives then sp->cv sweeps x, y, z
    double* wa = us;  // ping (state area, free by now)
    double* wb = Fs;  // pong
    if (live) {
        double wv[NVP];
        prim(un, wv, a.gm);
#pragma unroll
        for (int v = 0; v < NVP; v++) wa[v * TPE + s] = wv[v];
    }
    __syncthreads();
    double* src = wa;
    double* dst = wb;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        if (live) {
            const int id = d == 0 ? ii : (d == 1 ? jj : kk);
#pragma unroll
            for (int v = 0; v < NVP; v++) {
                double x = 0;
#pragma unroll
                for (int l = 0; l < N; l++) {
                    int sl = d == 0 ? s - ii + l : (d == 1 ? s + (l - jj) * N : s + (l - kk) * N * N);
                    x += src[v * TPE + sl] * a.S2C[id * N + l];
                }
                dst[v * TPE + s] = x;
            }
        }
        __syncthreads();
        double* tmp = src;
        src = dst;
        dst = tmp;
    }
    // src now holds W_cv of the element
    double r = __builtin_inf();
    if (live) {
        double wc[NVP];
#pragma unroll
        for (int v = 0; v < NVP; v++) wc[v] = src[v * TPE + s];
        if (a.W_cv) {
#pragma unroll
            for (int v = 0; v < NVP; v++) a.W_cv[(size_t)v * slabE + (size_t)e * TPE + s] = wc[v];
            a.W_cv[(size_t)NVP * slabE + (size_t)e * TPE + s] = 0.0;
        }
        bool active = (!a.act[0] || (ei >= 1 && ei <= a.Ne[0] - 2)) && (!a.act[1] || (ej >= 1 && ej <= a.Ne[1] - 2)) &&
                      (!a.act[2] || (ek >= 1 && ek <= a.Ne[2] - 2));
        if (active) {  // src/hydro.cpp:226-256 (inviscid)
            double c_s = sound_speed(wc[0], wc[4], a.gm);
            double c_max = 0, dx_min = 1;
            if (a.act[0]) { c_max += (fabs(wc[1]) + c_s); dx_min = smin(dx_min, a.h[0]); }
            if (a.act[1]) { c_max += (fabs(wc[2]) + c_s); dx_min = smin(dx_min, a.h[1]); }
            if (a.act[2]) { c_max += (fabs(wc[3]) + c_s); dx_min = smin(dx_min, a.h[2]); }
            r = a.cfl * dx_min / c_max / N;
        }
    }
    __syncthreads();
    double* red = lds;
    red[tid] = r;
    __syncthreads();
    for (int wdt = G::BLOCK / 2; wdt > 0; wdt >>= 1) {
        if (tid < wdt) red[tid] = nan_min(red[tid], red[tid + wdt]);
        __syncthreads();
    }
    if (tid == 0) a.part[blockIdx.x] = red[0];
}

// Face traces of slice blockIdx.y of U (in_nv variables per slice): of the
// solution U_sp (one slice), the input of a step's first sweep, or of given
// predictor slices.
template <int N, int DIM>
__global__ void __launch_bounds__(256) k_trace0(const double* __restrict__ U, int in_nv, double* tr0,
                                                               double* tr1, double* tr2, const double* __restrict__ S2F,
                                                               int Eint) {
    using G = Tile<N, DIM>;
    constexpr int M = G::M, NT = G::NT, TPE = G::TPE, EPB = G::EPB;
    __shared__ double us[EPB * NVP * TPE];
    const int tid = threadIdx.x, el = tid / TPE, s = tid % TPE;
    const int e = blockIdx.x * EPB + el;
    const bool live = el < EPB && e < Eint;
    const size_t E = (size_t)Eint, slabE = E * TPE, t = blockIdx.y;
    if (live)
#pragma unroll
        for (int v = 0; v < NVP; v++) us[(el * NVP + v) * TPE + s] = U[(t * in_nv + v) * slabE + (size_t)e * TPE + s];
    __syncthreads();
    for (int task = s; live && task < DIM * NT; task += TPE) {
        const int d = task / NT, q = task % NT;
        double* trd = d == 0 ? tr0 : (d == 1 ? tr1 : tr2);
#pragma unroll
        for (int v = 0; v < NVP; v++) {
            double x0 = 0, x1 = 0;
#pragma unroll
            for (int l = 0; l < N; l++) {
                double p = us[(el * NVP + v) * TPE + spidx<N, DIM>(d, l, q)];
                x0 += p * S2F[l];
            }
#pragma unroll
            for (int l = 0; l < N; l++) {
                double p = us[(el * NVP + v) * TPE + spidx<N, DIM>(d, l, q)];
                x1 += p * S2F[(M - 1) * N + l];
            }
            double* to = trd + (((t * NVP + v) * E + e) * 2) * NT + q;
            to[0] = x0;
            to[NT] = x1;
        }
    }
}

__global__ void k_partial_min(double* part, int n) {
    __shared__ double red[1024];
    double r = __builtin_inf();
    for (int i = threadIdx.x; i < n; i += blockDim.x) r = nan_min(r, part[i]);
    red[threadIdx.x] = r;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = nan_min(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[0] = red[0];
}

// ---------------------------------------------------------------------------
struct Layout {
    size_t E = 0, tpe = 0, nt = 0, T = 0, dim = 0;
    size_t off_uader = 0, off_trA = 0, off_trB = 0, off_u0 = 0, off_part = 0, total = 0;
    size_t tr_axis = 0;  // doubles per axis per trace buffer
    unsigned blocks = 0;
};

template <int N, int DIM>
Layout layout_for(size_t E, int T) {
    using G = Tile<N, DIM>;
    Layout L;
    L.E = E;
    L.tpe = G::TPE;
    L.nt = G::NT;
    L.T = T;
    L.dim = DIM;
    L.blocks = (unsigned)((E + G::EPB - 1) / G::EPB);
    size_t st = (size_t)T * NVP * E * G::TPE;
    L.tr_axis = (size_t)T * NVP * E * 2 * G::NT;
    size_t o = 0;
    L.off_uader = o; o += st;
    L.off_trA = o; o += DIM * L.tr_axis;
    L.off_trB = o; o += DIM * L.tr_axis;
    L.off_u0 = o; o += (size_t)NVP * E * G::TPE;
    L.off_part = o; o += L.blocks + 64;
    L.total = o;
    return L;
}

// The predictor slices (T, NVP, E, TPE) in the stage layout of U_ader (T, nv,
// E, TPE); slots past the physical variables are zero, as the stage
// predictor leaves them.
__global__ void k_export_ader(const double* __restrict__ Ua, double* __restrict__ out, int nv, size_t slabE,
                              size_t total) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const size_t r = i % slabE, tv = i / slabE;
        const int v = (int)(tv % nv), t = (int)(tv / nv);
        out[i] = v < NVP ? Ua[((size_t)t * NVP + v) * slabE + r] : 0.0;
    }
}

// What a call produces instead of a whole step (all null: one step).
struct Extra {
    double* U_export = nullptr;  // the T-1 predictor sweeps from the zero-order guess, slices written here
    int nv_export = 0;
    double* F_out[3] = {nullptr, nullptr, nullptr};  // the predictor sweeps, then the flux evaluation of the
    int out_nv = 0;                                  // final slices written here (stage flux-point layout)
    const double* U_slices = nullptr;  // with F_out: the flux evaluation of these slices (stage layout), no
    int slices_nv = 0;                 // predictor; the workspace then holds only their traces
    double* const* Ffv = nullptr;      // the flux evaluation's sub-cell face fluxes, [t * 3 + d] (FV layout)
    int fvn[3] = {1, 1, 1}, fgh[3] = {0, 0, 0};
};

template <int N, int DIM>
size_t slice_trace_doubles(size_t E, int T) {
    return (size_t)DIM * T * NVP * E * 2 * Tile<N, DIM>::NT;
}

template <int N, int DIM, int T>
int run_step(spdk_ctx* c, double* U_sp, const int Ne[3], void* ws, double dt, double* W_cv, double* F_x,
             double* dt_next, const Extra* x) {
    using G = Tile<N, DIM>;
    size_t E = (size_t)Ne[0] * Ne[1] * Ne[2];
    Layout L = layout_for<N, DIM>(E, T);
    double* W = (double*)ws;
    double* trA[3] = {nullptr, nullptr, nullptr};
    double* trB[3] = {nullptr, nullptr, nullptr};
    for (int d = 0; d < DIM; d++) {
        trA[d] = W + L.off_trA + d * L.tr_axis;
        trB[d] = W + L.off_trB + d * L.tr_axis;
    }
    SweepArgs a;
    std::memset(&a, 0, sizeof(a));
    a.U_sp = U_sp;
    a.U_sp_out = U_sp;
    a.U0 = W + L.off_u0;
    a.part = W + L.off_part;
    a.Dm = c->mat[SPDK_M_DFP_TO_SP];
    a.S2F = c->mat[SPDK_M_SP_TO_FP];
    a.S2C = c->mat[SPDK_M_SP_TO_CV];
    a.inv = c->mat[SPDK_M_INVADER];
    a.w = c->wt;
    for (int d = 0; d < 3; d++) {
        a.Ne[d] = Ne[d];
        a.act[d] = c->prm.active[d];
        a.bc[d] = c->prm.bc[d];
        a.h[d] = c->prm.h[d];
    }
    a.E = (int)E;
    a.in_nv = NVP;
    a.dt = dt;
    a.gm = c->prm.gamma;
    a.cfl = c->prm.cfl;
    for (int d = 0; d < 3; d++) a.g[d] = c->prm.g[d];
    a.grav = (a.g[0] != 0.0 || a.g[1] != 0.0 || a.g[2] != 0.0);
    size_t lds = sizeof(double) * (size_t)G::EPB * G::LDS_ELEM;
    if (lds < sizeof(double) * G::BLOCK) lds = sizeof(double) * G::BLOCK;
    dim3 grid(L.blocks), block(G::BLOCK);
    // timing: one event pair per sweep launch, harvested after the step's own sync
    auto launch_sweep = [&](const SweepArgs& args) -> int {
        bool rec = c->timing && !x && c->ev_used < spdk_ctx::kEvPool;
        if (rec) SPDK_CHECK(hipEventRecord(c->evs[c->ev_used][0], c->stream));
        if (args.Ffv[0][0])
            hipLaunchKernelGGL((k_sweep<N, DIM, T, true>), dim3((L.blocks + 7) / 8 * 8), block, lds, c->stream, args);
        else
            hipLaunchKernelGGL((k_sweep<N, DIM, T, false>), grid, block, lds, c->stream, args);
        SPDK_LAUNCH_CHECK("k_sweep");
        if (rec) SPDK_CHECK(hipEventRecord(c->evs[c->ev_used++][1], c->stream));
        return SPDK_OK;
    };
    auto traces0 = [&]() -> int {
        hipLaunchKernelGGL((k_trace0<N, DIM>), grid, block, 0, c->stream, U_sp, NV, trA[0], trA[1], trA[2], a.S2F, (int)E);
        SPDK_LAUNCH_CHECK("k_trace0");
        return SPDK_OK;
    };
    int rc;
    const bool flux_out = x && (x->F_out[0] || x->Ffv);
    auto set_fv = [&](SweepArgs& sa) {
        if (!x || !x->Ffv) return;
        for (int t = 0; t < T; t++)
            for (int d = 0; d < 3; d++) sa.Ffv[t][d] = x->Ffv[t * 3 + d];
        for (int d = 0; d < 3; d++) { sa.fvn[d] = x->fvn[d]; sa.fgh[d] = x->fgh[d]; }
    };
    if (T > 1 && x && x->U_slices) {  // flux evaluation of given slices: their traces, then one flux sweep
        double* trS[3] = {nullptr, nullptr, nullptr};
        for (int d = 0; d < DIM; d++) trS[d] = W + (size_t)d * L.tr_axis;
        hipLaunchKernelGGL((k_trace0<N, DIM>), dim3(L.blocks, T), block, 0, c->stream, x->U_slices, x->slices_nv, trS[0],
                           trS[1], trS[2], a.S2F, (int)E);
        SPDK_LAUNCH_CHECK("k_trace0");
        SweepArgs s = a;
        s.U0 = nullptr;
        s.part = nullptr;
        for (int d = 0; d < 3; d++) { s.tr_in[d] = trS[d]; s.tr_out[d] = nullptr; s.F_out[d] = x->F_out[d]; }
        s.tin_slices = T;
        s.in_ader = 1;
        s.U_in = x->U_slices;
        s.in_nv = x->slices_nv;
        s.flux_only = 1;
        s.out_nv = x->out_nv;
        set_fv(s);
        return launch_sweep(s);
    }
    if (T > 1) {  // ADER: K = T sweeps, the last one with the corrector (or the flux evaluation)
        if ((rc = traces0())) return rc;
        const bool pred_only = x && x->U_export;
        for (int k = 0; k < (pred_only ? T - 1 : T); k++) {
            SweepArgs s = a;
            bool last = k == T - 1;
            double** tin = (k % 2 == 0) ? trA : trB;
            double** tout = (k % 2 == 0) ? trB : trA;
            for (int d = 0; d < 3; d++) { s.tr_in[d] = tin[d]; s.tr_out[d] = tout[d]; }
            s.tin_slices = k == 0 ? 1 : T;
            s.in_ader = k > 0;
            s.U_in = W + L.off_uader;
            s.U_out = W + L.off_uader;
            s.out_pred = !last;
            s.want_dt = last;
            s.W_cv = last ? W_cv : nullptr;
            s.F_x = last ? F_x : nullptr;
            if (last && flux_out) {
                s.want_dt = 0;
                s.W_cv = nullptr;
                s.F_x = nullptr;
                s.flux_only = 1;
                s.out_nv = x->out_nv;
                for (int d = 0; d < 3; d++) s.F_out[d] = x->F_out[d];
                set_fv(s);
            }
            if ((rc = launch_sweep(s))) return rc;
        }
        if (flux_out) return SPDK_OK;
        if (pred_only) {
            const size_t slabE = E * G::TPE, total = (size_t)T * x->nv_export * slabE;
            hipLaunchKernelGGL(k_export_ader, dim3((unsigned)std::min<size_t>((total + 255) / 256, 65536)), dim3(256), 0,
                               c->stream, W + L.off_uader, x->U_export, x->nv_export, slabE, total);
            SPDK_LAUNCH_CHECK("k_export_ader");
            return SPDK_OK;
        }
    } else {  // SSP-RK stages (src/hydro_ader.hpp:313-325)
        int order = c->prm.rk_order;
        double rk[3] = {0, 0, 0};
        if (order == 2) rk[1] = 0.5;
        if (order == 3) { rk[1] = 0.75; rk[2] = 1.0 / 3.0; }
        for (int st = 0; st < order; st++) {
            if ((rc = traces0())) return rc;
            SweepArgs s = a;
            for (int d = 0; d < 3; d++) { s.tr_in[d] = trA[d]; s.tr_out[d] = nullptr; }
            s.tin_slices = 1;
            s.in_ader = 0;
            s.out_pred = 0;
            s.rk_a = rk[st];
            s.rk_first = st == 0 && order > 1;
            bool last = st == order - 1;
            s.want_dt = last;
            s.W_cv = last ? W_cv : nullptr;
            s.F_x = last ? F_x : nullptr;
            if ((rc = launch_sweep(s))) return rc;
        }
    }
    hipLaunchKernelGGL(k_partial_min, dim3(1), dim3(1024), 0, c->stream, a.part, (int)L.blocks);
    SPDK_LAUNCH_CHECK("k_partial_min");
    SPDK_CHECK(hipMemcpyAsync(c->red_host, a.part, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SPDK_CHECK(hipStreamSynchronize(c->stream));
    if (dt_next) *dt_next = c->red_host[0];
    for (int i = 0; i < c->ev_used; i++) {
        float ms = 0;
        SPDK_CHECK(hipEventElapsedTime(&ms, c->evs[i][0], c->evs[i][1]));
        c->kernel_ms += ms;
        c->kernel_launches++;
    }
    c->ev_used = 0;
    return SPDK_OK;
}

typedef int (*step_fn)(spdk_ctx*, double*, const int*, void*, double, double*, double*, double*, const Extra*);
typedef Layout (*layout_fn)(size_t, int);
typedef size_t (*trace_fn)(size_t, int);

struct Variant {
    int n, dim;
    step_fn ader, rk;
    layout_fn layout;
    trace_fn slice_traces;
};
#define SPDK_VARIANT(Nn, D) {Nn, D, run_step<Nn, D, Nn>, run_step<Nn, D, 1>, layout_for<Nn, D>, slice_trace_doubles<Nn, D>}
const Variant kVariants[] = {SPDK_VARIANT(2, 1), SPDK_VARIANT(3, 1), SPDK_VARIANT(4, 1), SPDK_VARIANT(5, 1),
                             SPDK_VARIANT(2, 2), SPDK_VARIANT(3, 2), SPDK_VARIANT(4, 2), SPDK_VARIANT(5, 2),
                             SPDK_VARIANT(2, 3), SPDK_VARIANT(3, 3), SPDK_VARIANT(4, 3)};

const Variant* find_variant(spdk_ctx* c) {
    int dim = c->prm.active[0] + c->prm.active[1] + c->prm.active[2];
    // the fused tile assumes active axes are x, x-y or x-y-z
    if (!c->prm.active[0] || (c->prm.active[2] && !c->prm.active[1])) return nullptr;
    for (const Variant& v : kVariants)
        if (v.n == c->n && v.dim == dim) return &v;
    return nullptr;
}

// Workspace of the allocation-free step entry: one array per GPU, resident
// until the process ends. A kernel takes its address as seen from the
// device, read per device with hipGetSymbolAddress (the array's name in host
// code is a host-side stand-in, not a device address).
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (looked up once per device).
int scratch_for_current_device(void** out) {
    int dev = -1;
    SPDK_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) { set_error("fused noalloc: device index out of range"); return SPDK_E_ARG; }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("fused noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return SPDK_E_HIP;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return SPDK_OK;
}

}  // namespace

extern "C" {

size_t spdk_fused_workspace_bytes(spdk_ctx* c, const int N[3]) {
    const Variant* v = c ? find_variant(c) : nullptr;
    if (!v) {
        set_error("fused path: no tile variant for this degree/dimensionality");
        return 0;
    }
    if (c->T > kMaxT || c->prm.nu != 0.0) {
        set_error("fused path covers the inviscid update");
        return 0;
    }
    Layout L = v->layout((size_t)N[0] * N[1] * N[2], c->T);
    return L.total * sizeof(double);
}

// Checks of a whole fused step on U_sp of shape s.
static int check_step_fused(spdk_ctx* c, const spdk_sd_shape& s, const Variant** vout) {
    if (c->prm.nu != 0.0) {
        set_error("fused path covers the inviscid update");
        return SPDK_E_CONFIG;
    }
    const Variant* v = find_variant(c);
    if (!v) { set_error("fused path: no tile variant for this degree/dimensionality"); return SPDK_E_CONFIG; }
    if (s.nv != NV || s.nt != 1 || s.nx != c->n || s.ny != (c->prm.active[1] ? c->n : 1) ||
        s.nz != (c->prm.active[2] ? c->n : 1)) {
        set_error("fused: U_sp shape");
        return SPDK_E_SHAPE;
    }
    for (int d = 0; d < 3; d++) {
        int Nd = d == 0 ? s.Nx : (d == 1 ? s.Ny : s.Nz);
        if (c->prm.active[d] && Nd < 3) { set_error("fused: axis needs ghost elements"); return SPDK_E_SHAPE; }
    }
    if ((uint64_t)s.Nx * s.Ny * s.Nz >= (1ull << 31)) { set_error("fused: too many elements"); return SPDK_E_SHAPE; }
    *vout = v;
    return SPDK_OK;
}

int spdk_hydro_step_fused(spdk_ctx* c, double* U_sp, spdk_sd_shape s, void* ws, double dt, double* W_cv, double* F_x,
                          double* dt_next) {
    if (!c || !U_sp || !ws) { set_error("fused: null argument"); return SPDK_E_ARG; }
    const Variant* v = nullptr;
    if (int rc = check_step_fused(c, s, &v)) return rc;
    int Ne[3] = {s.Nx, s.Ny, s.Nz};
    std::strncpy(c->kernel_name, "k_sweep", sizeof(c->kernel_name) - 1);
    if (c->T > 1) return v->ader(c, U_sp, Ne, ws, dt, W_cv, F_x, dt_next, nullptr);
    return v->rk(c, U_sp, Ne, ws, dt, W_cv, F_x, dt_next, nullptr);
}

size_t spdk_fused_noalloc_max_bytes(void) { return kScratchBytes; }

int spdk_hydro_step_fused_noalloc(spdk_ctx* c, double* U_sp, spdk_sd_shape s, double dt, double* W_cv, double* F_x,
                                  double* dt_next) {
    if (!c || !U_sp) { set_error("fused noalloc: null argument"); return SPDK_E_ARG; }
    const Variant* v = nullptr;
    if (int rc = check_step_fused(c, s, &v)) return rc;
    const size_t need = v->layout((size_t)s.Nx * s.Ny * s.Nz, c->T).total * sizeof(double);
    if (need > kScratchBytes) {
        set_error("fused noalloc: the step needs " + std::to_string(need) + " bytes of workspace, more than the " +
                  std::to_string(kScratchBytes) + " of the static scratch; use spdk_hydro_step_fused");
        return SPDK_E_SHAPE;
    }
    void* ws = nullptr;
    if (int rc = scratch_for_current_device(&ws)) return rc;
    int Ne[3] = {s.Nx, s.Ny, s.Nz};
    std::strncpy(c->kernel_name, "k_sweep", sizeof(c->kernel_name) - 1);
    if (c->T > 1) return v->ader(c, U_sp, Ne, ws, dt, W_cv, F_x, dt_next, nullptr);
    return v->rk(c, U_sp, Ne, ws, dt, W_cv, F_x, dt_next, nullptr);
}

int spdk_fused_predict(spdk_ctx* c, double* U_sp, spdk_sd_shape s, void* ws, double dt, double* U_ader,
                       spdk_sd_shape s_ader) {
    if (!c || !U_sp || !ws || !U_ader) { set_error("fused predict: null argument"); return SPDK_E_ARG; }
    if (c->prm.nu != 0.0 || c->T < 2) {
        set_error("fused predict covers the inviscid ADER predictor");
        return SPDK_E_CONFIG;
    }
    const Variant* v = find_variant(c);
    if (!v) { set_error("fused path: no tile variant for this degree/dimensionality"); return SPDK_E_CONFIG; }
    if (s.nv != NV || s.nt != 1 || s.nx != c->n || s.ny != (c->prm.active[1] ? c->n : 1) ||
        s.nz != (c->prm.active[2] ? c->n : 1) || s_ader.nt != c->T || s_ader.nv < NVP || s_ader.Nx != s.Nx ||
        s_ader.Ny != s.Ny || s_ader.Nz != s.Nz || s_ader.nx != s.nx || s_ader.ny != s.ny || s_ader.nz != s.nz) {
        set_error("fused predict: U_sp / U_ader shape");
        return SPDK_E_SHAPE;
    }
    for (int d = 0; d < 3; d++) {
        int Nd = d == 0 ? s.Nx : (d == 1 ? s.Ny : s.Nz);
        if (c->prm.active[d] && Nd < 3) { set_error("fused: axis needs ghost elements"); return SPDK_E_SHAPE; }
    }
    if ((uint64_t)s.Nx * s.Ny * s.Nz >= (1ull << 31)) { set_error("fused: too many elements"); return SPDK_E_SHAPE; }
    int Ne[3] = {s.Nx, s.Ny, s.Nz};
    std::strncpy(c->kernel_name, "k_sweep", sizeof(c->kernel_name) - 1);
    Extra x;
    x.U_export = U_ader;
    x.nv_export = s_ader.nv;
    return v->ader(c, U_sp, Ne, ws, dt, nullptr, nullptr, nullptr, &x);
}

// Checks shared by the flux-evaluation entries: the ADER slices shape and the
// flux-point arrays of the active axes.
static int check_flux_args(spdk_ctx* c, spdk_sd_shape s_ader, double* const F[3], const Variant** vout,
                           bool need_F = true) {
    if (c->prm.nu != 0.0 || c->T < 2) {
        set_error("fused flux evaluation covers the inviscid ADER step");
        return SPDK_E_CONFIG;
    }
    const Variant* v = find_variant(c);
    if (!v) { set_error("fused path: no tile variant for this degree/dimensionality"); return SPDK_E_CONFIG; }
    if (s_ader.nt != c->T || s_ader.nv < NVP || s_ader.nx != c->n || s_ader.ny != (c->prm.active[1] ? c->n : 1) ||
        s_ader.nz != (c->prm.active[2] ? c->n : 1)) {
        set_error("fused fluxes: slice shape");
        return SPDK_E_SHAPE;
    }
    for (int d = 0; d < 3; d++) {
        int Nd = d == 0 ? s_ader.Nx : (d == 1 ? s_ader.Ny : s_ader.Nz);
        if (c->prm.active[d] && Nd < 3) { set_error("fused: axis needs ghost elements"); return SPDK_E_SHAPE; }
        if (need_F && c->prm.active[d] && !F[d]) {
            set_error("fused fluxes: null flux array of an active axis");
            return SPDK_E_ARG;
        }
    }
    if ((uint64_t)s_ader.Nx * s_ader.Ny * s_ader.Nz >= (1ull << 31)) { set_error("fused: too many elements"); return SPDK_E_SHAPE; }
    *vout = v;
    return SPDK_OK;
}

int spdk_fused_predict_fluxes(spdk_ctx* c, double* U_sp, spdk_sd_shape s, void* ws, double dt, double* F_x,
                              double* F_y, double* F_z, spdk_sd_shape s_ader) {
    if (!c || !U_sp || !ws) { set_error("fused predict fluxes: null argument"); return SPDK_E_ARG; }
    double* F[3] = {F_x, F_y, F_z};
    const Variant* v = nullptr;
    int rc = check_flux_args(c, s_ader, F, &v);
    if (rc) return rc;
    if (s.nv != NV || s.nt != 1 || s.nx != s_ader.nx || s.ny != s_ader.ny || s.nz != s_ader.nz || s.Nx != s_ader.Nx ||
        s.Ny != s_ader.Ny || s.Nz != s_ader.Nz) {
        set_error("fused predict fluxes: U_sp shape");
        return SPDK_E_SHAPE;
    }
    int Ne[3] = {s.Nx, s.Ny, s.Nz};
    Extra x;
    for (int d = 0; d < 3; d++) x.F_out[d] = F[d];
    x.out_nv = s_ader.nv;
    std::strncpy(c->kernel_name, "k_sweep", sizeof(c->kernel_name) - 1);
    return v->ader(c, U_sp, Ne, ws, dt, nullptr, nullptr, nullptr, &x);
}

int spdk_fused_predict_fv_fluxes(spdk_ctx* c, double* U_sp, spdk_sd_shape s, void* ws, double dt, double* const* F_fv,
                                 spdk_fv_shape sc, double* F_x, spdk_sd_shape s_ader) {
    if (!c || !U_sp || !ws || !F_fv) { set_error("fused predict fv fluxes: null argument"); return SPDK_E_ARG; }
    double* F[3] = {F_x, nullptr, nullptr};  // flux-point arrays are optional here
    const Variant* v = nullptr;
    int rc = check_flux_args(c, s_ader, F, &v, false);
    if (rc) return rc;
    if (s.nv != NV || s.nt != 1 || s.nx != s_ader.nx || s.ny != s_ader.ny || s.nz != s_ader.nz || s.Nx != s_ader.Nx ||
        s.Ny != s_ader.Ny || s.Nz != s_ader.Nz) {
        set_error("fused predict fv fluxes: U_sp shape");
        return SPDK_E_SHAPE;
    }
    const int Ne[3] = {s.Nx, s.Ny, s.Nz}, nc[3] = {sc.Nx, sc.Ny, sc.Nz};
    for (int d = 0; d < 3; d++) {
        const bool act = c->prm.active[d] != 0;
        const bool ok = act ? (Ne[d] > 2 * c->ngh[d] && nc[d] == (Ne[d] - 2 * c->ngh[d]) * c->n + 2 * c->fgh[d] &&
                               c->ngh[d] == 1)
                            : (Ne[d] == 1 && nc[d] == 1);
        if (!ok || sc.nv != NV) {
            set_error("fused predict fv fluxes: FV cell shape disagrees with U_sp along axis " + std::to_string(d));
            return SPDK_E_SHAPE;
        }
        for (int t = 0; t < c->T; t++)
            if (act && !F_fv[t * 3 + d]) { set_error("fused predict fv fluxes: null face array"); return SPDK_E_ARG; }
    }
    Extra x;
    x.F_out[0] = F_x;  // optional: the x flux-point fluxes (outputs)
    x.out_nv = s_ader.nv;
    x.Ffv = F_fv;
    for (int d = 0; d < 3; d++) { x.fvn[d] = nc[d]; x.fgh[d] = c->fgh[d]; }
    std::strncpy(c->kernel_name, "k_sweep", sizeof(c->kernel_name) - 1);
    return v->ader(c, U_sp, Ne, ws, dt, nullptr, nullptr, nullptr, &x);
}

size_t spdk_fused_fluxes_workspace_bytes(spdk_ctx* c, const int N[3]) {
    const Variant* v = c ? find_variant(c) : nullptr;
    if (!v || c->T < 2) {
        set_error("fused fluxes: no tile variant for this degree/dimensionality, or not ADER");
        return 0;
    }
    return v->slice_traces((size_t)N[0] * N[1] * N[2], c->T) * sizeof(double);
}

int spdk_fused_fluxes(spdk_ctx* c, const double* U_ader, spdk_sd_shape s_ader, void* ws, double* F_x, double* F_y,
                      double* F_z) {
    if (!c || !U_ader || !ws) { set_error("fused fluxes: null argument"); return SPDK_E_ARG; }
    double* F[3] = {F_x, F_y, F_z};
    const Variant* v = nullptr;
    int rc = check_flux_args(c, s_ader, F, &v);
    if (rc) return rc;
    int Ne[3] = {s_ader.Nx, s_ader.Ny, s_ader.Nz};
    Extra x;
    for (int d = 0; d < 3; d++) x.F_out[d] = F[d];
    x.out_nv = s_ader.nv;
    x.U_slices = U_ader;
    x.slices_nv = s_ader.nv;
    std::strncpy(c->kernel_name, "k_sweep", sizeof(c->kernel_name) - 1);
    return v->ader(c, nullptr, Ne, ws, 0.0, nullptr, nullptr, nullptr, &x);
}

}  // extern "C"
