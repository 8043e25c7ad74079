// SD stage kernels: one thread per output (element, point), looping over
// time slices and variables inside, in the reference's C-order layout.
// Each kernel restates the arithmetic of the reference function cited at
// its launcher; compiled with -ffp-contract=off they are bit-identical to it.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstring>
#include <mutex>
#include "ctx.hpp"
#include "physics.hpp"

using namespace spdk;

namespace spdk {
static thread_local std::string g_err;
void set_error(const std::string& m) { g_err = m; }
int hip_fail(hipError_t e, const char* what) {
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return SPDK_E_HIP;
}
}  // namespace spdk

static constexpr int BS = 256;

// ---------------------------------------------------------------------------
// transform_a_to_b_1d / _slice: src/transforms.cpp:120-177. One thread per
// pencil (time slice, variable, element, line of points along `dim`): it
// reads the line's q inputs once and writes all of its outputs, each the
// sum over ll of A[ll] M(id, ll) in ascending ll, the reference's order.
// Arrays are rows of points, row = (t, v, element); consecutive threads take
// consecutive lines of one element, so loads and stores stay within the
// element's contiguous block (coalesced along y and z, 32-byte strided along
// x with every line of the block read by the same wave).
template <int Q>
__global__ void __launch_bounds__(BS) k_transform_1d(const double* __restrict__ A, SD sa, double* __restrict__ B, SD sb,
                                                     const double* __restrict__ M, int ldm, int dim, int t_src,
                                                     size_t total) {
    size_t idx = (size_t)blockIdx.x * BS + threadIdx.x;
    if (idx >= total) return;
    const int nxa = sa.nx, nya = sa.ny, nxb = sb.nx, nyb = sb.ny;
    const int nout = dim == 0 ? sb.nx : (dim == 1 ? sb.ny : sb.nz);
    // transverse extent: lines per element
    const int c0 = dim == 0 ? sb.ny : sb.nx, c1 = dim == 2 ? sb.ny : sb.nz;
    const int P = c0 * c1;
    const size_t rb = idx / P;  // output row: (t, v, element) of B
    const int line = (int)(idx % P);
    const int a0 = line % c0, a1 = line / c0;
    size_t ra = rb;
    if (t_src >= 0) {  // B holds one slice; A's row is that of slice t_src
        const size_t E = (size_t)sb.Nz * sb.Ny * sb.Nx;
        ra = ((size_t)t_src * sa.nv) * E + rb;
    }
    const double* Ar = A + ra * sa.pts();
    double* Br = B + rb * sb.pts();
    // offsets of point ll along the line in A and of output f in B
    int abase, astep, bbase, bstep;
    if (dim == 0) { abase = (a1 * nya + a0) * nxa; astep = 1; bbase = (a1 * nyb + a0) * nxb; bstep = 1; }
    else if (dim == 1) { abase = a1 * nya * nxa + a0; astep = nxa; bbase = a1 * nyb * nxb + a0; bstep = nxb; }
    else { abase = a1 * nxa + a0; astep = nxa * nya; bbase = a1 * nxb + a0; bstep = nxb * nyb; }
    double in[Q];
#pragma unroll
    for (int ll = 0; ll < Q; ll++) in[ll] = Ar[abase + ll * astep];
    for (int f = 0; f < nout; f++) {
        const double* Mrow = M + (size_t)f * ldm;
        double u = 0;
#pragma unroll
        for (int ll = 0; ll < Q; ll++) u += in[ll] * Mrow[ll];
        Br[bbase + f * bstep] = u;
    }
}

// ---------------------------------------------------------------------------
// transform_a_to_b (src/transforms.cpp:99-117) with the sweeps of every active
// axis (x, x-y or x-y-z) of an element in one pass: a workgroup owns EPB whole
// elements; per (slice, variable) plane their values go through LDS and the
// x, y, z sweeps run there, each output the sum over ll of in[ll] M(id, ll)
// in ascending ll as in k_transform_1d. The result equals the ping-pong
// sequence bit for bit, with one read of A and one write of B.
template <int Q, int D>
__global__ void __launch_bounds__(BS) k_transform_el(const double* __restrict__ A, double* __restrict__ B,
                                                     const double* __restrict__ M, int ldm, uint32_t E, int planes) {
    constexpr int TPE = D == 1 ? Q : (D == 2 ? Q * Q : Q * Q * Q);
    constexpr int EPB = TPE >= BS ? 1 : BS / TPE;
    static_assert(TPE <= BS, "one thread per point of an element");
    __shared__ double buf[2][EPB * TPE];
    __shared__ double sM[Q * Q];
    const int tid = threadIdx.x, el = tid / TPE, s = tid % TPE;
    const uint32_t e = blockIdx.x * EPB + el;
    const bool live = el < EPB && e < E;
    if (tid < Q * Q) sM[tid] = M[(tid / Q) * ldm + tid % Q];
    const int ii = s % Q, jj = (s / Q) % Q, kk = s / (Q * Q);
    double* b0 = buf[0] + (live ? el : 0) * TPE;
    double* b1 = buf[1] + (live ? el : 0) * TPE;
    const size_t slab = (size_t)E * TPE;
    __syncthreads();
    for (int pl = 0; pl < planes; pl++) {
        const size_t o = (size_t)pl * slab + (size_t)e * TPE + s;
        if (live) b0[s] = A[o];
        __syncthreads();
        double u = 0;
        if (live) {
#pragma unroll
            for (int l = 0; l < Q; l++) u += b0[s - ii + l] * sM[ii * Q + l];
        }
        if (D >= 2) {
            if (live) b1[s] = u;
            __syncthreads();
            u = 0;
            if (live) {
#pragma unroll
                for (int l = 0; l < Q; l++) u += b1[s + (l - jj) * Q] * sM[jj * Q + l];
            }
            if (D == 3) {
                if (live) b0[s] = u;
                __syncthreads();
                u = 0;
                if (live) {
#pragma unroll
                    for (int l = 0; l < Q; l++) u += b0[s + (l - kk) * Q * Q] * sM[kk * Q + l];
                }
            }
        }
        if (live) B[o] = u;
        __syncthreads();
    }
}

template <int D>
static bool launch_transform_el(spdk_ctx* c, const double* A, double* B, const spdk_sd_shape& s, int mat) {
    const uint32_t E = (uint32_t)((size_t)s.Nz * s.Ny * s.Nx);
    const int planes = s.nt * s.nv;
    const int q = s.nx;
#define SPDK_TEL(QQ)                                                                                                  \
    case QQ: {                                                                                                        \
        constexpr int TPE = D == 1 ? QQ : (D == 2 ? QQ * QQ : QQ * QQ * QQ);                                          \
        constexpr int EPB = TPE >= BS ? 1 : BS / TPE;                                                                 \
        hipLaunchKernelGGL((k_transform_el<QQ, D>), dim3((E + EPB - 1) / EPB), dim3(BS), 0, c->stream, A, B, c->mat[mat], \
                           c->mcols[mat], E, planes);                                                                 \
        return true;                                                                                                  \
    }
    switch (q) {
        SPDK_TEL(1) SPDK_TEL(2) SPDK_TEL(3) SPDK_TEL(4) SPDK_TEL(5) SPDK_TEL(6)
        default: return false;
    }
#undef SPDK_TEL
}

// ---------------------------------------------------------------------------
// compute_conservatives / compute_primitives: src/hydro.cpp:47-72, :108-132
__global__ void __launch_bounds__(BS) k_cons_prim(const double* __restrict__ A, double* __restrict__ B, SD s, double gm,
                                                  int to_prim, uint32_t total) {
    uint32_t idx = blockIdx.x * BS + threadIdx.x;
    if (idx >= total) return;
    size_t slab = s.slab();
    for (int t = 0; t < s.nt; t++) {
        size_t o = (size_t)t * s.nv * slab + idx;
        double a[NVP], b[NVP];
#pragma unroll
        for (int v = 0; v < NVP; v++) a[v] = A[o + v * slab];
        if (to_prim) prim(a, b, gm);
        else cons(a, b, gm);
#pragma unroll
        for (int v = 0; v < NVP; v++) B[o + v * slab] = b[v];
        if (s.nv > NVP) B[o + NVP * slab] = 0.0;  // slot: zero-initialised local in the reference
    }
}

// ---------------------------------------------------------------------------
// compute_fluxes_t<V1,V2,V3>: src/hydro.cpp:167-206
template <int D>
__global__ void __launch_bounds__(BS) k_flux(const double* __restrict__ U, double* __restrict__ F, SD s, double gm,
                                             uint32_t total) {
    uint32_t idx = blockIdx.x * BS + threadIdx.x;
    if (idx >= total) return;
    size_t slab = s.slab();
    for (int t = 0; t < s.nt; t++) {
        size_t o = (size_t)t * s.nv * slab + idx;
        double u[NVP], f[NVP];
#pragma unroll
        for (int v = 0; v < NVP; v++) u[v] = U[o + v * slab];
        state_flux<D>(u, f, gm);
#pragma unroll
        for (int v = 0; v < NVP; v++) F[o + v * slab] = f[v];
        if (s.nv > NVP) F[o + NVP * slab] = 0.0;
    }
}

// ---------------------------------------------------------------------------
// boundaries (SD interface traces): src/boundary.cpp:48-113. One thread per
// transverse (element, point) location of the ghost slab.
__global__ void __launch_bounds__(BS) k_sd_bc(double* __restrict__ U, SD s, int dim, int type, uint32_t total) {
    uint32_t idx = blockIdx.x * BS + threadIdx.x;
    if (idx >= total) return;
    int Mx = dim == 0 ? 1 : s.Nx, My = dim == 1 ? 1 : s.Ny, Mz = dim == 2 ? 1 : s.Nz;
    int px = dim == 0 ? 1 : s.nx, py = dim == 1 ? 1 : s.ny, pz = dim == 2 ? 1 : s.nz;
    int k, j, i, kk, jj, ii;
    split6(idx, Mz, My, Mx, pz, py, px, k, j, i, kk, jj, ii);
    int N = dim == 0 ? s.Nx : (dim == 1 ? s.Ny : s.Nz);
    int n = dim == 0 ? s.nx : (dim == 1 ? s.ny : s.nz);
    auto at = [&](int t, int v, int l, int ll) -> size_t {
        if (dim == 0) return s.off(t, v, k, j, l, kk, jj, ll);
        if (dim == 1) return s.off(t, v, k, l, i, kk, ll, ii);
        return s.off(t, v, l, j, i, ll, jj, ii);
    };
    for (int t = 0; t < s.nt; t++)
        for (int v = 0; v < s.nv; v++) {
            double BL, BR;
            if (type == SPDK_BC_PERIODIC) {
                BL = U[at(t, v, N - 2, n - 1)];
                BR = U[at(t, v, 1, 0)];
            } else if (type == SPDK_BC_GRADFREE) {
                BL = U[at(t, v, 1, 0)];
                BR = U[at(t, v, N - 2, n - 1)];
            } else {
                double sgn = (v == 1 + dim) ? -1.0 : 1.0;
                BL = sgn * U[at(t, v, 1, 0)];
                BR = sgn * U[at(t, v, N - 2, n - 1)];
            }
            U[at(t, v, 0, n - 1)] = BL;
            U[at(t, v, N - 1, 0)] = BR;
        }
}

// ---------------------------------------------------------------------------
// sd_riemann_solver_t<D,...>: src/hydro.cpp:368-418
template <int D>
__global__ void __launch_bounds__(BS) k_riemann(double* __restrict__ U, double* __restrict__ F, SD s, double gm,
                                                int viscous, uint32_t total) {
    uint32_t idx = blockIdx.x * BS + threadIdx.x;
    if (idx >= total) return;
    int Nx = s.Nx - (D == 0), Ny = s.Ny - (D == 1), Nz = s.Nz - (D == 2);
    int px = D == 0 ? 1 : s.nx, py = D == 1 ? 1 : s.ny, pz = D == 2 ? 1 : s.nz;
    int k, j, i, kk, jj, ii;
    split6(idx, Nz, Ny, Nx, pz, py, px, k, j, i, kk, jj, ii);
    int n = D == 0 ? s.nx : (D == 1 ? s.ny : s.nz);
    size_t oL, oR;
    if (D == 0) { oL = s.off(0, 0, k, j, i, kk, jj, n - 1); oR = s.off(0, 0, k, j, i + 1, kk, jj, 0); }
    else if (D == 1) { oL = s.off(0, 0, k, j, i, kk, n - 1, ii); oR = s.off(0, 0, k, j + 1, i, kk, 0, ii); }
    else { oL = s.off(0, 0, k, j, i, n - 1, jj, ii); oR = s.off(0, 0, k + 1, j, i, 0, jj, ii); }
    size_t slab = s.slab();
    for (int t = 0; t < s.nt; t++) {
        size_t tb = (size_t)t * s.nv * slab;
        double uL[NVP], uR[NVP], f[NVP];
#pragma unroll
        for (int v = 0; v < NVP; v++) { uL[v] = U[tb + v * slab + oL]; uR[v] = U[tb + v * slab + oR]; }
        riemann_llf<D>(uL, uR, f, gm);
#pragma unroll
        for (int v = 0; v < NVP; v++) { F[tb + v * slab + oL] = f[v]; F[tb + v * slab + oR] = f[v]; }
        if (s.nv > NVP) {
            // slot: 0.5*(0+0) - 0.5*c*(uR5-uL5) with the wave speed of this face
            double wL[NVP], wR[NVP];
            prim(uL, wL, gm);
            prim(uR, wR, gm);
            double c = smax(sound_speed(wL[0], wL[4], gm) + fabs(wL[1 + D]), sound_speed(wR[0], wR[4], gm) + fabs(wR[1 + D]));
            double s5 = 0.5 * (0.0 + 0.0) - 0.5 * c * (U[tb + NVP * slab + oR] - U[tb + NVP * slab + oL]);
            F[tb + NVP * slab + oL] = s5;
            F[tb + NVP * slab + oR] = s5;
        }
        if (viscous) {  // riemann_wind(u_L, u_R, 0.5) into both face states (src/hydro.cpp:402-409)
            for (int v = 0; v < s.nv; v++) {
                const double a = U[tb + v * slab + oL], b = U[tb + v * slab + oR];
                const double w = 0.5 * (b + a) + 0.5 * (b - a);
                U[tb + v * slab + oL] = w;
                U[tb + v * slab + oR] = w;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Viscous terms (src/hydro.cpp:420-555), active when nu > 0.
// sd_rusanov_solver: riemann_wind(-0.5) of (right element's face value, left
// element's) added to F on both sides of every element face along D.
template <int D>
__global__ void __launch_bounds__(BS) k_rusanov(const double* __restrict__ U, double* __restrict__ F, SD s,
                                                uint32_t total) {
    uint32_t idx = blockIdx.x * BS + threadIdx.x;
    if (idx >= total) return;
    int Nx = s.Nx - (D == 0), Ny = s.Ny - (D == 1), Nz = s.Nz - (D == 2);
    int px = D == 0 ? 1 : s.nx, py = D == 1 ? 1 : s.ny, pz = D == 2 ? 1 : s.nz;
    int k, j, i, kk, jj, ii;
    split6(idx, Nz, Ny, Nx, pz, py, px, k, j, i, kk, jj, ii);
    int n = D == 0 ? s.nx : (D == 1 ? s.ny : s.nz);
    size_t oL, oR;
    if (D == 0) { oL = s.off(0, 0, k, j, i, kk, jj, n - 1); oR = s.off(0, 0, k, j, i + 1, kk, jj, 0); }
    else if (D == 1) { oL = s.off(0, 0, k, j, i, kk, n - 1, ii); oR = s.off(0, 0, k, j + 1, i, kk, 0, ii); }
    else { oL = s.off(0, 0, k, j, i, n - 1, jj, ii); oR = s.off(0, 0, k + 1, j, i, 0, jj, ii); }
    const size_t slab = s.slab();
    for (int t = 0; t < s.nt; t++)
        for (int v = 0; v < s.nv; v++) {
            const size_t o = ((size_t)t * s.nv + v) * slab;
            const double uL = U[o + oR], uR = U[o + oL];
            const double f = 0.5 * (uR + uL) + (-0.5) * (uR - uL);
            F[o + oL] += f;
            F[o + oR] += f;
        }
}
// compute_gradient: du = sum_ll U(ll) D(id, ll) / h over the flux points along
// D, at the solution points of the active elements.
template <int D>
__global__ void __launch_bounds__(BS) k_gradient(const double* __restrict__ U, SD su, double* __restrict__ dU, SD sd,
                                                 const double* __restrict__ Dm, int ldm, double h, int gx, int gy,
                                                 int gz, uint32_t total) {
    uint32_t idx = blockIdx.x * BS + threadIdx.x;
    if (idx >= total) return;
    int k, j, i, kk, jj, ii;
    split6(idx, sd.Nz - 2 * gz, sd.Ny - 2 * gy, sd.Nx - 2 * gx, sd.nz, sd.ny, sd.nx, k, j, i, kk, jj, ii);
    k += gz; j += gy; i += gx;
    const int q = D == 0 ? sd.nx : (D == 1 ? sd.ny : sd.nz);
    const int id = D == 0 ? ii : (D == 1 ? jj : kk);
    for (int t = 0; t < sd.nt; t++)
        for (int v = 0; v < sd.nv; v++) {
            double du = 0;
            for (int ll = 0; ll < q + 1; ll++) {
                const double u = D == 0 ? U[su.off(t, v, k, j, i, kk, jj, ll)]
                                        : (D == 1 ? U[su.off(t, v, k, j, i, kk, ll, ii)] : U[su.off(t, v, k, j, i, ll, jj, ii)]);
                du += u * Dm[id * ldm + ll] / h;
            }
            dU[sd.off(t, v, k, j, i, kk, jj, ii)] = du;
        }
}
// compute_viscous_flux with viscous_fluxes, at every flux point along D.
template <int D>
__global__ void __launch_bounds__(BS) k_viscous_flux(double* __restrict__ U, SD su, const double* __restrict__ dU1,
                                                     const double* __restrict__ dU2, const double* __restrict__ dU3,
                                                     SD sd, const double* __restrict__ S, int lds, double gm, double nu,
                                                     double beta, uint32_t total) {
    uint32_t idx = blockIdx.x * BS + threadIdx.x;
    if (idx >= total) return;
    int k, j, i, kk, jj, ii;
    split6(idx, su.Nz, su.Ny, su.Nx, su.nz, su.ny, su.nx, k, j, i, kk, jj, ii);
    constexpr int v1 = 1 + D, v2 = 1 + (D + 1) % 3, v3 = 1 + (D + 2) % 3;
    const int q = D == 0 ? su.nx : (D == 1 ? su.ny : su.nz);
    const int id = D == 0 ? ii : (D == 1 ? jj : kk);
    for (int t = 0; t < su.nt; t++) {
        double u[NV] = {0, 0, 0, 0, 0, 0}, du1[NV], du2[NV], du3[NV];
        for (int v = 0; v < su.nv && v < NV; v++) {
            du1[v] = 0;
            du2[v] = 0;
            du3[v] = 0;
            u[v] = U[su.off(t, v, k, j, i, kk, jj, ii)];
            for (int ll = 0; ll < q - 1; ll++) {
                const int a = D == 0 ? ll : ii, b = D == 1 ? ll : jj, c = D == 2 ? ll : kk;
                const size_t o = sd.off(t, v, k, j, i, c, b, a);
                const double sm = S[id * lds + ll];
                du1[v] += dU1[o] * sm;
                du2[v] += dU2[o] * sm;
                du3[v] += dU3[o] * sm;
            }
        }
        double w[NVP];
        prim(u, w, gm);
        double f[NV] = {0, 0, 0, 0, 0, 0};
        f[0] = 0;
        f[4] = 0;
        f[v1] = 2 * nu * du1[v1] + beta * du1[v1];
        f[v1] += beta * du2[v2];
        f[v2] = nu * (du1[v2] + du2[v1]);
        f[4] += w[v2] * f[v2];
        f[v1] += beta * du3[v3];
        f[v3] = nu * (du1[v3] + du3[v1]);
        f[4] += w[v3] * f[v3];
        f[4] += w[v1] * f[v1];
        for (int v = 0; v < su.nv && v < NV; v++) U[su.off(t, v, k, j, i, kk, jj, ii)] = f[v];
    }
}

// ---------------------------------------------------------------------------
// copy_ader: src/hydro_ader.hpp:428-444
__global__ void __launch_bounds__(BS) k_copy_ader(const double* __restrict__ U, double* __restrict__ Ua, SD s,
                                                  uint32_t total) {
    uint32_t idx = blockIdx.x * BS + threadIdx.x;
    if (idx >= total) return;
    size_t slab = s.slab();
    for (int t = 0; t < s.nt; t++)
        for (int v = 0; v < s.nv; v++) Ua[((size_t)t * s.nv + v) * slab + idx] = U[(size_t)v * slab + idx];
}

// gravity_source: src/transforms.cpp:17-27 (U is the ADER state, slice t)
__device__ __forceinline__ double grav_src(const double* U, size_t tb, size_t slab, size_t o, int var, const double* g) {
    if (var == 1) return U[tb + o] * g[0];
    if (var == 2) return U[tb + o] * g[1];
    if (var == 3) return U[tb + o] * g[2];
    if (var == 4) return U[tb + slab + o] * g[0] + U[tb + 2 * slab + o] * g[1] + U[tb + 3 * slab + o] * g[2];
    return 0.0;
}

struct PredArgs {
    double g[3];
    double h[3];
    int act[3];
    int grav;
    double dt;
};

// dudt of one (point, var, t): the flux divergence with x, y, z interleaved
// per flux point (src/transforms.cpp:301-307)
__device__ __forceinline__ double divergence(const double* Fx, const double* Fy, const double* Fz, SD fx, SD fy, SD fz,
                                             const double* Dm, int ldd, int q, const PredArgs& a, int t, int v, int k,
                                             int j, int i, int kk, int jj, int ii) {
    double d = 0;
    for (int ll = 0; ll < q; ll++) {
        if (a.act[0]) d += Fx[fx.off(t, v, k, j, i, kk, jj, ll)] * Dm[ii * ldd + ll] / a.h[0];
        if (a.act[1]) d += Fy[fy.off(t, v, k, j, i, kk, ll, ii)] * Dm[jj * ldd + ll] / a.h[1];
        if (a.act[2]) d += Fz[fz.off(t, v, k, j, i, ll, jj, ii)] * Dm[kk * ldd + ll] / a.h[2];
    }
    return d;
}

// update_prediction: src/transforms.cpp:263-325
__global__ void __launch_bounds__(BS) k_predict(const double* __restrict__ U, double* __restrict__ Ua,
                                                const double* __restrict__ Fx, const double* __restrict__ Fy,
                                                const double* __restrict__ Fz, SD sa, SD fx, SD fy, SD fz,
                                                const double* __restrict__ Dm, int ldd, const double* __restrict__ inv,
                                                const double* __restrict__ w, int q, PredArgs a, uint32_t total) {
    uint32_t idx = blockIdx.x * BS + threadIdx.x;
    if (idx >= total) return;
    int k, j, i, kk, jj, ii;
    split6(idx, sa.Nz, sa.Ny, sa.Nx, sa.nz, sa.ny, sa.nx, k, j, i, kk, jj, ii);
    int T = sa.nt;
    size_t slab = sa.slab();
    for (int v = 0; v < sa.nv; v++) {
        double dudt[10];
        double u_old = U[(size_t)v * slab + idx];
        for (int t = 0; t < T; t++) {
            dudt[t] = divergence(Fx, Fy, Fz, fx, fy, fz, Dm, ldd, q, a, t, v, k, j, i, kk, jj, ii);
            if (a.grav) dudt[t] -= grav_src(Ua, (size_t)t * sa.nv * slab, slab, idx, v, a.g);
        }
        for (int t = 0; t < T; t++) {
            double du = 0;
            for (int ll = 0; ll < T; ll++) du += dudt[ll] * inv[t * T + ll] * w[ll] * a.dt;
            Ua[((size_t)t * sa.nv + v) * slab + idx] = u_old - du;
        }
    }
}

// update_solution: src/transforms.cpp:327-374
__global__ void __launch_bounds__(BS) k_correct(double* __restrict__ U, const double* __restrict__ Ua,
                                                const double* __restrict__ Fx, const double* __restrict__ Fy,
                                                const double* __restrict__ Fz, SD sa, SD fx, SD fy, SD fz,
                                                const double* __restrict__ Dm, int ldd, const double* __restrict__ w,
                                                int q, PredArgs a, uint32_t total) {
    uint32_t idx = blockIdx.x * BS + threadIdx.x;
    if (idx >= total) return;
    int k, j, i, kk, jj, ii;
    split6(idx, sa.Nz, sa.Ny, sa.Nx, sa.nz, sa.ny, sa.nx, k, j, i, kk, jj, ii);
    int T = sa.nt;
    size_t slab = sa.slab();
    for (int v = 0; v < sa.nv; v++) {
        double du = 0;
        for (int t = 0; t < T; t++) {
            double dudt = divergence(Fx, Fy, Fz, fx, fy, fz, Dm, ldd, q, a, t, v, k, j, i, kk, jj, ii);
            if (a.grav) dudt -= grav_src(Ua, (size_t)t * sa.nv * slab, slab, idx, v, a.g);
            du += dudt * w[t] * a.dt;
        }
        U[(size_t)v * slab + idx] -= du;
    }
}

// combine_solution: src/transforms.cpp:252-261
__global__ void __launch_bounds__(BS) k_combine(double* __restrict__ U, const double* __restrict__ U0, SD s, double a,
                                                uint32_t total) {
    uint32_t idx = blockIdx.x * BS + threadIdx.x;
    if (idx >= total) return;
    size_t slab = s.slab();
    for (int v = 0; v < s.nv; v++) {
        size_t o = (size_t)v * slab + idx;
        U[o] = a * U0[o] + (1.0 - a) * U[o];
    }
}

// compute_dt: src/hydro.cpp:208-262, per-block min then a one-block pass.
__global__ void __launch_bounds__(BS) k_dt_partial(const double* __restrict__ W, SD s, int gx, int gy, int gz,
                                                   double gm, double cfl, double dx, double dy, double dz, double nu,
                                                   int ax, int ay, int az, double* __restrict__ part, uint32_t total) {
    __shared__ double red[BS];
    double r = __builtin_inf();
    int Mz = s.Nz - 2 * gz, My = s.Ny - 2 * gy, Mx = s.Nx - 2 * gx;
    size_t slab = s.slab();
    for (uint32_t idx = blockIdx.x * BS + threadIdx.x; idx < total; idx += gridDim.x * BS) {
        int k, j, i, kk, jj, ii;
        split6(idx, Mz, My, Mx, s.nz, s.ny, s.nx, k, j, i, kk, jj, ii);
        size_t o = s.off(0, 0, k + gz, j + gy, i + gx, kk, jj, ii);
        double c_s = sound_speed(W[o], W[4 * slab + o], gm);
        double c_max = 0, dx_min = 1, dt_min;
        if (ax) { c_max += (fabs(W[slab + o]) + c_s); dx_min = smin(dx_min, dx); }
        if (ay) { c_max += (fabs(W[2 * slab + o]) + c_s); dx_min = smin(dx_min, dy); }
        if (az) { c_max += (fabs(W[3 * slab + o]) + c_s); dx_min = smin(dx_min, dz); }
        dt_min = cfl * dx_min / c_max / s.nx;
        if (nu > 0.0) {
            double dx_sub = dx_min / s.nx;
            double dt_visc = 0.25 * cfl * dx_sub * dx_sub / nu;
            dt_min = nan_min(dt_min, dt_visc);
        }
        r = nan_min(r, dt_min);
    }
    red[threadIdx.x] = r;
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = nan_min(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
__global__ void __launch_bounds__(BS) k_min_final(double* __restrict__ part, int n) {
    __shared__ double red[BS];
    double r = __builtin_inf();
    for (int i = threadIdx.x; i < n; i += BS) r = nan_min(r, part[i]);
    red[threadIdx.x] = r;
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = nan_min(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[0] = red[0];
}

// ===========================================================================
// C ABI
extern "C" {

const char* spdk_last_error(void) { return spdk::g_err.c_str(); }

int spdk_device_info(char* name, size_t len, int* cus) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) { set_error("no HIP device"); return SPDK_E_NOGPU; }
    hipDeviceProp_t p;
    SPDK_CHECK(hipGetDeviceProperties(&p, dev));
    if (name && len) { std::strncpy(name, p.gcnArchName, len - 1); name[len - 1] = 0; }
    if (cus) *cus = p.multiProcessorCount;
    return SPDK_OK;
}

int spdk_ctx_create(const spdk_params* prm, const spdk_operators* ops, void* stream, spdk_ctx** out) {
    if (!prm || !ops || !out) { set_error("null argument"); return SPDK_E_ARG; }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) { set_error("no HIP device"); return SPDK_E_NOGPU; }
    hipDeviceProp_t dp;
    int dev = 0;
    SPDK_CHECK(hipGetDevice(&dev));
    SPDK_CHECK(hipGetDeviceProperties(&dp, dev));
    if (std::strncmp(dp.gcnArchName, "gfx950", 6) != 0) {
        set_error(std::string("libspdk is built for gfx950, device is ") + dp.gcnArchName);
        return SPDK_E_NOGPU;
    }
    if (prm->p < 1 || prm->p > 8 || prm->n_ader < 1 || prm->n_ader > 10) { set_error("bad degree"); return SPDK_E_CONFIG; }
    spdk_ctx* c = new spdk_ctx();
    c->prm = *prm;
    c->p = prm->p;
    c->n = prm->p + 1;
    c->m = prm->p + 2;
    c->T = prm->n_ader;
    for (int d = 0; d < 3; d++) {
        c->ngh[d] = prm->active[d] ? 1 : 0;
        c->fgh[d] = prm->active[d] ? 2 : 0;
    }
    if (stream) c->stream = (hipStream_t)stream;
    else {
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; set_error("stream"); return SPDK_E_HIP; }
        c->own_stream = true;
    }
    int n = c->n, m = c->m, T = c->T;
    int rows[SPDK_M_COUNT] = {m, n, n, n, n, n, T}, cols[SPDK_M_COUNT] = {n, m, m, n, n, m, T};
    for (int q = 0; q < SPDK_M_COUNT; q++) {
        c->mrows[q] = rows[q];
        c->mcols[q] = cols[q];
        size_t b = sizeof(double) * rows[q] * cols[q];
        SPDK_CHECK(hipMalloc(&c->mat[q], b));
        if (ops->mat[q]) SPDK_CHECK(hipMemcpy(c->mat[q], ops->mat[q], b, hipMemcpyHostToDevice));
        else SPDK_CHECK(hipMemset(c->mat[q], 0, b));
        if (ops->mat[q]) c->mat_host[q].assign(ops->mat[q], ops->mat[q] + rows[q] * cols[q]);
        else c->mat_host[q].assign((size_t)rows[q] * cols[q], 0.0);
    }
    SPDK_CHECK(hipMalloc(&c->wt, sizeof(double) * T));
    for (int t = 0; t < T; t++) c->wt_host[t] = ops->wt[t];
    SPDK_CHECK(hipMemcpy(c->wt, ops->wt, sizeof(double) * T, hipMemcpyHostToDevice));
    for (int d = 0; d < 3; d++) {
        c->fvnf[d] = ops->fv_nfaces[d];
        c->fvn[d] = ops->fv_ncells[d];
        if (ops->fv_faces[d] && c->fvnf[d] > 0) {
            SPDK_CHECK(hipMalloc(&c->fvf[d], sizeof(double) * c->fvnf[d]));
            SPDK_CHECK(hipMemcpy(c->fvf[d], ops->fv_faces[d], sizeof(double) * c->fvnf[d], hipMemcpyHostToDevice));
        }
        if (ops->fv_centers[d] && c->fvn[d] > 0) {
            SPDK_CHECK(hipMalloc(&c->fvc[d], sizeof(double) * c->fvn[d]));
            SPDK_CHECK(hipMemcpy(c->fvc[d], ops->fv_centers[d], sizeof(double) * c->fvn[d], hipMemcpyHostToDevice));
        }
    }
    SPDK_CHECK(hipMalloc(&c->red, sizeof(double) * (c->red_blocks + 8)));
    SPDK_CHECK(hipHostMalloc(&c->red_host, sizeof(double) * 8, hipHostMallocDefault));
    SPDK_CHECK(hipEventCreate(&c->ev0));
    SPDK_CHECK(hipEventCreate(&c->ev1));
    for (int i = 0; i < spdk_ctx::kEvPool; i++)
        for (int j = 0; j < 2; j++) SPDK_CHECK(hipEventCreate(&c->evs[i][j]));
    *out = c;
    return SPDK_OK;
}

int spdk_ctx_destroy(spdk_ctx* c) {
    if (!c) return SPDK_OK;
    hipStreamSynchronize(c->stream);
    for (int q = 0; q < SPDK_M_COUNT; q++) hipFree(c->mat[q]);
    hipFree(c->wt);
    for (int d = 0; d < 3; d++) { hipFree(c->fvf[d]); hipFree(c->fvc[d]); }
    hipFree(c->red);
    hipHostFree(c->red_host);
    for (void* p : c->owned) hipFree(p);
    for (int i = 0; i < spdk_ctx::kEvPool; i++)
        for (int j = 0; j < 2; j++)
            if (c->evs[i][j]) hipEventDestroy(c->evs[i][j]);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->own_stream) hipStreamDestroy(c->stream);
    delete c;
    return SPDK_OK;
}

int spdk_ctx_sync(spdk_ctx* c) {
    SPDK_CHECK(hipStreamSynchronize(c->stream));
    return SPDK_OK;
}
void* spdk_ctx_stream(spdk_ctx* c) { return (void*)c->stream; }

int spdk_malloc(void** ptr, size_t bytes) {
    if (!ptr) return SPDK_E_ARG;
    SPDK_CHECK(hipMalloc(ptr, bytes ? bytes : 8));
    return SPDK_OK;
}
int spdk_free(void* ptr) {
    SPDK_CHECK(hipFree(ptr));
    return SPDK_OK;
}
int spdk_memset0(spdk_ctx* c, void* ptr, size_t bytes) {
    SPDK_CHECK(hipMemsetAsync(ptr, 0, bytes, c->stream));
    return SPDK_OK;
}
int spdk_memcpy_h2d(spdk_ctx* c, void* dst, const void* src, size_t bytes) {
    SPDK_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    SPDK_CHECK(hipStreamSynchronize(c->stream));
    return SPDK_OK;
}
int spdk_memcpy_d2h(spdk_ctx* c, void* dst, const void* src, size_t bytes) {
    SPDK_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    SPDK_CHECK(hipStreamSynchronize(c->stream));
    return SPDK_OK;
}
int spdk_memcpy_d2d(spdk_ctx* c, void* dst, const void* src, size_t bytes) {
    SPDK_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    return SPDK_OK;
}

static bool shape_ok(const spdk_sd_shape& s) {
    return s.nt > 0 && s.nv > 0 && s.Nz > 0 && s.Ny > 0 && s.Nx > 0 && s.nz > 0 && s.ny > 0 && s.nx > 0 &&
           (uint64_t)s.Nz * s.Ny * s.Nx * s.nz * s.ny * s.nx < (1ull << 32);
}
#define REQUIRE_SHAPE(s)                                                  \
    do {                                                                  \
        if (!shape_ok(s)) { set_error("bad SD shape"); return SPDK_E_SHAPE; } \
    } while (0)

int spdk_transform_1d(spdk_ctx* c, const double* A, spdk_sd_shape sa, double* B, spdk_sd_shape sb, int mat, int dim,
                      int t_src) {
    REQUIRE_SHAPE(sa);
    REQUIRE_SHAPE(sb);
    if (mat < 0 || mat >= SPDK_M_COUNT || dim < 0 || dim > 2 || !A || !B) { set_error("bad transform args"); return SPDK_E_ARG; }
    int q = dim == 0 ? sa.nx : (dim == 1 ? sa.ny : sa.nz);
    int nb = dim == 0 ? sb.nx : (dim == 1 ? sb.ny : sb.nz);
    if (q != c->mcols[mat] || nb > c->mrows[mat] || sa.Nx != sb.Nx || sa.Ny != sb.Ny || sa.Nz != sb.Nz || sa.nv != sb.nv ||
        (t_src >= sa.nt) || (t_src < 0 && sa.nt != sb.nt)) {
        set_error("transform shape mismatch");
        return SPDK_E_SHAPE;
    }
    for (int d = 0; d < 3; d++) {
        int pa = d == 0 ? sa.nx : (d == 1 ? sa.ny : sa.nz), pb = d == 0 ? sb.nx : (d == 1 ? sb.ny : sb.nz);
        if (d != dim && pa != pb) { set_error("transform transverse mismatch"); return SPDK_E_SHAPE; }
    }
    // one thread per (slice, variable, element, line along dim)
    const int lines = (dim == 0 ? sb.ny * sb.nz : (dim == 1 ? sb.nx * sb.nz : sb.nx * sb.ny));
    const size_t rows = (size_t)(t_src < 0 ? sb.nt : 1) * sb.nv * sb.Nz * sb.Ny * sb.Nx;
    const size_t total = rows * lines;
    if (total == 0) return SPDK_OK;
    dim3 g((unsigned)((total + BS - 1) / BS));
    switch (q) {
        case 1: hipLaunchKernelGGL(k_transform_1d<1>, g, dim3(BS), 0, c->stream, A, SD(sa), B, SD(sb), c->mat[mat], c->mcols[mat], dim, t_src, total); break;
        case 2: hipLaunchKernelGGL(k_transform_1d<2>, g, dim3(BS), 0, c->stream, A, SD(sa), B, SD(sb), c->mat[mat], c->mcols[mat], dim, t_src, total); break;
        case 3: hipLaunchKernelGGL(k_transform_1d<3>, g, dim3(BS), 0, c->stream, A, SD(sa), B, SD(sb), c->mat[mat], c->mcols[mat], dim, t_src, total); break;
        case 4: hipLaunchKernelGGL(k_transform_1d<4>, g, dim3(BS), 0, c->stream, A, SD(sa), B, SD(sb), c->mat[mat], c->mcols[mat], dim, t_src, total); break;
        case 5: hipLaunchKernelGGL(k_transform_1d<5>, g, dim3(BS), 0, c->stream, A, SD(sa), B, SD(sb), c->mat[mat], c->mcols[mat], dim, t_src, total); break;
        case 6: hipLaunchKernelGGL(k_transform_1d<6>, g, dim3(BS), 0, c->stream, A, SD(sa), B, SD(sb), c->mat[mat], c->mcols[mat], dim, t_src, total); break;
        case 7: hipLaunchKernelGGL(k_transform_1d<7>, g, dim3(BS), 0, c->stream, A, SD(sa), B, SD(sb), c->mat[mat], c->mcols[mat], dim, t_src, total); break;
        case 8: hipLaunchKernelGGL(k_transform_1d<8>, g, dim3(BS), 0, c->stream, A, SD(sa), B, SD(sb), c->mat[mat], c->mcols[mat], dim, t_src, total); break;
        case 9: hipLaunchKernelGGL(k_transform_1d<9>, g, dim3(BS), 0, c->stream, A, SD(sa), B, SD(sb), c->mat[mat], c->mcols[mat], dim, t_src, total); break;
        case 10: hipLaunchKernelGGL(k_transform_1d<10>, g, dim3(BS), 0, c->stream, A, SD(sa), B, SD(sb), c->mat[mat], c->mcols[mat], dim, t_src, total); break;
        default: set_error("transform: more than 10 points along the axis"); return SPDK_E_SHAPE;
    }
    SPDK_LAUNCH_CHECK("transform_1d");
    return SPDK_OK;
}

int spdk_transform(spdk_ctx* c, const double* A, double* B, double* T, spdk_sd_shape s, int mat) {
    int dims[3], nd = 0;
    for (int d = 0; d < 3; d++)
        if (c->prm.active[d]) dims[nd++] = d;
    // every active-axis sweep of an element in one pass: active axes x, x-y or
    // x-y-z, square matrix, the same point count along every active axis
    REQUIRE_SHAPE(s);
    if (mat < 0 || mat >= SPDK_M_COUNT || !A || !B) { set_error("bad transform args"); return SPDK_E_ARG; }
    const int q = s.nx, act_ok = nd >= 1 && dims[0] == 0 && (nd < 2 || dims[1] == 1) && (nd < 3 || dims[2] == 2);
    const bool pts_ok = (s.ny == (nd >= 2 ? q : 1)) && (s.nz == (nd >= 3 ? q : 1));
    if (act_ok && pts_ok && c->mcols[mat] == q && c->mrows[mat] == q && A != B && (size_t)s.Nz * s.Ny * s.Nx < (1ull << 31)) {
        bool launched = nd == 1 ? launch_transform_el<1>(c, A, B, s, mat)
                                : (nd == 2 ? launch_transform_el<2>(c, A, B, s, mat) : launch_transform_el<3>(c, A, B, s, mat));
        if (launched) {
            SPDK_LAUNCH_CHECK("transform_el");
            return SPDK_OK;
        }
    }
    const double* src = A;
    for (int q = 0; q < nd; q++) {
        double* dst = ((nd - 1 - q) % 2 == 0) ? B : T;
        int rc = spdk_transform_1d(c, src, s, dst, s, mat, dims[q], -1);
        if (rc) return rc;
        src = dst;
    }
    return SPDK_OK;
}

int spdk_conservatives(spdk_ctx* c, const double* W, double* U, spdk_sd_shape s) {
    REQUIRE_SHAPE(s);
    if (s.nv < NVP) return SPDK_E_SHAPE;
    uint32_t total = (uint32_t)SD(s).total_cells();
    hipLaunchKernelGGL(k_cons_prim, dim3(nblocks(total, BS)), dim3(BS), 0, c->stream, W, U, SD(s), c->prm.gamma, 0, total);
    SPDK_LAUNCH_CHECK("conservatives");
    return SPDK_OK;
}
int spdk_primitives(spdk_ctx* c, const double* U, double* W, spdk_sd_shape s) {
    REQUIRE_SHAPE(s);
    if (s.nv < NVP) return SPDK_E_SHAPE;
    uint32_t total = (uint32_t)SD(s).total_cells();
    hipLaunchKernelGGL(k_cons_prim, dim3(nblocks(total, BS)), dim3(BS), 0, c->stream, U, W, SD(s), c->prm.gamma, 1, total);
    SPDK_LAUNCH_CHECK("primitives");
    return SPDK_OK;
}

int spdk_hydro_flux(spdk_ctx* c, const double* U, double* F, spdk_sd_shape s, int dim) {
    REQUIRE_SHAPE(s);
    if (s.nv < NVP || dim < 0 || dim > 2) return SPDK_E_ARG;
    uint32_t total = (uint32_t)SD(s).total_cells();
    dim3 g(nblocks(total, BS));
    if (dim == 0) hipLaunchKernelGGL(k_flux<0>, g, dim3(BS), 0, c->stream, U, F, SD(s), c->prm.gamma, total);
    if (dim == 1) hipLaunchKernelGGL(k_flux<1>, g, dim3(BS), 0, c->stream, U, F, SD(s), c->prm.gamma, total);
    if (dim == 2) hipLaunchKernelGGL(k_flux<2>, g, dim3(BS), 0, c->stream, U, F, SD(s), c->prm.gamma, total);
    SPDK_LAUNCH_CHECK("hydro_flux");
    return SPDK_OK;
}

int spdk_fill_sd_traces(spdk_ctx* c, double* U, spdk_sd_shape s, int dim) {
    REQUIRE_SHAPE(s);
    if (dim < 0 || dim > 2) return SPDK_E_ARG;
    int N = dim == 0 ? s.Nx : (dim == 1 ? s.Ny : s.Nz);
    if (N < 3) { set_error("trace fill needs ghost elements"); return SPDK_E_SHAPE; }
    uint32_t total = (uint32_t)((size_t)(dim == 0 ? 1 : s.Nx) * (dim == 1 ? 1 : s.Ny) * (dim == 2 ? 1 : s.Nz) *
                                (dim == 0 ? 1 : s.nx) * (dim == 1 ? 1 : s.ny) * (dim == 2 ? 1 : s.nz));
    hipLaunchKernelGGL(k_sd_bc, dim3(nblocks(total, BS)), dim3(BS), 0, c->stream, U, SD(s), dim, c->prm.bc[dim], total);
    SPDK_LAUNCH_CHECK("fill_sd_traces");
    return SPDK_OK;
}

int spdk_sd_riemann(spdk_ctx* c, double* U, double* F, spdk_sd_shape s, int dim, int viscous) {
    REQUIRE_SHAPE(s);
    if (s.nv < NVP || dim < 0 || dim > 2 || !U || !F) return SPDK_E_ARG;
    int Nx = s.Nx - (dim == 0), Ny = s.Ny - (dim == 1), Nz = s.Nz - (dim == 2);
    int px = dim == 0 ? 1 : s.nx, py = dim == 1 ? 1 : s.ny, pz = dim == 2 ? 1 : s.nz;
    uint32_t total = (uint32_t)((size_t)Nz * Ny * Nx * pz * py * px);
    if (total == 0) return SPDK_OK;
    dim3 g(nblocks(total, BS));
    const int vis = viscous ? 1 : 0;
    if (dim == 0) hipLaunchKernelGGL(k_riemann<0>, g, dim3(BS), 0, c->stream, U, F, SD(s), c->prm.gamma, vis, total);
    if (dim == 1) hipLaunchKernelGGL(k_riemann<1>, g, dim3(BS), 0, c->stream, U, F, SD(s), c->prm.gamma, vis, total);
    if (dim == 2) hipLaunchKernelGGL(k_riemann<2>, g, dim3(BS), 0, c->stream, U, F, SD(s), c->prm.gamma, vis, total);
    SPDK_LAUNCH_CHECK("riemann_llf");
    return SPDK_OK;
}

int spdk_riemann_llf(spdk_ctx* c, const double* U, double* F, spdk_sd_shape s, int dim) {
    // the inviscid solver never writes U
    return spdk_sd_riemann(c, const_cast<double*>(U), F, s, dim, 0);
}

int spdk_sd_rusanov(spdk_ctx* c, const double* U, double* F, spdk_sd_shape s, int dim) {
    REQUIRE_SHAPE(s);
    if (dim < 0 || dim > 2 || !U || !F) return SPDK_E_ARG;
    int Nx = s.Nx - (dim == 0), Ny = s.Ny - (dim == 1), Nz = s.Nz - (dim == 2);
    int px = dim == 0 ? 1 : s.nx, py = dim == 1 ? 1 : s.ny, pz = dim == 2 ? 1 : s.nz;
    uint32_t total = (uint32_t)((size_t)Nz * Ny * Nx * pz * py * px);
    if (total == 0) return SPDK_OK;
    dim3 g(nblocks(total, BS));
    if (dim == 0) hipLaunchKernelGGL(k_rusanov<0>, g, dim3(BS), 0, c->stream, U, F, SD(s), total);
    if (dim == 1) hipLaunchKernelGGL(k_rusanov<1>, g, dim3(BS), 0, c->stream, U, F, SD(s), total);
    if (dim == 2) hipLaunchKernelGGL(k_rusanov<2>, g, dim3(BS), 0, c->stream, U, F, SD(s), total);
    SPDK_LAUNCH_CHECK("sd_rusanov");
    return SPDK_OK;
}

// The flux-point array of axis dim and a solution-point array of the same
// mesh and slices: the same element counts, one more point along dim.
static bool fp_sp_pair(const spdk_sd_shape& f, const spdk_sd_shape& s, int dim) {
    const int nf[3] = {f.nx, f.ny, f.nz}, ns[3] = {s.nx, s.ny, s.nz};
    if (f.nt != s.nt || f.nv != s.nv || f.Nx != s.Nx || f.Ny != s.Ny || f.Nz != s.Nz) return false;
    for (int d = 0; d < 3; d++)
        if (nf[d] != (d == dim ? ns[d] + 1 : ns[d])) return false;
    return true;
}

int spdk_sd_gradient(spdk_ctx* c, const double* U, spdk_sd_shape s_fp, double* dU, spdk_sd_shape s_sp, int dim) {
    REQUIRE_SHAPE(s_fp);
    REQUIRE_SHAPE(s_sp);
    if (dim < 0 || dim > 2 || !U || !dU || !c->prm.active[dim]) return SPDK_E_ARG;
    if (!fp_sp_pair(s_fp, s_sp, dim) || c->mcols[SPDK_M_DFP_TO_SP] != (dim == 0 ? s_fp.nx : (dim == 1 ? s_fp.ny : s_fp.nz))) {
        set_error("sd_gradient: flux-point / solution-point shapes");
        return SPDK_E_SHAPE;
    }
    int gx = c->ngh[0], gy = c->ngh[1], gz = c->ngh[2];
    uint32_t total = (uint32_t)((size_t)(s_sp.Nz - 2 * gz) * (s_sp.Ny - 2 * gy) * (s_sp.Nx - 2 * gx) * s_sp.nz * s_sp.ny *
                                s_sp.nx);
    if (total == 0) return SPDK_OK;
    dim3 g(nblocks(total, BS));
    const double h = c->prm.h[dim];
    const double* Dm = c->mat[SPDK_M_DFP_TO_SP];
    const int ld = c->mcols[SPDK_M_DFP_TO_SP];
    if (dim == 0) hipLaunchKernelGGL(k_gradient<0>, g, dim3(BS), 0, c->stream, U, SD(s_fp), dU, SD(s_sp), Dm, ld, h, gx, gy, gz, total);
    if (dim == 1) hipLaunchKernelGGL(k_gradient<1>, g, dim3(BS), 0, c->stream, U, SD(s_fp), dU, SD(s_sp), Dm, ld, h, gx, gy, gz, total);
    if (dim == 2) hipLaunchKernelGGL(k_gradient<2>, g, dim3(BS), 0, c->stream, U, SD(s_fp), dU, SD(s_sp), Dm, ld, h, gx, gy, gz, total);
    SPDK_LAUNCH_CHECK("sd_gradient");
    return SPDK_OK;
}

int spdk_viscous_flux(spdk_ctx* c, double* U, spdk_sd_shape s_fp, const double* dU1, const double* dU2, const double* dU3,
                      spdk_sd_shape s_sp, int dim) {
    REQUIRE_SHAPE(s_fp);
    REQUIRE_SHAPE(s_sp);
    if (dim < 0 || dim > 2 || !U || !dU1 || !dU2 || !dU3 || s_fp.nv < NVP || s_fp.nv > NV) return SPDK_E_ARG;
    if (!fp_sp_pair(s_fp, s_sp, dim) || c->mcols[SPDK_M_SP_TO_FP] + 1 != (dim == 0 ? s_fp.nx : (dim == 1 ? s_fp.ny : s_fp.nz))) {
        set_error("viscous_flux: flux-point / solution-point shapes");
        return SPDK_E_SHAPE;
    }
    uint32_t total = (uint32_t)SD(s_fp).total_cells();
    if (total == 0) return SPDK_OK;
    dim3 g(nblocks(total, BS));
    const double* S = c->mat[SPDK_M_SP_TO_FP];
    const int ld = c->mcols[SPDK_M_SP_TO_FP];
    const double gm = c->prm.gamma, nu = c->prm.nu, beta = c->prm.beta;
    if (dim == 0) hipLaunchKernelGGL(k_viscous_flux<0>, g, dim3(BS), 0, c->stream, U, SD(s_fp), dU1, dU2, dU3, SD(s_sp), S, ld, gm, nu, beta, total);
    if (dim == 1) hipLaunchKernelGGL(k_viscous_flux<1>, g, dim3(BS), 0, c->stream, U, SD(s_fp), dU1, dU2, dU3, SD(s_sp), S, ld, gm, nu, beta, total);
    if (dim == 2) hipLaunchKernelGGL(k_viscous_flux<2>, g, dim3(BS), 0, c->stream, U, SD(s_fp), dU1, dU2, dU3, SD(s_sp), S, ld, gm, nu, beta, total);
    SPDK_LAUNCH_CHECK("viscous_flux");
    return SPDK_OK;
}

int spdk_copy_ader(spdk_ctx* c, const double* U, double* Ua, spdk_sd_shape s) {
    REQUIRE_SHAPE(s);
    uint32_t total = (uint32_t)SD(s).total_cells();
    hipLaunchKernelGGL(k_copy_ader, dim3(nblocks(total, BS)), dim3(BS), 0, c->stream, U, Ua, SD(s), total);
    SPDK_LAUNCH_CHECK("copy_ader");
    return SPDK_OK;
}

static void flux_shapes(spdk_ctx* c, const spdk_sd_shape& s, SD* f) {
    for (int d = 0; d < 3; d++) {
        f[d] = SD(s);
        bool a = c->prm.active[d];
        if (d == 0) f[d].nx = a ? c->m : s.nx;
        if (d == 1) f[d].ny = a ? c->m : s.ny;
        if (d == 2) f[d].nz = a ? c->m : s.nz;
    }
}
static PredArgs pred_args(spdk_ctx* c, double dt) {
    PredArgs a;
    for (int d = 0; d < 3; d++) {
        a.g[d] = c->prm.g[d];
        a.h[d] = c->prm.h[d];
        a.act[d] = c->prm.active[d];
    }
    a.grav = (a.g[0] != 0.0 || a.g[1] != 0.0 || a.g[2] != 0.0);
    a.dt = dt;
    return a;
}

int spdk_ader_predict(spdk_ctx* c, const double* U, double* Ua, const double* Fx, const double* Fy, const double* Fz,
                      spdk_sd_shape s, int q, double dt) {
    REQUIRE_SHAPE(s);
    if (s.nt > 10 || s.nt != c->T) { set_error("predict: slice count"); return SPDK_E_SHAPE; }
    SD f[3];
    flux_shapes(c, s, f);
    uint32_t total = (uint32_t)SD(s).total_cells();
    hipLaunchKernelGGL(k_predict, dim3(nblocks(total, BS)), dim3(BS), 0, c->stream, U, Ua, Fx, Fy ? Fy : Fx, Fz ? Fz : Fx,
                       SD(s), f[0], f[1], f[2], c->mat[SPDK_M_DFP_TO_SP], c->mcols[SPDK_M_DFP_TO_SP],
                       c->mat[SPDK_M_INVADER], c->wt, q, pred_args(c, dt), total);
    SPDK_LAUNCH_CHECK("ader_predict");
    return SPDK_OK;
}

int spdk_ader_correct(spdk_ctx* c, double* U, const double* Ua, const double* Fx, const double* Fy, const double* Fz,
                      spdk_sd_shape s, int q, double dt) {
    REQUIRE_SHAPE(s);
    if (s.nt != c->T) { set_error("correct: slice count"); return SPDK_E_SHAPE; }
    SD f[3];
    flux_shapes(c, s, f);
    uint32_t total = (uint32_t)SD(s).total_cells();
    hipLaunchKernelGGL(k_correct, dim3(nblocks(total, BS)), dim3(BS), 0, c->stream, U, Ua, Fx, Fy ? Fy : Fx, Fz ? Fz : Fx,
                       SD(s), f[0], f[1], f[2], c->mat[SPDK_M_DFP_TO_SP], c->mcols[SPDK_M_DFP_TO_SP], c->wt, q,
                       pred_args(c, dt), total);
    SPDK_LAUNCH_CHECK("ader_correct");
    return SPDK_OK;
}

int spdk_rk_combine(spdk_ctx* c, double* U, const double* U0, spdk_sd_shape s, double a) {
    REQUIRE_SHAPE(s);
    uint32_t total = (uint32_t)SD(s).total_cells();
    hipLaunchKernelGGL(k_combine, dim3(nblocks(total, BS)), dim3(BS), 0, c->stream, U, U0, SD(s), a, total);
    SPDK_LAUNCH_CHECK("rk_combine");
    return SPDK_OK;
}

int spdk_min_dt(spdk_ctx* c, const double* W, spdk_sd_shape s, double* dt_host) {
    REQUIRE_SHAPE(s);
    int gx = c->ngh[0], gy = c->ngh[1], gz = c->ngh[2];
    uint32_t total = (uint32_t)((size_t)(s.Nz - 2 * gz) * (s.Ny - 2 * gy) * (s.Nx - 2 * gx) * s.nz * s.ny * s.nx);
    int nb = (int)std::min<size_t>(c->red_blocks, nblocks(total, BS));
    if (nb < 1) nb = 1;
    const spdk_params& P = c->prm;
    hipLaunchKernelGGL(k_dt_partial, dim3(nb), dim3(BS), 0, c->stream, W, SD(s), gx, gy, gz, P.gamma, P.cfl, P.h[0], P.h[1],
                       P.h[2], P.nu, P.active[0], P.active[1], P.active[2], c->red, total);
    SPDK_LAUNCH_CHECK("dt_partial");
    hipLaunchKernelGGL(k_min_final, dim3(1), dim3(BS), 0, c->stream, c->red, nb);
    SPDK_LAUNCH_CHECK("min_final");
    SPDK_CHECK(hipMemcpyAsync(c->red_host, c->red, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SPDK_CHECK(hipStreamSynchronize(c->stream));
    if (dt_host) *dt_host = c->red_host[0];
    return SPDK_OK;
}

}  // extern "C"

extern "C" {
int spdk_timing_enable(spdk_ctx* c, int on) {
    if (!c) return SPDK_E_ARG;
    c->timing = on != 0;
    return SPDK_OK;
}
int spdk_timing_reset(spdk_ctx* c) {
    if (!c) return SPDK_E_ARG;
    c->kernel_ms = 0;
    c->kernel_launches = 0;
    return SPDK_OK;
}
int spdk_timing_read(spdk_ctx* c, double* ms, int* launches, char* name, size_t len) {
    if (!c) return SPDK_E_ARG;
    if (ms) *ms = c->kernel_ms;
    if (launches) *launches = c->kernel_launches;
    if (name && len) {
        std::strncpy(name, c->kernel_name, len - 1);
        name[len - 1] = 0;
    }
    return SPDK_OK;
}
}  // extern "C"
