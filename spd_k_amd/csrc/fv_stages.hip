// FV fallback stage kernels: SD flux integration over sub-cell faces, the FV
// update, FV ghost fills, trouble detection (NAD / SED / relax / PAD), theta
// blending and the blended MUSCL-Hancock + HLLC face fluxes. Each kernel
// restates the reference function cited at its launcher.
#include <hip/hip_runtime.h>
#include <algorithm>
#include "ctx.hpp"
#include "physics.hpp"

// Launch-shape constants of the 3-D FV kernels (measured at 96^3 elements):
// cells per z column of the update and trouble-detection kernels, and the
// x and z extents of the fallback's cell tile (y makes up the 256 threads).
#ifndef SPDK_UPD_KC
#define SPDK_UPD_KC 8
#endif
#ifndef SPDK_NAD_KC
#define SPDK_NAD_KC 8
#endif
#ifndef SPDK_FB_TX
#define SPDK_FB_TX 8
#endif
#ifndef SPDK_FB_TZ
#define SPDK_FB_TZ 4
#endif
// waves per SIMD the fallback tile kernel is compiled for (4: 126 VGPRs, no
// spills; 3.75 -> 3.45 ms per call at 96^3 against the unbounded 160)
#ifndef SPDK_FB_WAVES
#define SPDK_FB_WAVES 4
#endif

using namespace spdk;
static constexpr int BS = 256;

// Workgroups are dispatched to the 8 XCDs in turn. The stencil kernels give
// XCD x the x-th contiguous eighth of their blocks, so the rows and planes
// that neighbouring cells share are read through one L2; the grid is padded
// to a multiple of 8 (xcd_grid) and blocks past the range exit on the index
// check.
constexpr unsigned kXcd = 8;
static inline unsigned xcd_grid(size_t total) { return (nblocks(total, BS) + kXcd - 1) / kXcd * kXcd; }
__device__ __forceinline__ uint32_t xcd_block() { return (blockIdx.x % kXcd) * (gridDim.x / kXcd) + blockIdx.x / kXcd; }

struct Geo {  // FV sub-grid coordinates per axis (device pointers)
    const double* f[3];
    const double* c[3];
};

// ---------------------------------------------------------------------------
// face_integral_ref: src/transforms.cpp:378-427 (at most one transverse axis active)
__global__ void __launch_bounds__(BS) k_face_int_ref(const double* __restrict__ Ffp, SD s, double* __restrict__ F, FV sf,
                                                     const double* __restrict__ S, int lds, int t_id, int dim, int dim1,
                                                     int dim2, int a1, int a2, int gx, int gy, int gz, int fx, int fy,
                                                     int fz, uint32_t total) {
    uint32_t idx = blockIdx.x * BS + threadIdx.x;
    if (idx >= total) return;
    int Nx = s.Nx + (dim == 0), Ny = s.Ny + (dim == 1), Nz = s.Nz + (dim == 2);
    int px = s.nx - (dim == 0), py = s.ny - (dim == 1), pz = s.nz - (dim == 2);
    int k, j, i, kk, jj, ii;
    split6(idx, Nz - 2 * gz, Ny - 2 * gy, Nx - 2 * gx, pz, py, px, k, j, i, kk, jj, ii);
    k += gz; j += gy; i += gx;
    int I = ii + fx + (i - gx) * px, J = jj + fy + (j - gy) * py, K = kk + fz + (k - gz) * pz;
    if (!(K < sf.Nz && J < sf.Ny && I < sf.Nx)) return;
    int q1 = dim1 == 0 ? px : (dim1 == 1 ? py : pz), q2 = dim2 == 0 ? px : (dim2 == 1 ? py : pz);
    int id1 = dim1 == 0 ? ii : (dim1 == 1 ? jj : kk), id2 = dim2 == 0 ? ii : (dim2 == 1 ? jj : kk);
    for (int v = 0; v < s.nv; v++) {
        double f = 0;
        for (int nn = 0; nn < q2; nn++) {
            int a0 = dim2 == 0 ? nn : ii, b0 = dim2 == 1 ? nn : jj, c0 = dim2 == 2 ? nn : kk;
            for (int ll = 0; ll < q1; ll++) {
                int a = dim1 == 0 ? ll : a0, b = dim1 == 1 ? ll : b0, c = dim1 == 2 ? ll : c0;
                double x = Ffp[s.off(t_id, v, k, j, i, c, b, a)];
                if (a1) x *= S[id1 * lds + ll];
                if (a2) x *= S[id2 * lds + nn];
                f += x;
            }
        }
        F[sf.off(v, K, J, I)] = f;
    }
}

// face_integral with both transverse axes active, the two sweeps of
// src/transforms.cpp:435-480 on element tiles: a workgroup owns EX consecutive
// elements of the (k, j, i) element range; it stages their flux-point panels
// of one variable in LDS with contiguous loads, forms the dim1 sweep for every
// (face, dim1 output, dim2 point) once, and writes the dim2 sweep with x
// fastest, so the FV face array is written in runs of EX*N values. Each sum
// runs in the reference's order (the dim1 sum over ll of F S(o1, ll), then
// the dim2 sum over nn of those times S(o2, nn)), so the result equals the
// two-pass one bit for bit without the intermediate array.
template <int Q>
struct FaceTile {
    static constexpr int EX = Q <= 4 ? 16 : (Q <= 6 ? 8 : 4);  // elements per workgroup (LDS below 64 KB)
    static constexpr int P = (Q + 1) * Q * Q;     // flux points of one element, one variable
    static constexpr int U = Q * Q * Q;           // dim1-swept values (= FV faces) per element
};
template <int Q>
__global__ void __launch_bounds__(BS) k_face_int_tile(const double* __restrict__ Ffp, SD s, double* __restrict__ F,
                                                      FV sf, const double* __restrict__ S, int lds, int t_id, int dim,
                                                      int gx, int gy, int gz, int fx, int fy, int fz, int ex, int ey,
                                                      uint32_t n_el) {
    using TL = FaceTile<Q>;
    __shared__ double sF[TL::EX * TL::P];
    __shared__ double sU[TL::EX * TL::U];
    __shared__ double sS[Q * Q];
    __shared__ size_t sbase[TL::EX];
    __shared__ int sIJK[TL::EX][3];
    const int tid = threadIdx.x;
    const uint32_t e0 = blockIdx.x * TL::EX;
    const int ne = (int)min((uint32_t)TL::EX, n_el - e0);
    if (tid < Q * Q) sS[tid] = S[(tid / Q) * lds + tid % Q];
    if (tid < TL::EX) {
        int i = -1, j = 0, k = 0;
        if (tid < ne) {
            const uint32_t e = e0 + tid;
            i = gx + (int)(e % ex); j = gy + (int)((e / ex) % ey); k = gz + (int)(e / ((uint32_t)ex * ey));
            // an element past the stored range (possible only along dim) has no FV face in range
            if (i >= s.Nx || j >= s.Ny || k >= s.Nz) i = -1;
        }
        sIJK[tid][0] = i; sIJK[tid][1] = j; sIJK[tid][2] = k;
        sbase[tid] = i < 0 ? 0 : s.off(t_id, 0, k, j, i, 0, 0, 0);
    }
    // panel extents: Q + 1 flux points along dim, Q along the others
    const int Lx = dim == 0 ? Q + 1 : Q, Ly = dim == 1 ? Q + 1 : Q;
    const int dim1 = dim == 0 ? 1 : (dim == 1 ? 2 : 0), dim2 = dim1 == 0 ? 1 : (dim1 == 1 ? 2 : 0);
    const int st[3] = {1, Lx, Lx * Ly};
    const int sd = st[dim], s1 = st[dim1], s2 = st[dim2];
    const size_t slab = s.slab();
    __syncthreads();
    for (int v = 0; v < s.nv; v++) {
        for (int q = tid; q < ne * TL::P; q += BS) {
            const int e = q / TL::P, p = q - e * TL::P;
            if (sIJK[e][0] >= 0) sF[q] = Ffp[sbase[e] + (size_t)v * slab + p];
        }
        __syncthreads();
        // dim1 sweep: sU[e][f][o1][nn] = sum_ll F(f, ll, nn) S(o1, ll)
        for (int q = tid; q < ne * TL::U; q += BS) {
            const int e = q / TL::U, r = q - e * TL::U;
            const int nn = r % Q, o1 = (r / Q) % Q, f = r / (Q * Q);
            const double* Fe = sF + e * TL::P + f * sd + nn * s2;
            double u = 0;
#pragma unroll
            for (int ll = 0; ll < Q; ll++) u += Fe[ll * s1] * sS[o1 * Q + ll];
            sU[q] = u;
        }
        __syncthreads();
        // dim2 sweep and the write, x fastest: q -> (kk, jj, e, ii)
        for (int q = tid; q < ne * TL::U; q += BS) {
            const int ii = q % Q, e = (q / Q) % ne, jj = (q / (Q * ne)) % Q, kk = q / (Q * Q * ne);
            const int i = sIJK[e][0];
            if (i < 0) continue;
            const int j = sIJK[e][1], k = sIJK[e][2];
            const int I = ii + fx + (i - gx) * Q, J = jj + fy + (j - gy) * Q, K = kk + fz + (k - gz) * Q;
            if (!(K < sf.Nz && J < sf.Ny && I < sf.Nx)) continue;
            const int f = dim == 0 ? ii : (dim == 1 ? jj : kk);
            const int o1 = dim1 == 0 ? ii : (dim1 == 1 ? jj : kk), o2 = dim2 == 0 ? ii : (dim2 == 1 ? jj : kk);
            const double* Ue = sU + e * TL::U + (f * Q + o1) * Q;
            double acc = 0;
#pragma unroll
            for (int nn = 0; nn < Q; nn++) acc += Ue[nn] * sS[o2 * Q + nn];
            F[sf.off(v, K, J, I)] = acc;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// fv_update_solution: src/transforms.cpp:483-543
struct FvUpd {
    int act[3], grav;
    double g[3];
    double wdt;  // w(t_id)*dt
    int update;
};
// One FV cell (K, J, I), SD position o: U_old = U_cv (write_old), U_new =
// U_old - w dt (sum_d dF_d / h_d - S), and with update U_cv = U_new.
__device__ __forceinline__ void fv_cell(double* __restrict__ Un, double* __restrict__ Uo, double* __restrict__ Ucv,
                                        const SD& s, const double* __restrict__ Fx, const double* __restrict__ Fy,
                                        const double* __restrict__ Fz, const FV& sc, const FV& sx, const FV& sy,
                                        const FV& sz, const Geo& geo, const FvUpd& a, int K, int J, int I, size_t o,
                                        bool write_old, bool update) {
    const size_t slab = s.slab();
    for (int v = 0; v < s.nv; v++) {
        double h, F = 0;
        double uo = Ucv[v * slab + o];
        if (write_old) Uo[sc.off(v, K, J, I)] = uo;
        if (a.act[0]) { h = geo.f[0][I + 1] - geo.f[0][I]; F += (Fx[sx.off(v, K, J, I + 1)] - Fx[sx.off(v, K, J, I)]) / h; }
        if (a.act[1]) { h = geo.f[1][J + 1] - geo.f[1][J]; F += (Fy[sy.off(v, K, J + 1, I)] - Fy[sy.off(v, K, J, I)]) / h; }
        if (a.act[2]) { h = geo.f[2][K + 1] - geo.f[2][K]; F += (Fz[sz.off(v, K + 1, J, I)] - Fz[sz.off(v, K, J, I)]) / h; }
        if (a.grav) {
            double S = 0.0;
            if (v == 1) S = Ucv[o] * a.g[0];
            else if (v == 2) S = Ucv[o] * a.g[1];
            else if (v == 3) S = Ucv[o] * a.g[2];
            else if (v == 4) S = Ucv[slab + o] * a.g[0] + Ucv[2 * slab + o] * a.g[1] + Ucv[3 * slab + o] * a.g[2];
            F -= S;
        }
        double un = uo - a.wdt * F;
        Un[sc.off(v, K, J, I)] = un;
        if (update) Ucv[v * slab + o] = un;
    }
}

// One thread per active FV cell, x fastest: the FV arrays (fluxes, old and
// new state) are read and written in contiguous runs; the SD-layout state in
// runs of one element row.
struct FvIdx {
    int K, J, I;
    size_t o;
};
__device__ __forceinline__ FvIdx fv_idx(uint32_t idx, const SD& s, int gx, int gy, int gz, int fx, int fy, int fz) {
    const int Mx = (s.Nx - 2 * gx) * s.nx, My = (s.Ny - 2 * gy) * s.ny;
    const int ic = (int)(idx % Mx), jc = (int)((idx / Mx) % My), kc = (int)(idx / ((uint32_t)Mx * My));
    const int i = gx + ic / s.nx, ii = ic % s.nx, j = gy + jc / s.ny, jj = jc % s.ny, k = gz + kc / s.nz, kk = kc % s.nz;
    FvIdx r;
    r.I = ii + fx + (i - gx) * s.nx;
    r.J = jj + fy + (j - gy) * s.ny;
    r.K = kk + fz + (k - gz) * s.nz;
    r.o = s.off(0, 0, k, j, i, kk, jj, ii);
    return r;
}

__global__ void __launch_bounds__(BS) k_fv_update(double* __restrict__ Un, double* __restrict__ Uo, double* __restrict__ Ucv,
                                                  SD s, const double* __restrict__ Fx, const double* __restrict__ Fy,
                                                  const double* __restrict__ Fz, FV sc, FV sx, FV sy, FV sz, Geo geo,
                                                  FvUpd a, int gx, int gy, int gz, int fx, int fy, int fz, uint32_t total) {
    uint32_t idx = xcd_block() * BS + threadIdx.x;
    if (idx >= total) return;
    const FvIdx c = fv_idx(idx, s, gx, gy, gz, fx, fy, fz);
    fv_cell(Un, Uo, Ucv, s, Fx, Fy, Fz, sc, sx, sy, sz, geo, a, c.K, c.J, c.I, c.o, true, a.update != 0);
}

// The update = 0 call followed by compute_primitives of U_new and U_old on the
// active cells: W_new = prim(U_new), W_old = prim(U_cv), U_new written, U_old
// not (its only reader is that conversion). With ghosts filled on W instead
// of U the result is the same: a ghost is a copy of an active cell, with the
// normal momentum negated for reflective walls, and the conversion maps the
// negated momentum to the exactly negated velocity and the same pressure.
__global__ void __launch_bounds__(BS) k_fv_update_prims(double* __restrict__ Un, double* __restrict__ Wn,
                                                        double* __restrict__ Wo, const double* __restrict__ Ucv, SD s,
                                                        const double* __restrict__ Fx, const double* __restrict__ Fy,
                                                        const double* __restrict__ Fz, FV sc, FV sx, FV sy, FV sz,
                                                        Geo geo, FvUpd a, double gm, int gx, int gy, int gz, int fx,
                                                        int fy, int fz, uint32_t total) {
    uint32_t idx = xcd_block() * BS + threadIdx.x;
    if (idx >= total) return;
    const FvIdx c = fv_idx(idx, s, gx, gy, gz, fx, fy, fz);
    const int K = c.K, J = c.J, I = c.I;
    const size_t o = c.o, slab = s.slab();
    double uo[NV], un[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
        double h, F = 0;
        uo[v] = Ucv[v * slab + o];
        if (a.act[0]) { h = geo.f[0][I + 1] - geo.f[0][I]; F += (Fx[sx.off(v, K, J, I + 1)] - Fx[sx.off(v, K, J, I)]) / h; }
        if (a.act[1]) { h = geo.f[1][J + 1] - geo.f[1][J]; F += (Fy[sy.off(v, K, J + 1, I)] - Fy[sy.off(v, K, J, I)]) / h; }
        if (a.act[2]) { h = geo.f[2][K + 1] - geo.f[2][K]; F += (Fz[sz.off(v, K + 1, J, I)] - Fz[sz.off(v, K, J, I)]) / h; }
        if (a.grav) {
            double S = 0.0;
            if (v == 1) S = Ucv[o] * a.g[0];
            else if (v == 2) S = Ucv[o] * a.g[1];
            else if (v == 3) S = Ucv[o] * a.g[2];
            else if (v == 4) S = Ucv[slab + o] * a.g[0] + Ucv[2 * slab + o] * a.g[1] + Ucv[3 * slab + o] * a.g[2];
            F -= S;
        }
        un[v] = uo[v] - a.wdt * F;
        Un[sc.off(v, K, J, I)] = un[v];
    }
    double wn[NVP], wo[NVP];
    prim(un, wn, gm);
    prim(uo, wo, gm);
#pragma unroll
    for (int v = 0; v < NVP; v++) {
        Wn[sc.off(v, K, J, I)] = wn[v];
        Wo[sc.off(v, K, J, I)] = wo[v];
    }
#pragma unroll
    for (int v = NVP; v < NV; v++) {
        Wn[sc.off(v, K, J, I)] = 0.0;
        Wo[sc.off(v, K, J, I)] = 0.0;
    }
}

// The update = 1 call after the update = 0 call of the same slice and the
// blend: a face flux changed only where its blend weight max(theta of the two
// cells) is nonzero, so a cell whose theta and face neighbours' theta are all
// zero keeps the update = 0 result, already in U_new, and only U_cv = U_new
// is written; the other cells are recomputed. U_old already holds U_cv. With
// gravity every cell is recomputed: the update = 1 form takes the source of
// the momenta and energy from the density it has just written to U_cv.
__global__ void __launch_bounds__(BS) k_fv_update_blended(double* __restrict__ Un, double* __restrict__ Ucv, SD s,
                                                          const double* __restrict__ Fx, const double* __restrict__ Fy,
                                                          const double* __restrict__ Fz, const double* __restrict__ th,
                                                          FV sc, FV sx, FV sy, FV sz, Geo geo, FvUpd a, int gx, int gy,
                                                          int gz, int fx, int fy, int fz, uint32_t total) {
    uint32_t idx = xcd_block() * BS + threadIdx.x;
    if (idx >= total) return;
    const FvIdx c = fv_idx(idx, s, gx, gy, gz, fx, fy, fz);
    const size_t sy1 = (size_t)sc.Nx, sz1 = (size_t)sc.Nx * sc.Ny;
    const size_t t0 = ((size_t)c.K * sc.Ny + c.J) * sc.Nx + c.I;
    bool touched = a.grav || th[t0] != 0.0;
    if (a.act[0]) touched = touched || th[t0 - 1] != 0.0 || th[t0 + 1] != 0.0;
    if (a.act[1]) touched = touched || th[t0 - sy1] != 0.0 || th[t0 + sy1] != 0.0;
    if (a.act[2]) touched = touched || th[t0 - sz1] != 0.0 || th[t0 + sz1] != 0.0;
    if (touched) {
        fv_cell(Un, nullptr, Ucv, s, Fx, Fy, Fz, sc, sx, sy, sz, geo, a, c.K, c.J, c.I, c.o, false, true);
        return;
    }
    const size_t slab = s.slab();
    for (int v = 0; v < s.nv; v++) Ucv[v * slab + c.o] = Un[sc.off(v, c.K, c.J, c.I)];
}

// ---------------------------------------------------------------------------
// The FV state across the ADER nodes, in the FV cell layout
// (src/hydro_ader.hpp:500-575). The cell values of the step's start are
// formed once from U_sp and written at their FV cell positions; every node
// then reads and writes contiguous rows, and two state buffers take turns:
// the node's update = 0 result goes to the other buffer, the blended update
// rewrites only the cells whose face fluxes the blend changed, and the
// buffers swap. At the end the cell values go back to the solution points.
// The arithmetic of every value is that of the SD-layout sequence
// (transform_a_to_b, fv_update_solution with update = 0 and 1, U_cv = U_new).

// Elements of one workgroup: EPB consecutive elements of one row along x.
template <int Q, int D>
struct ElTile {
    static constexpr int TPE = D == 1 ? Q : (D == 2 ? Q * Q : Q * Q * Q);
    static constexpr int EPB = TPE >= BS ? 1 : BS / TPE;
};

// transform_a_to_b of one element (src/transforms.cpp:81-118 per axis, x then
// y then z): point s of the element enters as `in`, its transformed value is
// returned. b0/b1 are the element's two LDS planes; every thread of the block
// calls it (it holds barriers).
template <int Q, int D>
__device__ __forceinline__ double el_transform(double in, double* b0, double* b1, const double* sM, int s, bool live) {
    const int ii = s % Q, jj = (s / Q) % Q, kk = s / (Q * Q);
    __syncthreads();
    if (live) b0[s] = in;
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
    return u;
}

// Slot q of a workgroup's FV rows: the row (jj, kk) of the EPB elements, x
// fastest, so the elements' cells of one row are one contiguous run.
template <int Q, int D>
__device__ __forceinline__ void fv_slot(int q, int& e, int& pt, int& ii, int& jj, int& kk) {
    constexpr int EPB = ElTile<Q, D>::EPB;
    ii = q % Q;
    e = (q / Q) % EPB;
    const int r = q / (Q * EPB);
    jj = D >= 2 ? r % Q : 0;
    kk = D == 3 ? r / Q : 0;
    pt = (kk * Q + jj) * Q + ii;
}

// U_sp (SD layout, active elements) -> cell values at the FV cell positions
// of U_fv. Workgroup g: row g / gpr of active elements, elements
// (g % gpr) * EPB ... of that row.
template <int Q, int D>
__global__ void __launch_bounds__(BS) k_fv_state_in(const double* __restrict__ Usp, SD s, double* __restrict__ Ufv, FV sc,
                                                    const double* __restrict__ M, int ldm, int gx, int gy, int gz,
                                                    int fx, int fy, int fz, int ex, int ey, int gpr, uint32_t groups) {
    using TL = ElTile<Q, D>;
    constexpr int TPE = TL::TPE, EPB = TL::EPB;
    __shared__ double b0[EPB * TPE], b1[EPB * TPE], sM[Q * Q];
    const uint32_t g = xcd_block();  // neighbouring groups share FV row lines: one L2
    if (g >= groups) return;
    const int tid = threadIdx.x, el = tid / TPE, sp = tid % TPE;
    const int row = (int)(g / gpr), i0 = (int)(g % gpr) * EPB;
    const int j = row % ey, k = row / ey;
    const int i = i0 + el;
    const bool live = el < EPB && i < ex;
    if (tid < Q * Q) sM[tid] = M[(tid / Q) * ldm + tid % Q];
    const size_t slab = s.slab();
    const size_t eo = ((((size_t)(k + gz) * s.Ny + (j + gy)) * s.Nx) + (i + gx)) * TPE + sp;
    double* e0 = b0 + (live ? el : 0) * TPE;
    double* e1 = b1 + (live ? el : 0) * TPE;
    for (int v = 0; v < NVP; v++) {
        const double u = el_transform<Q, D>(live ? Usp[v * slab + eo] : 0.0, e0, e1, sM, sp, live);
        __syncthreads();
        if (live) b1[el * TPE + sp] = u;
        __syncthreads();
        if (tid < EPB * TPE) {
            int e, pt, ii, jj, kk;
            fv_slot<Q, D>(tid, e, pt, ii, jj, kk);
            if (i0 + e < ex) {
                const int I = fx + (i0 + e) * Q + ii;
                const int J = D >= 2 ? fy + j * Q + jj : fy + j;
                const int K = D == 3 ? fz + k * Q + kk : fz + k;
                Ufv[sc.off(v, K, J, I)] = b1[e * TPE + pt];
            }
        }
    }
}

// The update = 0 call on the FV state (k_fv_update_prims with U_cv read in
// the FV cell layout): U_next, W_new = prim(U_next), W_old = prim(U_fv).
// Variable 5 (zero throughout) is neither read nor written; of W_new only
// the variables in wn_vars are written (trouble detection reads density and
// pressure).
__global__ void __launch_bounds__(BS) k_fv_state_update_prims(double* __restrict__ Un, double* __restrict__ Wn,
                                                              double* __restrict__ Wo, const double* __restrict__ Ucv,
                                                              const double* __restrict__ Fx,
                                                              const double* __restrict__ Fy,
                                                              const double* __restrict__ Fz, FV sc, FV sx, FV sy, FV sz,
                                                              Geo geo, FvUpd a, double gm, int fx, int fy, int fz, int Mx,
                                                              int My, unsigned wn_vars, uint32_t total) {
    uint32_t idx = xcd_block() * BS + threadIdx.x;
    if (idx >= total) return;
    const int I = fx + (int)(idx % Mx), J = fy + (int)((idx / Mx) % My), K = fz + (int)(idx / ((uint32_t)Mx * My));
    const size_t o = sc.off(0, K, J, I), pl = sc.plane();
    double uo[NVP], un[NVP];
#pragma unroll
    for (int v = 0; v < NVP; v++) {
        double h, F = 0;
        uo[v] = Ucv[v * pl + o];
        if (a.act[0]) { h = geo.f[0][I + 1] - geo.f[0][I]; F += (Fx[sx.off(v, K, J, I + 1)] - Fx[sx.off(v, K, J, I)]) / h; }
        if (a.act[1]) { h = geo.f[1][J + 1] - geo.f[1][J]; F += (Fy[sy.off(v, K, J + 1, I)] - Fy[sy.off(v, K, J, I)]) / h; }
        if (a.act[2]) { h = geo.f[2][K + 1] - geo.f[2][K]; F += (Fz[sz.off(v, K + 1, J, I)] - Fz[sz.off(v, K, J, I)]) / h; }
        if (a.grav) {
            double S = 0.0;
            if (v == 1) S = uo[0] * a.g[0];
            else if (v == 2) S = uo[0] * a.g[1];
            else if (v == 3) S = uo[0] * a.g[2];
            else if (v == 4) S = uo[1] * a.g[0] + uo[2] * a.g[1] + uo[3] * a.g[2];
            F -= S;
        }
        un[v] = uo[v] - a.wdt * F;
        Un[v * pl + o] = un[v];
    }
    double wn[NVP], wo[NVP];
    prim(un, wn, gm);
    prim(uo, wo, gm);
#pragma unroll
    for (int v = 0; v < NVP; v++) {
        if ((wn_vars >> v) & 1u) Wn[v * pl + o] = wn[v];
        Wo[v * pl + o] = wo[v];
    }
}

// The same update with z active, one thread per (x, y) column walking KC
// cells up in z: the z face shared by two cells is read once and carried in
// registers (in the one-thread-per-cell form the two readers are a plane
// apart in dispatch order, and the second read misses L2). Same arithmetic
// per cell.
template <int KC>
__global__ void __launch_bounds__(BS) k_fv_state_update_prims_z(double* __restrict__ Un, double* __restrict__ Wn,
                                                                double* __restrict__ Wo, const double* __restrict__ Ucv,
                                                                const double* __restrict__ Fx,
                                                                const double* __restrict__ Fy,
                                                                const double* __restrict__ Fz, FV sc, FV sx, FV sy,
                                                                FV sz, Geo geo, FvUpd a, double gm, int fx, int fy,
                                                                int fz, int Mx, int My, int Mz, unsigned wn_vars,
                                                                uint32_t total) {
    uint32_t idx = xcd_block() * BS + threadIdx.x;
    if (idx >= total) return;  // total = Mx * My * ceil(Mz / KC)
    const int I = fx + (int)(idx % Mx), J = fy + (int)((idx / Mx) % My);
    const int k0 = (int)(idx / ((uint32_t)Mx * My)) * KC;
    const size_t pl = sc.plane();
    double flo[NVP];  // the lower z face of the current cell
#pragma unroll
    for (int v = 0; v < NVP; v++) flo[v] = Fz[sz.off(v, fz + k0, J, I)];
    for (int kk = 0; kk < KC && k0 + kk < Mz; kk++) {
        const int K = fz + k0 + kk;
        const size_t o = sc.off(0, K, J, I);
        double uo[NVP], un[NVP];
#pragma unroll
        for (int v = 0; v < NVP; v++) {
            double h, F = 0;
            uo[v] = Ucv[v * pl + o];
            if (a.act[0]) { h = geo.f[0][I + 1] - geo.f[0][I]; F += (Fx[sx.off(v, K, J, I + 1)] - Fx[sx.off(v, K, J, I)]) / h; }
            if (a.act[1]) { h = geo.f[1][J + 1] - geo.f[1][J]; F += (Fy[sy.off(v, K, J + 1, I)] - Fy[sy.off(v, K, J, I)]) / h; }
            const double fup = Fz[sz.off(v, K + 1, J, I)];
            h = geo.f[2][K + 1] - geo.f[2][K];
            F += (fup - flo[v]) / h;
            flo[v] = fup;
            if (a.grav) {
                double S = 0.0;
                if (v == 1) S = uo[0] * a.g[0];
                else if (v == 2) S = uo[0] * a.g[1];
                else if (v == 3) S = uo[0] * a.g[2];
                else if (v == 4) S = uo[1] * a.g[0] + uo[2] * a.g[1] + uo[3] * a.g[2];
                F -= S;
            }
            un[v] = uo[v] - a.wdt * F;
            Un[v * pl + o] = un[v];
        }
        double wn[NVP], wo[NVP];
        prim(un, wn, gm);
        prim(uo, wo, gm);
#pragma unroll
        for (int v = 0; v < NVP; v++) {
            if ((wn_vars >> v) & 1u) Wn[v * pl + o] = wn[v];
            Wo[v * pl + o] = wo[v];
        }
    }
}

// The update = 1 call on the FV state: a cell whose theta and face
// neighbours' theta are all zero keeps the update = 0 result already in
// U_next (its face fluxes did not change); the other cells are recomputed
// from U_fv with the blended fluxes. With gravity every cell is recomputed,
// the momentum and energy sources taken from the values just formed (the
// update = 1 form writes U_cv variable by variable and reads it back).
__global__ void __launch_bounds__(BS) k_fv_state_update_blended(double* __restrict__ Un, const double* __restrict__ Ucv,
                                                                const double* __restrict__ Fx,
                                                                const double* __restrict__ Fy,
                                                                const double* __restrict__ Fz,
                                                                const double* __restrict__ th, FV sc, FV sx, FV sy,
                                                                FV sz, Geo geo, FvUpd a, int fx, int fy, int fz, int Mx,
                                                                int My, uint32_t total) {
    uint32_t idx = xcd_block() * BS + threadIdx.x;
    if (idx >= total) return;
    const int I = fx + (int)(idx % Mx), J = fy + (int)((idx / Mx) % My), K = fz + (int)(idx / ((uint32_t)Mx * My));
    const size_t sy1 = (size_t)sc.Nx, sz1 = (size_t)sc.Nx * sc.Ny;
    const size_t o = ((size_t)K * sc.Ny + J) * sc.Nx + I, pl = sc.plane();
    bool touched = a.grav || th[o] != 0.0;
    if (a.act[0]) touched = touched || th[o - 1] != 0.0 || th[o + 1] != 0.0;
    if (a.act[1]) touched = touched || th[o - sy1] != 0.0 || th[o + sy1] != 0.0;
    if (a.act[2]) touched = touched || th[o - sz1] != 0.0 || th[o + sz1] != 0.0;
    if (!touched) return;
    double un[NVP];
#pragma unroll
    for (int v = 0; v < NVP; v++) {
        double h, F = 0;
        const double uo = Ucv[v * pl + o];
        if (a.act[0]) { h = geo.f[0][I + 1] - geo.f[0][I]; F += (Fx[sx.off(v, K, J, I + 1)] - Fx[sx.off(v, K, J, I)]) / h; }
        if (a.act[1]) { h = geo.f[1][J + 1] - geo.f[1][J]; F += (Fy[sy.off(v, K, J + 1, I)] - Fy[sy.off(v, K, J, I)]) / h; }
        if (a.act[2]) { h = geo.f[2][K + 1] - geo.f[2][K]; F += (Fz[sz.off(v, K + 1, J, I)] - Fz[sz.off(v, K, J, I)]) / h; }
        if (a.grav) {
            double S = 0.0;
            if (v == 1) S = un[0] * a.g[0];
            else if (v == 2) S = un[0] * a.g[1];
            else if (v == 3) S = un[0] * a.g[2];
            else if (v == 4) S = un[1] * a.g[0] + un[2] * a.g[1] + un[3] * a.g[2];
            F -= S;
        }
        un[v] = uo - a.wdt * F;
        Un[v * pl + o] = un[v];
    }
}

// FV state -> U_sp of every element (active elements from their FV cells,
// ghost elements by the sp -> cv -> sp round trip the SD-layout sequence
// gives them), and optionally W_cv = cv(prim(U_sp)) and the per-workgroup
// minimum of the next dt over the active points (src/hydro.cpp:108-132,
// src/hydro_ader.hpp:235-243, the k_dt_partial formula). Workgroups stride
// over the element groups (row g / gpr of all elements, x fastest).
struct DtArgs {
    double gm, cfl, h[3], nu;
    int act[3];
};
template <int Q, int D>
__global__ void __launch_bounds__(BS) k_fv_state_out(const double* __restrict__ Ufv, FV sc, double* __restrict__ Usp,
                                                     double* __restrict__ Wcv, SD s, const double* __restrict__ Mcs,
                                                     const double* __restrict__ Msc, int ldm, int gx, int gy, int gz,
                                                     int fx, int fy, int fz, int gpr, uint32_t groups, DtArgs dta,
                                                     double* __restrict__ part) {
    using TL = ElTile<Q, D>;
    constexpr int TPE = TL::TPE, EPB = TL::EPB;
    __shared__ double b0[EPB * TPE], b1[EPB * TPE], sI[EPB * TPE], sMc[Q * Q], sMs[Q * Q];
    __shared__ double red[BS];
    const int tid = threadIdx.x, el = tid / TPE, sp = tid % TPE;
    if (tid < Q * Q) {
        sMc[tid] = Mcs[(tid / Q) * ldm + tid % Q];
        sMs[tid] = Msc[(tid / Q) * ldm + tid % Q];
    }
    const bool want_w = Wcv != nullptr, want_dt = part != nullptr;
    const size_t slab = s.slab();
    double r = __builtin_inf();
    for (uint32_t g = blockIdx.x; g < groups; g += gridDim.x) {
        const int row = (int)(g / gpr), i0 = (int)(g % gpr) * EPB;
        const int j = row % s.Ny, k = row / s.Ny, i = i0 + el;
        const bool live = el < EPB && i < s.Nx;
        const bool act_row = j >= gy && j < s.Ny - gy && k >= gz && k < s.Nz - gz;
        const bool active = live && act_row && i >= gx && i < s.Nx - gx;
        const size_t eo = (((size_t)k * s.Ny + j) * s.Nx + i) * TPE + sp;
        double* e0 = b0 + (live ? el : 0) * TPE;
        double* e1 = b1 + (live ? el : 0) * TPE;
        double u[NVP];
        for (int v = 0; v < NVP; v++) {
            // the cell values: FV rows for active elements, the forward
            // transform of U_sp for ghost elements
            __syncthreads();
            if (act_row && tid < EPB * TPE) {
                int e, pt, ii, jj, kk;
                fv_slot<Q, D>(tid, e, pt, ii, jj, kk);
                const int ie = i0 + e;
                if (ie >= gx && ie < s.Nx - gx) {
                    const int I = fx + (ie - gx) * Q + ii;
                    const int J = D >= 2 ? fy + (j - gy) * Q + jj : fy + (j - gy);
                    const int K = D == 3 ? fz + (k - gz) * Q + kk : fz + (k - gz);
                    sI[e * TPE + pt] = Ufv[sc.off(v, K, J, I)];
                }
            }
            const double cg = el_transform<Q, D>(live && !active ? Usp[v * slab + eo] : 0.0, e0, e1, sMs, sp,
                                                 live && !active);
            const double cv = active ? sI[el * TPE + sp] : cg;
            u[v] = el_transform<Q, D>(cv, e0, e1, sMc, sp, live);
            if (live) Usp[v * slab + eo] = u[v];
        }
        if (!want_w && !want_dt) continue;
        double w[NVP];
        if (live) prim(u, w, dta.gm);
        else
            for (int v = 0; v < NVP; v++) w[v] = 0.0;
        double wc[NVP];
        for (int v = 0; v < NVP; v++) {
            wc[v] = el_transform<Q, D>(w[v], e0, e1, sMs, sp, live);
            if (live && want_w) Wcv[v * slab + eo] = wc[v];
        }
        if (active && want_dt) {
            const double c_s = sound_speed(wc[0], wc[4], dta.gm);
            double c_max = 0, dx_min = 1, dt_min;
            if (dta.act[0]) { c_max += (fabs(wc[1]) + c_s); dx_min = smin(dx_min, dta.h[0]); }
            if (dta.act[1]) { c_max += (fabs(wc[2]) + c_s); dx_min = smin(dx_min, dta.h[1]); }
            if (dta.act[2]) { c_max += (fabs(wc[3]) + c_s); dx_min = smin(dx_min, dta.h[2]); }
            dt_min = dta.cfl * dx_min / c_max / Q;
            if (dta.nu > 0.0) {
                const double dx_sub = dx_min / Q;
                const double dt_visc = 0.25 * dta.cfl * dx_sub * dx_sub / dta.nu;
                dt_min = nan_min(dt_min, dt_visc);
            }
            r = nan_min(r, dt_min);
        }
    }
    if (!want_dt) return;
    red[tid] = r;
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if (tid < w) red[tid] = nan_min(red[tid], red[tid + w]);
        __syncthreads();
    }
    if (tid == 0) part[blockIdx.x] = red[0];
}

// ---------------------------------------------------------------------------
// boundaries, FV cell-centred (src/boundary.cpp:123-190) along one axis.
// Reads come from active cells and writes go to ghost cells, so one pass is race-free.
__global__ void __launch_bounds__(BS) k_fv_bc(double* __restrict__ U, FV s, int dim, int type, uint32_t total) {
    uint32_t idx = blockIdx.x * BS + threadIdx.x;
    if (idx >= total) return;
    int Nx = dim == 0 ? 2 : s.Nx, Ny = dim == 1 ? 2 : s.Ny;
    int k, j, i;
    split3(idx, Ny, Nx, k, j, i);
    int l = dim == 0 ? i : (dim == 1 ? j : k);
    int N = dim == 0 ? s.Nx : (dim == 1 ? s.Ny : s.Nz);
    auto at = [&](int v, int ll) -> size_t {
        if (dim == 0) return s.off(v, k, j, ll);
        if (dim == 1) return s.off(v, k, ll, i);
        return s.off(v, ll, j, i);
    };
    for (int v = 0; v < s.nv; v++) {
        double BL, BR;
        if (type == SPDK_BC_PERIODIC) { BL = U[at(v, N - 4 + l)]; BR = U[at(v, 2 + l)]; }
        else if (type == SPDK_BC_GRADFREE) { BL = U[at(v, 2 + l)]; BR = U[at(v, N - 4 + l)]; }
        else {
            double sgn = (v == 1 + dim) ? -1.0 : 1.0;
            BL = sgn * U[at(v, 3 - l)];
            BR = sgn * U[at(v, N - 3 - l)];
        }
        U[at(v, l)] = BL;
        U[at(v, N - 2 + l)] = BR;
    }
}

// The same ghost layers of every active axis in one launch: each ghost cell
// (enumerated once: the z layers, then the y layers of the interior z
// planes, then the x layers of the interior rows) takes the interior cell
// that the x, then y, then z fills of k_fv_bc would leave in it, times the
// reflective sign of each axis it is a ghost of. Same bits: the values are
// copies with exact sign flips.
struct BcAxes {
    int type[3], act[3];
};
__device__ __forceinline__ int bc_src(int c, int n, int type, bool& flip) {
    flip = false;
    if (c >= 2 && c <= n - 3) return c;
    const bool lo = c < 2;
    const int l = lo ? c : c - (n - 2);  // layer 0 or 1 of its side
    if (type == SPDK_BC_PERIODIC) return lo ? n - 4 + l : 2 + l;
    if (type == SPDK_BC_GRADFREE) return lo ? 2 + l : n - 4 + l;
    flip = true;
    return lo ? 3 - l : n - 3 - l;
}
__global__ void __launch_bounds__(BS) k_fv_bc_all(double* __restrict__ U, FV s, BcAxes b, uint32_t nzl, uint32_t nyl,
                                                  uint32_t total) {
    uint32_t idx = blockIdx.x * BS + threadIdx.x;
    if (idx >= total) return;
    const int kz0 = b.act[2] ? 2 : 0, kz1 = b.act[2] ? s.Nz - 3 : 0;   // interior z planes
    const int jy0 = b.act[1] ? 2 : 0, jy1 = b.act[1] ? s.Ny - 3 : 0;   // interior y rows
    int k, j, i;
    if (idx < nzl) {  // z layers: 4 planes of Ny x Nx
        i = idx % s.Nx;
        j = (idx / s.Nx) % s.Ny;
        const int L = idx / (s.Nx * s.Ny);
        k = L < 2 ? L : s.Nz - 4 + L;
    } else if ((idx -= nzl) < nyl) {  // y layers of the interior planes: 4 rows of Nx
        i = idx % s.Nx;
        const int L = (idx / s.Nx) % 4;
        j = L < 2 ? L : s.Ny - 4 + L;
        k = kz0 + (int)(idx / (s.Nx * 4));
    } else {  // x layers of the interior rows
        idx -= nyl;
        const int L = idx % 4;
        i = L < 2 ? L : s.Nx - 4 + L;
        j = jy0 + (int)((idx / 4) % (jy1 - jy0 + 1));
        k = kz0 + (int)(idx / (4 * (jy1 - jy0 + 1)));
    }
    (void)kz1;
    bool fx = false, fy = false, fz = false;
    const int si = b.act[0] ? bc_src(i, s.Nx, b.type[0], fx) : i;
    const int sj = b.act[1] ? bc_src(j, s.Ny, b.type[1], fy) : j;
    const int sk = b.act[2] ? bc_src(k, s.Nz, b.type[2], fz) : k;
    for (int v = 0; v < s.nv; v++) {
        double u = U[s.off(v, sk, sj, si)];
        if (fx && v == 1) u = -1.0 * u;
        if (fy && v == 2) u = -1.0 * u;
        if (fz && v == 3) u = -1.0 * u;
        U[s.off(v, k, j, i)] = u;
    }
}

// compute_primitives, FV: src/hydro.cpp:134-153
__global__ void __launch_bounds__(BS) k_fv_prim(const double* __restrict__ U, double* __restrict__ W, FV s, double gm,
                                                uint32_t total) {
    uint32_t idx = blockIdx.x * BS + threadIdx.x;
    if (idx >= total) return;
    size_t pl = s.plane();
    double u[NVP], w[NVP];
#pragma unroll
    for (int v = 0; v < NVP; v++) u[v] = U[v * pl + idx];
    prim(u, w, gm);
#pragma unroll
    for (int v = 0; v < NVP; v++) W[v * pl + idx] = w[v];
    for (int v = NVP; v < s.nv; v++) W[v * pl + idx] = 0.0;
}

// ---------------------------------------------------------------------------
// NAD (src/trouble_detection.cpp:35-104) on cells 1..N-2 and SED
// (:107-148) on cells 2..N-3 of every active axis, one launch.
__device__ __forceinline__ double alpha_fn(double dv, double dUm, double dU, double dUp) {
    double vL = dU - dUm;
    double ap = smin(1.0, smax(vL, 0.0) / dv);
    double am = smin(1.0, smin(vL, 0.0) / dv);
    double aL = (dv < 0.0 ? am : ap);
    aL = (dv == 0.0 ? 1.0 : aL);
    double vR = dUp - dU;
    ap = smin(1.0, smax(vR, 0.0) / dv);
    am = smin(1.0, smin(vR, 0.0) / dv);
    double aR = (dv < 0.0 ? am : ap);
    aR = (dv == 0.0 ? 1.0 : aR);
    return smin(aL, aR);
}
struct Det {
    int act[3];
    int moore, delta, use_sed, pad, mask;
    double tol;
};
// The active axes are template arguments, so the Moore box and the SED
// stencils unroll with their neighbour offsets formed once per thread; the
// box is visited in the same (z, y, x) order as the reference's loops.
// BITS: instead of the alpha arrays, one 16-bit word per cell whose bit
// v * 3 + d says alpha_d(v) >= 1; relax_NAD only asks whether the smallest
// alpha around a cell is below 1, so the words give the same flags (for
// finite states) with 2 bytes per cell instead of 48. Cells outside 2..N-3
// are never written and read as 0, the reference's never-written alpha = 0.
template <int AX, int AY, int AZ, bool BITS = false>
__global__ void __launch_bounds__(BS) k_nad_sed(const double* __restrict__ Wn, const double* __restrict__ Wo,
                                                double* __restrict__ tr, double* __restrict__ alx, double* __restrict__ aly,
                                                double* __restrict__ alz, uint16_t* __restrict__ abits, FV s, Det a,
                                                Geo geo, uint32_t total) {
    uint32_t idx = xcd_block() * BS + threadIdx.x;
    if (idx >= total) return;
    int k, j, i;
    split3(idx, s.Ny, s.Nx, k, j, i);
    const int nvar = s.nv - 1;
    const bool in1 = (!AX || (i >= 1 && i <= s.Nx - 2)) && (!AY || (j >= 1 && j <= s.Ny - 2)) &&
                     (!AZ || (k >= 1 && k <= s.Nz - 2));
    const bool in2 = (!AX || (i >= 2 && i <= s.Nx - 3)) && (!AY || (j >= 2 && j <= s.Ny - 3)) &&
                     (!AZ || (k >= 2 && k <= s.Nz - 3));
    const size_t pl = s.plane(), c0 = ((size_t)k * s.Ny + j) * s.Nx + i;
    const long sy = s.Nx, sz = (long)s.Nx * s.Ny;
    uint32_t bits = 0;
    for (int v = 0; v < nvar; v++) {
        if (!((a.mask >> v) & 1)) continue;
        const double* wo = Wo + v * pl + c0;
        const double* wn = Wn + v * pl + c0;
        if (in1) {
            double un = wn[0];
            double mx = wo[0], mn = mx;
            if (a.moore) {
#pragma unroll
                for (int dk = -AZ; dk <= AZ; dk++)
#pragma unroll
                    for (int dj = -AY; dj <= AY; dj++)
#pragma unroll
                        for (int di = -AX; di <= AX; di++) {
                            double u = wo[dk * sz + dj * sy + di];
                            mx = smax(mx, u);
                            mn = smin(mn, u);
                        }
            } else {
                double uL, uR;
                if (AX) { uL = wo[-1]; uR = wo[1]; mx = smax(smax(uL, mx), uR); mn = smin(smin(uL, mn), uR); }
                if (AY) { uL = wo[-sy]; uR = wo[sy]; mx = smax(smax(uL, mx), uR); mn = smin(smin(uL, mn), uR); }
                if (AZ) { uL = wo[-sz]; uR = wo[sz]; mx = smax(smax(uL, mx), uR); mn = smin(smin(uL, mn), uR); }
            }
            if (a.delta) {
                double eps = a.tol * (mx - mn);
                mn -= eps;
                mx += eps;
            } else {
                mn -= fabs(mn) * a.tol;
                mx += fabs(mx) * a.tol;
            }
            tr[v * pl + c0] = (un > mx || un < mn) ? 1.0 : 0.0;
        }
        if (a.use_sed && in2) {
            double u = wn[0];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                if (!(d == 0 ? AX : (d == 1 ? AY : AZ))) continue;
                const double* c = geo.c[d];
                const double* f = geo.f[d];
                const int l = d == 0 ? i : (d == 1 ? j : k);
                const long st = d == 0 ? 1 : (d == 1 ? sy : sz);
                double h = (c[l + 1] - c[l - 1]);
                double hp = (c[l + 2] - c[l]);
                double hm = (c[l] - c[l - 2]);
                double du = (wn[st] - wn[-st]) / h;
                double dup = (wn[2 * st] - u) / hp;
                double dum = (u - wn[-2 * st]) / hm;
                double d2u = (dup - dum) / (c[l + 1] - c[l - 1]);
                double dv = 0.5 * d2u * (f[l + 1] - f[l]);
                const double al_v = alpha_fn(dv, dum, du, dup);
                if (BITS) {
                    if (al_v >= 1.0) bits |= 1u << (v * 3 + d);
                } else {
                    double* al = d == 0 ? alx : (d == 1 ? aly : alz);
                    al[v * pl + c0] = al_v;
                }
            }
        }
    }
    if (BITS && a.use_sed && in2) abits[c0] = (uint16_t)bits;
}

// The same detection in 3-D (Moore box, alpha bits), one thread per (x, y)
// column walking KC cells up in z: the box's three W_old planes and the SED
// z stencil of W_new slide along the column in registers, so each plane is
// loaded once per column instead of three (box) or five (SED) times, and
// the one-thread-per-cell form's plane-apart re-reads (L2 misses) go away.
// Cells 1..N-2 of every axis; every value is formed by the same expression
// as in k_nad_sed, and the box's max and min by plane (each plane reduced
// once, in (y, x) order, then the three planes): for finite values max and
// min are exact in any grouping, so the flags are the same; a non-finite
// state stops the run at the next dt anyway.
template <int KC>
__global__ void __launch_bounds__(BS) k_nad_sed_z(const double* __restrict__ Wn, const double* __restrict__ Wo,
                                                  double* __restrict__ tr, uint16_t* __restrict__ abits, FV s, Det a,
                                                  Geo geo, uint32_t total) {
    uint32_t idx = xcd_block() * BS + threadIdx.x;
    if (idx >= total) return;  // total = (Nx - 2) * (Ny - 2) * ceil((Nz - 2) / KC)
    const int mx1 = s.Nx - 2, my1 = s.Ny - 2;
    const int i = 1 + (int)(idx % mx1), j = 1 + (int)((idx / mx1) % my1);
    const int k0 = 1 + (int)(idx / ((uint32_t)mx1 * my1)) * KC;
    const int kn = min(KC, s.Nz - 1 - k0);  // cells k0 .. k0 + kn - 1 (<= Nz - 2)
    const size_t pl = s.plane();
    const long sy = s.Nx, sz = (long)s.Nx * s.Ny;
    const bool in2xy = i >= 2 && i <= s.Nx - 3 && j >= 2 && j <= s.Ny - 3;
    const bool sed = a.use_sed && in2xy;
    static_assert(KC <= 8, "alpha words of the column packed in two 64-bit registers");
    uint64_t bits[2] = {0, 0};  // 16-bit word of cell kk at bits 16 * (kk % 4) of bits[kk / 4]
    for (int v = 0; v < s.nv - 1; v++) {
        if (!((a.mask >> v) & 1)) continue;
        const double* wo = Wo + v * pl + ((size_t)k0 * s.Ny + j) * s.Nx + i;  // cell (k0, j, i)
        const double* wn = Wn + v * pl + ((size_t)k0 * s.Ny + j) * s.Nx + i;
        // max and min of the 3 x 3 W_old patch of planes k-1, k, k+1 (index dk + 1)
        double pmx[3], pmn[3];
        auto patch = [&](const double* q, double& hi, double& lo) {
            hi = lo = q[-sy - 1];
#pragma unroll
            for (int dj = -1; dj <= 1; dj++)
#pragma unroll
                for (int di = -1; di <= 1; di++) {
                    const double u = q[dj * sy + di];
                    hi = smax(hi, u);
                    lo = smin(lo, u);
                }
        };
        patch(wo - sz, pmx[1], pmn[1]);
        patch(wo, pmx[2], pmn[2]);
        double Z[5];  // W_new of the column at k-2 .. k+2 (planes below 0 are never used)
#pragma unroll
        for (int q = 0; q < 4; q++) Z[q + 1] = (k0 - 2 + q >= 0) ? wn[(q - 2) * sz] : 0.0;
#pragma unroll 1
        for (int kk = 0; kk < kn; kk++) {
            const int k = k0 + kk;
            const double* o = wo + (long)kk * sz;
            const double* n = wn + (long)kk * sz;
            // slide: the box gains plane k+1, the column gains k+2 (k + 2 <= Nz - 1 whenever SED is formed;
            // at k = Nz - 2 the plane k + 2 is outside the array and not loaded)
            pmx[0] = pmx[1]; pmn[0] = pmn[1];
            pmx[1] = pmx[2]; pmn[1] = pmn[2];
            patch(o + sz, pmx[2], pmn[2]);
#pragma unroll
            for (int q = 0; q < 4; q++) Z[q] = Z[q + 1];
            Z[4] = (k + 2 <= s.Nz - 1) ? n[2 * sz] : 0.0;
            // NAD (k_nad_sed, Moore box)
            {
                double un = Z[2];
                double mx = smax(smax(pmx[0], pmx[1]), pmx[2]), mn = smin(smin(pmn[0], pmn[1]), pmn[2]);
                if (a.delta) {
                    double eps = a.tol * (mx - mn);
                    mn -= eps;
                    mx += eps;
                } else {
                    mn -= fabs(mn) * a.tol;
                    mx += fabs(mx) * a.tol;
                }
                tr[v * pl + ((size_t)k * s.Ny + j) * s.Nx + i] = (un > mx || un < mn) ? 1.0 : 0.0;
            }
            if (sed && k >= 2 && k <= s.Nz - 3) {
                double u = Z[2];
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const double* c = geo.c[d];
                    const double* f = geo.f[d];
                    const int l = d == 0 ? i : (d == 1 ? j : k);
                    double wm2, wm1, wp1, wp2;
                    if (d == 2) {
                        wm2 = Z[0]; wm1 = Z[1]; wp1 = Z[3]; wp2 = Z[4];
                    } else {
                        const long st = d == 0 ? 1 : sy;
                        wm2 = n[-2 * st]; wm1 = n[-st]; wp1 = n[st]; wp2 = n[2 * st];
                    }
                    double h = (c[l + 1] - c[l - 1]);
                    double hp = (c[l + 2] - c[l]);
                    double hm = (c[l] - c[l - 2]);
                    double du = (wp1 - wm1) / h;
                    double dup = (wp2 - u) / hp;
                    double dum = (u - wm2) / hm;
                    double d2u = (dup - dum) / (c[l + 1] - c[l - 1]);
                    double dv = 0.5 * d2u * (f[l + 1] - f[l]);
                    if (alpha_fn(dv, dum, du, dup) >= 1.0) {
                        const uint64_t bit = 1ull << (16 * (kk & 3) + v * 3 + d);
                        if (kk < 4) bits[0] |= bit;
                        else bits[1] |= bit;
                    }
                }
            }
        }
    }
    if (sed)
#pragma unroll 1
        for (int kk = 0; kk < kn; kk++) {
            const int k = k0 + kk;
            const uint64_t w = (kk < 4 ? bits[0] : bits[1]) >> (16 * (kk & 3));
            if (k >= 2 && k <= s.Nz - 3) abits[((size_t)k * s.Ny + j) * s.Nx + i] = (uint16_t)(w & 0xFFFFu);
        }
}

// relax_NAD (src/trouble_detection.cpp:152-183) then PAD_criteria (:244-259), cells 1..N-2
template <int AX, int AY, int AZ, bool BITS = false>
__global__ void __launch_bounds__(BS) k_relax_pad(const double* __restrict__ Wn, double* __restrict__ tr,
                                                  const double* __restrict__ alx, const double* __restrict__ aly,
                                                  const double* __restrict__ alz, const uint16_t* __restrict__ abits,
                                                  FV s, Det a, uint32_t total) {
    uint32_t idx = xcd_block() * BS + threadIdx.x;
    if (idx >= total) return;
    int k, j, i;
    split3(idx, s.Ny - 2 * AY, s.Nx - 2 * AX, k, j, i);
    k += AZ; j += AY; i += AX;
    const int nvar = s.nv - 1;
    const size_t pl = s.plane(), c0 = ((size_t)k * s.Ny + j) * s.Nx + i;
    const long sy = s.Nx, sz = (long)s.Nx * s.Ny;
    // BITS: bit v*3+d of all_ge1 is set when alpha_d(v) >= 1 at the cell and
    // both of its neighbours along d (each axis' bits from its own neighbours)
    uint32_t all_ge1 = 0x7fffu;
    if (BITS && a.use_sed) {
        constexpr uint32_t kX = 0x1249u, kY = 0x2492u, kZ = 0x4924u;  // bits v*3+0, v*3+1, v*3+2
        const uint32_t m0 = abits[c0];
        uint32_t ax = kX, ay = kY, az = kZ;
        if (AX) ax = (uint32_t)abits[c0 - 1] & m0 & (uint32_t)abits[c0 + 1] & kX;
        if (AY) ay = (uint32_t)abits[c0 - sy] & m0 & (uint32_t)abits[c0 + sy] & kY;
        if (AZ) az = (uint32_t)abits[c0 - sz] & m0 & (uint32_t)abits[c0 + sz] & kZ;
        all_ge1 = ax | ay | az;
    }
    for (int v = 0; v < nvar; v++) {
        if (!((a.mask >> v) & 1)) continue;
        const size_t o = v * pl + c0;
        bool small;  // the smallest alpha around the cell is below 1
        if (BITS) {
            const uint32_t need = (AX ? 1u << (v * 3) : 0u) | (AY ? 1u << (v * 3 + 1) : 0u) | (AZ ? 1u << (v * 3 + 2) : 0u);
            small = (all_ge1 & need) != need;
        } else {
            double alpha = 1;
            if (AX) alpha = smin(smin(alx[o - 1], alx[o]), alx[o + 1]);
            if (AY) alpha = smin(alpha, smin(smin(aly[o - sy], aly[o]), aly[o + sy]));
            if (AZ) alpha = smin(alpha, smin(smin(alz[o - sz], alz[o]), alz[o + sz]));
            small = alpha < 1;
        }
        double t = tr[o];
        if (a.use_sed) tr[o] = small ? t : 0;
    }
    double t = 0;
    for (int v = 0; v < nvar; v++)
        if ((a.mask >> v) & 1) t = smax(t, tr[v * pl + c0]);
    tr[nvar * pl + c0] = t;
    if (a.pad) {
        double d = Wn[c0], p = Wn[4 * pl + c0];
        if (d < kRhoMin || d > kRhoMax) tr[nvar * pl + c0] = 1;
        if (p < kPMin || p > kPMax) tr[nvar * pl + c0] = 1;
    }
}

// apply_blending (src/trouble_detection.cpp:191-210), cells 1..N-2
template <int AX, int AY, int AZ>
__global__ void __launch_bounds__(BS) k_blend(const double* __restrict__ tr, double* __restrict__ th, FV s,
                                              uint32_t total) {
    uint32_t idx = xcd_block() * BS + threadIdx.x;
    if (idx >= total) return;
    int k, j, i;
    split3(idx, s.Ny - 2 * AY, s.Nx - 2 * AX, k, j, i);
    k += AZ; j += AY; i += AX;
    const double w[4] = {1.0, 0.75, 0.5, 0.375};
    const size_t c0 = ((size_t)k * s.Ny + j) * s.Nx + i;
    const long sy = s.Nx, sz = (long)s.Nx * s.Ny;
    const double* tv = tr + (size_t)(s.nv - 1) * s.plane() + c0;
    double t = 0;
#pragma unroll
    for (int dk = -AZ; dk <= AZ; dk++)
#pragma unroll
        for (int dj = -AY; dj <= AY; dj++)
#pragma unroll
            for (int di = -AX; di <= AX; di++) {
                const int n0 = (di != 0) + (dj != 0) + (dk != 0);
                t = smax(t, w[n0] * tv[dk * sz + dj * sy + di]);
            }
    th[c0] = t;
}
// blending_ring (src/trouble_detection.cpp:226-242), cells 1..N-2
template <int AX, int AY, int AZ>
__global__ void __launch_bounds__(BS) k_ring(const double* __restrict__ ti, double* __restrict__ to, FV s,
                                             uint32_t total) {
    uint32_t idx = xcd_block() * BS + threadIdx.x;
    if (idx >= total) return;
    int k, j, i;
    split3(idx, s.Ny - 2 * AY, s.Nx - 2 * AX, k, j, i);
    k += AZ; j += AY; i += AX;
    const size_t c0 = ((size_t)k * s.Ny + j) * s.Nx + i;
    const long sy = s.Nx, sz = (long)s.Nx * s.Ny;
    const double* tp = ti + c0;
    double t = tp[0], ring = 0;
#pragma unroll
    for (int dk = -AZ; dk <= AZ; dk++)
#pragma unroll
        for (int dj = -AY; dj <= AY; dj++)
#pragma unroll
            for (int di = -AX; di <= AX; di++)
                if (tp[dk * sz + dj * sy + di] > 0) ring = 0.25;
    to[c0] = smax(t, ring);
}
// apply_blending then blending_ring (3-D) in one pass on 16 x 4 x 4 cell
// tiles: the aggregate troubles of the tile and a two-cell halo are staged in
// LDS, theta_tmp is formed on the tile and a one-cell halo (0 outside cells
// 1..N-2, where apply_blending does not write it), then the ring. The
// troubles are 0 or 1 and every theta_tmp is >= +0, so the maxima do not
// depend on the visiting order; k_blend's order is kept anyway. theta_tmp is
// not written.
struct ThTile {
    static constexpr int TX = 16, TY = 4, TZ = 4;
    static constexpr int RX = TX + 4, RY = TY + 4, RZ = TZ + 4;  // troubles with the two-cell halo
    static constexpr int PX = TX + 2, PY = TY + 2, PZ = TZ + 2;  // theta_tmp with the one-cell halo
    static_assert(TX * TY * TZ == BS, "one cell per thread");
};
__global__ void __launch_bounds__(BS) k_blend_ring_tile(const double* __restrict__ tr, double* __restrict__ th, FV s,
                                                        int bx, int by, int bz) {
    using TT = ThTile;
    constexpr int NR = TT::RX * TT::RY * TT::RZ, NP = TT::PX * TT::PY * TT::PZ;
    __shared__ double st[NR];
    __shared__ double sp[NP];
    const uint32_t b = xcd_block();
    if (b >= (uint32_t)bx * by * bz) return;  // padding blocks of the XCD grid (uniform per block)
    const int bi = (int)(b % bx), bj = (int)((b / bx) % by), bk = (int)(b / ((uint32_t)bx * by));
    const int i0 = 1 + bi * TT::TX, j0 = 1 + bj * TT::TY, k0 = 1 + bk * TT::TZ;  // first output cell
    const int tid = threadIdx.x;
    const double* tv = tr + (size_t)(s.nv - 1) * s.plane();
    for (int q = tid; q < NR; q += BS) {
        const int gi = i0 - 2 + q % TT::RX, gj = j0 - 2 + (q / TT::RX) % TT::RY, gk = k0 - 2 + q / (TT::RX * TT::RY);
        if (gi >= 0 && gi < s.Nx && gj >= 0 && gj < s.Ny && gk >= 0 && gk < s.Nz)
            st[q] = tv[((size_t)gk * s.Ny + gj) * s.Nx + gi];
    }
    __syncthreads();
    const double w[4] = {1.0, 0.75, 0.5, 0.375};
    for (int q = tid; q < NP; q += BS) {
        const int x = q % TT::PX, y = (q / TT::PX) % TT::PY, z = q / (TT::PX * TT::PY);
        const int gi = i0 - 1 + x, gj = j0 - 1 + y, gk = k0 - 1 + z;
        double t = 0;
        if (gi >= 1 && gi <= s.Nx - 2 && gj >= 1 && gj <= s.Ny - 2 && gk >= 1 && gk <= s.Nz - 2) {
            const double* c = st + ((z + 1) * TT::RY + (y + 1)) * TT::RX + (x + 1);
#pragma unroll
            for (int dk = -1; dk <= 1; dk++)
#pragma unroll
                for (int dj = -1; dj <= 1; dj++)
#pragma unroll
                    for (int di = -1; di <= 1; di++) {
                        const int n0 = (di != 0) + (dj != 0) + (dk != 0);
                        t = smax(t, w[n0] * c[(dk * TT::RY + dj) * TT::RX + di]);
                    }
        }
        sp[q] = t;
    }
    __syncthreads();
    const int tx = tid % TT::TX, ty = (tid / TT::TX) % TT::TY, tz = tid / (TT::TX * TT::TY);
    const int i = i0 + tx, j = j0 + ty, k = k0 + tz;
    if (i > s.Nx - 2 || j > s.Ny - 2 || k > s.Nz - 2) return;
    const double* p = sp + ((tz + 1) * TT::PY + (ty + 1)) * TT::PX + (tx + 1);
    double ring = 0;
#pragma unroll
    for (int dk = -1; dk <= 1; dk++)
#pragma unroll
        for (int dj = -1; dj <= 1; dj++)
#pragma unroll
            for (int di = -1; di <= 1; di++)
                if (p[(dk * TT::PY + dj) * TT::PX + di] > 0) ring = 0.25;
    th[((size_t)k * s.Ny + j) * s.Nx + i] = smax(p[0], ring);
}
// theta_from_troubles (src/trouble_detection.cpp:213-221)
__global__ void __launch_bounds__(BS) k_theta_raw(const double* __restrict__ tr, double* __restrict__ th, FV s, uint32_t total) {
    uint32_t idx = blockIdx.x * BS + threadIdx.x;
    if (idx >= total) return;
    th[idx] = tr[(size_t)(s.nv - 1) * s.plane() + idx];
}

// ---------------------------------------------------------------------------
// MUSCL-Hancock + HLLC fallback flux: src/muscl.hpp:2-13,
// src/hydro.cpp:591-788
__device__ __forceinline__ double minmod(double SL, double SR, double xL, double xR) {
    double ratio = SR / SL;
    ratio = smax(0.0, smin(ratio, 1.0));
    double slope = ratio * SL;
    return 0.5 * slope * (xR - xL);
}
// src/hydro.cpp:591-615
__device__ __forceinline__ void corrector(const double* W, double* dWt, const double* dWx, const double* dWy,
                                          const double* dWz, double gm) {
    dWt[0] = 0;
    dWt[4] = 0;
    dWt[0] -= W[1] * dWx[0] + W[0] * dWx[1];
    dWt[1] = -W[1] * dWx[1] - dWx[4] / W[0];
    dWt[4] -= W[1] * dWx[4] + dWx[1] * gm * W[4];
    dWt[1] -= W[2] * dWy[1];
    dWt[1] -= W[3] * dWz[1];
    dWt[0] -= W[2] * dWy[0] + W[0] * dWy[2];
    dWt[2] = -W[2] * dWy[2] - dWy[4] / W[0];
    dWt[4] -= W[2] * dWy[4] + dWy[2] * gm * W[4];
    dWt[2] -= W[1] * dWx[2];
    dWt[2] -= W[3] * dWz[2];
    dWt[0] -= W[3] * dWz[0] + W[0] * dWz[3];
    dWt[3] = -W[3] * dWz[3] - dWz[4] / W[0];
    dWt[4] -= W[3] * dWz[4] + dWz[3] * gm * W[4];
    dWt[3] -= W[1] * dWx[3];
    dWt[3] -= W[2] * dWy[3];
}
// slopes_d (src/hydro.cpp:622-681), split in two: the cell's reconstruction
// (its primitive state, minmod slopes along the active axes and the Hancock
// time derivative), which is the same for every face of the cell, and the
// state at one of its faces along D. A cell's reconstruction is computed once
// per thread and used for every face that needs it.
struct Recon {
    double w[NVP], dw[3][NVP], dwt[NVP];
};
__device__ __forceinline__ void recon(const double* __restrict__ W, const FV& s, const Geo& g, int k, int j, int i,
                                      int ay, int az, double gm, Recon& r) {
    const double *xc = g.c[0], *xf = g.f[0], *yc = g.c[1], *yf = g.f[1], *zc = g.c[2], *zf = g.f[2];
#pragma unroll
    for (int v = 0; v < NVP; v++) {
        const double w = W[s.off(v, k, j, i)];
        r.w[v] = w;
        double wL = W[s.off(v, k, j, i - 1)], wR = W[s.off(v, k, j, i + 1)];
        r.dw[0][v] = minmod((wR - w) / (xc[i + 1] - xc[i]), (w - wL) / (xc[i] - xc[i - 1]), xf[i], xf[i + 1]);
        r.dw[1][v] = 0;
        r.dw[2][v] = 0;
        if (ay) {
            wL = W[s.off(v, k, j - 1, i)];
            wR = W[s.off(v, k, j + 1, i)];
            r.dw[1][v] = minmod((wR - w) / (yc[j + 1] - yc[j]), (w - wL) / (yc[j] - yc[j - 1]), yf[j], yf[j + 1]);
        }
        if (az) {
            wL = W[s.off(v, k - 1, j, i)];
            wR = W[s.off(v, k + 1, j, i)];
            r.dw[2][v] = minmod((wR - w) / (zc[k + 1] - zc[k]), (w - wL) / (zc[k] - zc[k - 1]), zf[k], zf[k + 1]);
        }
    }
    corrector(r.w, r.dwt, r.dw[0], r.dw[1], r.dw[2], gm);
}
// The reconstructed state at the cell's upper (WL of the reference) or lower
// (WR) face along D.
template <int D>
__device__ __forceinline__ void face_state(const Recon& r, const Geo& g, int k, int j, int i, double dt, bool upper,
                                           double* out) {
    const double h = (D == 0 ? g.f[0][i + 1] - g.f[0][i] : (D == 1 ? g.f[1][j + 1] - g.f[1][j] : g.f[2][k + 1] - g.f[2][k]));
#pragma unroll
    for (int v = 0; v < NVP; v++) {
        const double dw = r.dw[D][v];
        out[v] = upper ? r.w[v] + dw + r.dwt[v] * dt / h : r.w[v] - dw + r.dwt[v] * dt / h;
    }
}
// compute_fluxes<D> (src/hydro.cpp:683-751) at the lower D face of cell
// (k,j,i), given that cell's lower-face state wR: the lower neighbour's
// reconstruction, HLLC, and the theta blend of F.
template <int D>
__device__ __forceinline__ void fb_face(const double* __restrict__ W, double* __restrict__ F, const FV& sF, const FV& s,
                                        const Geo& g, int k, int j, int i, double t, const double* wR, double dt, int ay,
                                        int az, double gm) {
    const int kn = k - (D == 2), jn = j - (D == 1), in = i - (D == 0);
    double wL[NVP], uL[NVP], uR[NVP], fL[NVP];
    {
        Recon rn;
        recon(W, s, g, kn, jn, in, ay, az, gm, rn);
        face_state<D>(rn, g, kn, jn, in, dt, true, wL);
    }
    cons(wL, uL, gm);
    cons(wR, uR, gm);
    riemann_hllc<D>(uL, uR, fL, gm);
#pragma unroll
    for (int v = 0; v < NVP; v++) {
        size_t o = sF.off(v, k, j, i);
        double f = F[o];
        f = f + t * (fL[v] - f);
        F[o] = f;
    }
    for (int v = NVP; v < sF.nv; v++) {  // slot: HLLC slot taken as zero
        size_t o = sF.off(v, k, j, i);
        double f = F[o];
        f = f + t * (0.0 - f);
        F[o] = f;
    }
}
// One thread per cell, its three lower faces. Blend weight 0: F + 0 * (F_HLLC
// - F) is F (up to the sign of a zero flux) for any finite HLLC flux, so such
// a face keeps F and its MUSCL-HLLC evaluation is skipped. Troubled cells and
// their ring are a small fraction of the mesh, and cluster, so most waves skip
// as a whole.
__global__ void __launch_bounds__(BS) k_fallback(const double* __restrict__ W, const double* __restrict__ th,
                                                 double* __restrict__ Fx, double* __restrict__ Fy, double* __restrict__ Fz,
                                                 FV s, FV sx, FV sy, FV sz, Geo g, double dt, int ax, int ay, int az,
                                                 double gm, uint32_t total) {
    uint32_t idx = xcd_block() * BS + threadIdx.x;
    if (idx >= total) return;
    int Mx = s.Nx - 4 * ax + ax, My = s.Ny - 4 * ay + ay;
    int k, j, i;
    split3(idx, My, Mx, k, j, i);
    k += 2 * az; j += 2 * ay; i += 2 * ax;
    const size_t c0 = ((size_t)k * s.Ny + j) * s.Nx + i;
    const double tc = th[c0];
    const double t0 = smax(tc, th[c0 - 1]);
    const double t1 = ay ? smax(tc, th[c0 - s.Nx]) : 0.0;
    const double t2 = az ? smax(tc, th[c0 - (size_t)s.Nx * s.Ny]) : 0.0;
    if (t0 == 0.0 && t1 == 0.0 && t2 == 0.0) return;
    double wR[3][NVP];  // this cell's lower-face states
    {
        Recon rc;
        recon(W, s, g, k, j, i, ay, az, gm, rc);
        if (t0 != 0.0) face_state<0>(rc, g, k, j, i, dt, false, wR[0]);
        if (t1 != 0.0) face_state<1>(rc, g, k, j, i, dt, false, wR[1]);
        if (t2 != 0.0) face_state<2>(rc, g, k, j, i, dt, false, wR[2]);
    }
    if (t0 != 0.0) fb_face<0>(W, Fx, sx, s, g, k, j, i, t0, wR[0], dt, ay, az, gm);
    if (t1 != 0.0) fb_face<1>(W, Fy, sy, s, g, k, j, i, t1, wR[1], dt, ay, az, gm);
    if (t2 != 0.0) fb_face<2>(W, Fz, sz, s, g, k, j, i, t2, wR[2], dt, ay, az, gm);
}

// The same fallback on tiles of TX x TY x TZ cells (x active; AY, AZ the other
// active axes): every cell that has a blended face, and the cells of the
// low-side halo layer whose upper face a tile cell needs, is reconstructed
// once; the upper-face states go through LDS to the cell above, so each
// face's two states come from one reconstruction per cell instead of one per
// face (the per-cell kernel above reconstructs a cell up to four times).
// The arithmetic of each state and flux is the per-cell kernel's, so the
// faces are the same bits. A tile with no nonzero theta in reach exits.
template <int AY, int AZ>
struct FbTile {
    static constexpr int TX = AZ ? SPDK_FB_TX : (AY ? 16 : 256), TY = AZ ? BS / (SPDK_FB_TX * SPDK_FB_TZ) : (AY ? 16 : 1),
                         TZ = AZ ? SPDK_FB_TZ : 1;
    static constexpr int NC = TX * TY * TZ;          // = BS
    static constexpr int HX = TY * TZ, HY = AY ? TX * TZ : 0, HZ = AZ ? TX * TY : 0;  // halo cells per layer
    static_assert(NC == BS, "one cell per thread");
    static_assert(HX + HY + HZ <= BS, "one halo cell per thread");
};
template <int AY, int AZ>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(SPDK_FB_WAVES))) k_fallback_tile(const double* __restrict__ W, const double* __restrict__ th,
                                                      double* __restrict__ Fx, double* __restrict__ Fy,
                                                      double* __restrict__ Fz, FV s, FV fx, FV fy, FV fz, Geo g,
                                                      double dt, double gm, int bx, int by, int bz) {
    using TL = FbTile<AY, AZ>;
    constexpr int TX = TL::TX, TY = TL::TY, TZ = TL::TZ;
    __shared__ double up[3][NVP][TL::NC];                    // upper-face states of the tile cells, per axis
    __shared__ double uph[NVP][TL::HX + TL::HY + TL::HZ];    // upper-face state of each halo cell (its one axis)
    const int tid = threadIdx.x;
    const uint32_t b = xcd_block();
    if (b >= (uint32_t)bx * by * bz) return;  // padding blocks of the XCD grid (uniform per block)
    const int bi = b % bx, bj = (b / bx) % by, bk = b / (bx * by);
    const int tx = tid % TX, ty = (tid / TX) % TY, tz = tid / (TX * TY);
    const int i0 = 2 + bi * TX, j0 = AY ? 2 + bj * TY : 0, k0 = AZ ? 2 + bk * TZ : 0;
    const int i = i0 + tx, j = j0 + ty, k = k0 + tz;
    // processed cells: 2..N-2 along each active axis
    const bool valid = i <= s.Nx - 2 && (!AY || j <= s.Ny - 2) && (!AZ || k <= s.Nz - 2);
    const long sy = s.Nx, sz = (long)s.Nx * s.Ny;
    const size_t c0 = ((size_t)k * s.Ny + j) * s.Nx + i;
    double tc = 0, tlo[3] = {0, 0, 0}, tf[3] = {0, 0, 0};
    if (valid) {
        tc = th[c0];
        tlo[0] = th[c0 - 1];
        if (AY) tlo[1] = th[c0 - sy];
        if (AZ) tlo[2] = th[c0 - sz];
        tf[0] = smax(tc, tlo[0]);
        if (AY) tf[1] = smax(tc, tlo[1]);
        if (AZ) tf[2] = smax(tc, tlo[2]);
    }
    const bool lower = tf[0] != 0.0 || tf[1] != 0.0 || tf[2] != 0.0;
    // the upper face along d feeds the tile cell above if that one is processed and blends it
    bool upper = false;
    double thi[3] = {0, 0, 0};
    if (valid) {
        if (tx + 1 < TX && i + 1 <= s.Nx - 2) { thi[0] = th[c0 + 1]; upper = upper || smax(thi[0], tc) != 0.0; }
        if (AY && ty + 1 < TY && j + 1 <= s.Ny - 2) { thi[1] = th[c0 + sy]; upper = upper || smax(thi[1], tc) != 0.0; }
        if (AZ && tz + 1 < TZ && k + 1 <= s.Nz - 2) { thi[2] = th[c0 + sz]; upper = upper || smax(thi[2], tc) != 0.0; }
    }
    if (!__syncthreads_or(lower)) return;  // no blended face in the tile
    // halo cell of this thread: layer x (i0-1, j0+.., k0+..), then y, then z
    int hd = -1, hi = 0, hj = 0, hk = 0, hslot = tid;
    if (tid < TL::HX) {
        hd = 0; hi = i0 - 1; hj = j0 + tid % TY; hk = k0 + tid / TY;
    } else if (AY && tid < TL::HX + TL::HY) {
        const int r = tid - TL::HX;
        hd = 1; hi = i0 + r % TX; hj = j0 - 1; hk = k0 + r / TX;
    } else if (AZ && tid < TL::HX + TL::HY + TL::HZ) {
        const int r = tid - TL::HX - TL::HY;
        hd = 2; hi = i0 + r % TX; hj = j0 + r / TX; hk = k0 - 1;
    }
    bool hneed = false;
    if (hd >= 0) {  // the tile cell above the halo cell is processed and blends that face
        const int ui = hi + (hd == 0), uj = hj + (hd == 1), uk = hk + (hd == 2);
        if (ui <= s.Nx - 2 && (!AY || uj <= s.Ny - 2) && (!AZ || uk <= s.Nz - 2)) {
            const size_t cu = ((size_t)uk * s.Ny + uj) * s.Nx + ui;
            const size_t ch = ((size_t)hk * s.Ny + hj) * s.Nx + hi;
            hneed = smax(th[cu], th[ch]) != 0.0;
        }
    }
    double wR[3][NVP];  // this cell's lower-face states
    if (valid && (lower || upper)) {
        Recon rc;
        recon(W, s, g, k, j, i, AY, AZ, gm, rc);
        if (tf[0] != 0.0) face_state<0>(rc, g, k, j, i, dt, false, wR[0]);
        if (AY && tf[1] != 0.0) face_state<1>(rc, g, k, j, i, dt, false, wR[1]);
        if (AZ && tf[2] != 0.0) face_state<2>(rc, g, k, j, i, dt, false, wR[2]);
        if (upper) {
            double wu[NVP];
            face_state<0>(rc, g, k, j, i, dt, true, wu);
#pragma unroll
            for (int v = 0; v < NVP; v++) up[0][v][tid] = wu[v];
            if (AY) {
                face_state<1>(rc, g, k, j, i, dt, true, wu);
#pragma unroll
                for (int v = 0; v < NVP; v++) up[1][v][tid] = wu[v];
            }
            if (AZ) {
                face_state<2>(rc, g, k, j, i, dt, true, wu);
#pragma unroll
                for (int v = 0; v < NVP; v++) up[2][v][tid] = wu[v];
            }
        }
    }
    if (hneed) {
        Recon rh;
        recon(W, s, g, hk, hj, hi, AY, AZ, gm, rh);
        double wu[NVP];
        if (hd == 0) face_state<0>(rh, g, hk, hj, hi, dt, true, wu);
        else if (hd == 1) face_state<1>(rh, g, hk, hj, hi, dt, true, wu);
        else face_state<2>(rh, g, hk, hj, hi, dt, true, wu);
#pragma unroll
        for (int v = 0; v < NVP; v++) uph[v][hslot] = wu[v];
    }
    __syncthreads();
    if (!valid) return;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if ((d == 1 && !AY) || (d == 2 && !AZ) || tf[d] == 0.0) continue;
        const int td = d == 0 ? tx : (d == 1 ? ty : tz);
        const int stride = d == 0 ? 1 : (d == 1 ? TX : TX * TY);
        const int hs = d == 0 ? ty + TY * tz : (d == 1 ? TL::HX + tx + TX * tz : TL::HX + TL::HY + tx + TX * ty);
        double wL[NVP], uL[NVP], uR[NVP], fL[NVP];
#pragma unroll
        for (int v = 0; v < NVP; v++) wL[v] = td > 0 ? up[d][v][tid - stride] : uph[v][hs];
        cons(wL, uL, gm);
        cons(wR[d], uR, gm);
        if (d == 0) riemann_hllc<0>(uL, uR, fL, gm);
        else if (d == 1) riemann_hllc<1>(uL, uR, fL, gm);
        else riemann_hllc<2>(uL, uR, fL, gm);
        double* F = d == 0 ? Fx : (d == 1 ? Fy : Fz);
        const FV& sF = d == 0 ? fx : (d == 1 ? fy : fz);
#pragma unroll
        for (int v = 0; v < NVP; v++) {
            size_t o = sF.off(v, k, j, i);
            double f = F[o];
            f = f + tf[d] * (fL[v] - f);
            F[o] = f;
        }
        for (int v = NVP; v < sF.nv; v++) {  // slot: HLLC slot taken as zero
            size_t o = sF.off(v, k, j, i);
            double f = F[o];
            f = f + tf[d] * (0.0 - f);
            F[o] = f;
        }
    }
}

// ===========================================================================
static Geo geo_of(spdk_ctx* c) {
    Geo g;
    for (int d = 0; d < 3; d++) {
        g.f[d] = c->fvf[d] ? c->fvf[d] : c->fvf[0];
        g.c[d] = c->fvc[d] ? c->fvc[d] : c->fvc[0];
    }
    return g;
}
static FV fv_stag(const spdk_fv_shape& s, int d, int active) {
    FV f(s);
    if (active) {
        if (d == 0) f.Nx += 1;
        if (d == 1) f.Ny += 1;
        if (d == 2) f.Nz += 1;
    }
    return f;
}

extern "C" {

int spdk_fv_face_integral(spdk_ctx* c, const double* Ffp, spdk_sd_shape s_fp, double* F, spdk_fv_shape sf, double* T,
                          int t_id, int dim) {
    int dim1 = dim == 0 ? 1 : (dim == 1 ? 2 : 0), dim2 = dim1 == 0 ? 1 : (dim1 == 1 ? 2 : 0);
    const int* act = c->prm.active;
    int gx = c->ngh[0], gy = c->ngh[1], gz = c->ngh[2];
    int px = s_fp.nx - (dim == 0), py = s_fp.ny - (dim == 1), pz = s_fp.nz - (dim == 2);
    int Nx = s_fp.Nx + (dim == 0), Ny = s_fp.Ny + (dim == 1), Nz = s_fp.Nz + (dim == 2);
    uint32_t total = (uint32_t)((size_t)(Nz - 2 * gz) * (Ny - 2 * gy) * (Nx - 2 * gx) * pz * py * px);
    const double* S = c->mat[SPDK_M_SP_TO_CV];
    int lds = c->mcols[SPDK_M_SP_TO_CV];
    if (!act[dim1] || !act[dim2]) {
        hipLaunchKernelGGL(k_face_int_ref, dim3(nblocks(total, BS)), dim3(BS), 0, c->stream, Ffp, SD(s_fp), F, FV(sf), S, lds,
                           t_id, dim, dim1, dim2, act[dim1], act[dim2], gx, gy, gz, c->fgh[0], c->fgh[1], c->fgh[2], total);
        SPDK_LAUNCH_CHECK("face_integral_ref");
        return SPDK_OK;
    }
    (void)T;  // the two sweeps run fused; no intermediate slice
    int q1 = dim1 == 0 ? s_fp.nx : (dim1 == 1 ? s_fp.ny : s_fp.nz), q2 = dim2 == 0 ? s_fp.nx : (dim2 == 1 ? s_fp.ny : s_fp.nz);
    if (q1 != q2 || q1 != lds) { set_error("face_integral: transverse point counts"); return SPDK_E_SHAPE; }
    if (t_id < 0 || t_id >= s_fp.nt) { set_error("face_integral: slice"); return SPDK_E_ARG; }
    if (s_fp.nt < 1 || px != q1 || py != q1 || pz != q1) { set_error("face_integral: face counts"); return SPDK_E_SHAPE; }
    const int ex = Nx - 2 * gx, ey = Ny - 2 * gy;
    const uint32_t n_el = (uint32_t)((size_t)ex * ey * (Nz - 2 * gz));
#define SPDK_FACE_TILE(QQ)                                                                                          \
    hipLaunchKernelGGL(k_face_int_tile<QQ>, dim3((n_el + FaceTile<QQ>::EX - 1) / FaceTile<QQ>::EX), dim3(BS), 0,       \
                       c->stream, Ffp, SD(s_fp), F, FV(sf), S, lds, t_id, dim, gx, gy, gz, c->fgh[0], c->fgh[1],     \
                       c->fgh[2], ex, ey, n_el)
    if (n_el == 0) return SPDK_OK;
    switch (q1) {
        case 2: SPDK_FACE_TILE(2); break;
        case 3: SPDK_FACE_TILE(3); break;
        case 4: SPDK_FACE_TILE(4); break;
        case 5: SPDK_FACE_TILE(5); break;
        case 6: SPDK_FACE_TILE(6); break;
        case 7: SPDK_FACE_TILE(7); break;
        case 8: SPDK_FACE_TILE(8); break;
        case 9: SPDK_FACE_TILE(9); break;
        default: set_error("face_integral: degree"); return SPDK_E_SHAPE;
    }
#undef SPDK_FACE_TILE
    SPDK_LAUNCH_CHECK("face_integral");
    return SPDK_OK;
}

// Shape agreement of an SD cell-value array and the FV arrays of one update
// (state, flux faces derived from it, theta): the same variable count (6),
// and along every axis the FV cells are the active elements' solution
// points plus the FV ghost layers; an inactive axis is one element, one
// point and one cell.
static int fv_shapes_ok(const spdk_ctx* c, const spdk_sd_shape& s_cv, const spdk_fv_shape& sc, const char* who) {
    if (s_cv.nt < 1 || s_cv.nv != NV || sc.nv != NV) {
        set_error(std::string(who) + ": SD and FV arrays need 6 variables");
        return SPDK_E_SHAPE;
    }
    const int Ne[3] = {s_cv.Nx, s_cv.Ny, s_cv.Nz}, np[3] = {s_cv.nx, s_cv.ny, s_cv.nz}, Nf[3] = {sc.Nx, sc.Ny, sc.Nz};
    for (int d = 0; d < 3; d++) {
        const bool act = c->prm.active[d] != 0;
        const bool ok = act ? (np[d] == c->n && Ne[d] > 2 * c->ngh[d] &&
                               Nf[d] == (Ne[d] - 2 * c->ngh[d]) * c->n + 2 * c->fgh[d])
                            : (np[d] == 1 && Ne[d] == 1 && Nf[d] == 1);
        if (!ok) {
            set_error(std::string(who) + ": SD cell-value and FV shapes disagree along axis " + std::to_string(d));
            return SPDK_E_SHAPE;
        }
    }
    return SPDK_OK;
}

int spdk_fv_update(spdk_ctx* c, double* Un, double* Uo, double* Ucv, spdk_sd_shape s_cv, const double* Fx, const double* Fy,
                   const double* Fz, spdk_fv_shape sc, int t_id, double dt, int update) {
    if (!Un || !Uo || !Ucv || !Fx) { set_error("fv_update: null argument"); return SPDK_E_ARG; }
    if (int rc = fv_shapes_ok(c, s_cv, sc, "fv_update")) return rc;
    FvUpd a;
    for (int d = 0; d < 3; d++) { a.act[d] = c->prm.active[d]; a.g[d] = c->prm.g[d]; }
    a.grav = (a.g[0] != 0.0 || a.g[1] != 0.0 || a.g[2] != 0.0);
    if (t_id < 0 || t_id >= c->T) { set_error("fv_update: slice"); return SPDK_E_ARG; }
    a.wdt = c->wt_host[t_id] * dt;
    a.update = update;
    int gx = c->ngh[0], gy = c->ngh[1], gz = c->ngh[2];
    uint32_t total = (uint32_t)((size_t)(s_cv.Nz - 2 * gz) * (s_cv.Ny - 2 * gy) * (s_cv.Nx - 2 * gx) * s_cv.nz * s_cv.ny * s_cv.nx);
    hipLaunchKernelGGL(k_fv_update, dim3(xcd_grid(total)), dim3(BS), 0, c->stream, Un, Uo, Ucv, SD(s_cv), Fx, Fy ? Fy : Fx,
                       Fz ? Fz : Fx, FV(sc), fv_stag(sc, 0, a.act[0]), fv_stag(sc, 1, a.act[1]), fv_stag(sc, 2, a.act[2]),
                       geo_of(c), a, gx, gy, gz, c->fgh[0], c->fgh[1], c->fgh[2], total);
    SPDK_LAUNCH_CHECK("fv_update");
    return SPDK_OK;
}

int spdk_fv_update_prims(spdk_ctx* c, double* Un, double* Wn, double* Wo, const double* Ucv, spdk_sd_shape s_cv,
                         const double* Fx, const double* Fy, const double* Fz, spdk_fv_shape sc, int t_id, double dt) {
    if (!Un || !Wn || !Wo || !Ucv || !Fx) { set_error("fv_update_prims: null argument"); return SPDK_E_ARG; }
    if (int rc = fv_shapes_ok(c, s_cv, sc, "fv_update_prims")) return rc;
    FvUpd a;
    for (int d = 0; d < 3; d++) { a.act[d] = c->prm.active[d]; a.g[d] = c->prm.g[d]; }
    a.grav = (a.g[0] != 0.0 || a.g[1] != 0.0 || a.g[2] != 0.0);
    if (t_id < 0 || t_id >= c->T) { set_error("fv_update_prims: slice"); return SPDK_E_ARG; }
    a.wdt = c->wt_host[t_id] * dt;
    a.update = 0;
    int gx = c->ngh[0], gy = c->ngh[1], gz = c->ngh[2];
    uint32_t total = (uint32_t)((size_t)(s_cv.Nz - 2 * gz) * (s_cv.Ny - 2 * gy) * (s_cv.Nx - 2 * gx) * s_cv.nz * s_cv.ny * s_cv.nx);
    hipLaunchKernelGGL(k_fv_update_prims, dim3(xcd_grid(total)), dim3(BS), 0, c->stream, Un, Wn, Wo, Ucv, SD(s_cv), Fx,
                       Fy ? Fy : Fx, Fz ? Fz : Fx, FV(sc), fv_stag(sc, 0, a.act[0]), fv_stag(sc, 1, a.act[1]),
                       fv_stag(sc, 2, a.act[2]), geo_of(c), a, c->prm.gamma, gx, gy, gz, c->fgh[0], c->fgh[1], c->fgh[2],
                       total);
    SPDK_LAUNCH_CHECK("fv_update_prims");
    return SPDK_OK;
}

int spdk_fv_update_blended(spdk_ctx* c, double* Un, double* Ucv, spdk_sd_shape s_cv, const double* Fx, const double* Fy,
                           const double* Fz, const double* theta, spdk_fv_shape sc, int t_id, double dt) {
    // theta has the FV cell shape sc (one variable is read)
    if (!theta || !Un || !Ucv || !Fx) { set_error("fv_update_blended: null argument"); return SPDK_E_ARG; }
    if (int rc = fv_shapes_ok(c, s_cv, sc, "fv_update_blended")) return rc;
    FvUpd a;
    for (int d = 0; d < 3; d++) { a.act[d] = c->prm.active[d]; a.g[d] = c->prm.g[d]; }
    a.grav = (a.g[0] != 0.0 || a.g[1] != 0.0 || a.g[2] != 0.0);
    if (t_id < 0 || t_id >= c->T) { set_error("fv_update_blended: slice"); return SPDK_E_ARG; }
    a.wdt = c->wt_host[t_id] * dt;
    a.update = 1;
    int gx = c->ngh[0], gy = c->ngh[1], gz = c->ngh[2];
    uint32_t total = (uint32_t)((size_t)(s_cv.Nz - 2 * gz) * (s_cv.Ny - 2 * gy) * (s_cv.Nx - 2 * gx) * s_cv.nz * s_cv.ny * s_cv.nx);
    hipLaunchKernelGGL(k_fv_update_blended, dim3(xcd_grid(total)), dim3(BS), 0, c->stream, Un, Ucv, SD(s_cv), Fx,
                       Fy ? Fy : Fx, Fz ? Fz : Fx, theta, FV(sc), fv_stag(sc, 0, a.act[0]), fv_stag(sc, 1, a.act[1]),
                       fv_stag(sc, 2, a.act[2]), geo_of(c), a, gx, gy, gz, c->fgh[0], c->fgh[1], c->fgh[2], total);
    SPDK_LAUNCH_CHECK("fv_update_blended");
    return SPDK_OK;
}

}  // extern "C"

// FV-state arrays have the context's FV cell shape (6 variables, its cell
// counts with ghost layers; one cell along an inactive axis).
static int fv_state_shape_ok(const spdk_ctx* c, const spdk_fv_shape& sc, const char* who) {
    const int Nf[3] = {sc.Nx, sc.Ny, sc.Nz};
    bool ok = sc.nv == NV;
    for (int d = 0; d < 3; d++)
        ok = ok && (c->prm.active[d] ? (Nf[d] == c->fvn[d] && Nf[d] > 2 * c->fgh[d]) : Nf[d] == 1);
    if (!ok) set_error(std::string(who) + ": FV state shape differs from the context's FV cells");
    return ok ? SPDK_OK : SPDK_E_SHAPE;
}
// The element-tile transforms: active axes x, x-y or x-y-z, Q points per
// axis (Q <= 6), square Q x Q cv/sp matrices. D = number of active axes.
static int fv_state_tile_ok(const spdk_ctx* c, const spdk_sd_shape& s, const spdk_fv_shape& sc, const char* who, int& D) {
    if (int rc = fv_shapes_ok(c, s, sc, who)) return rc;
    if (int rc = fv_state_shape_ok(c, sc, who)) return rc;
    const int* a = c->prm.active;
    D = (a[0] != 0) + (a[1] != 0) + (a[2] != 0);
    const int q = s.nx;
    bool ok = s.nt == 1 && a[0] && (D < 2 || a[1]) && (D < 3 || a[2]) && q >= 1 && q <= 6;
    for (int m : {SPDK_M_SP_TO_CV, SPDK_M_CV_TO_SP}) ok = ok && c->mrows[m] == q && c->mcols[m] == q;
    if (!ok) {
        set_error(std::string(who) + ": needs one slice, leading active axes and at most 6 points per axis");
        return SPDK_E_SHAPE;
    }
    return SPDK_OK;
}
static FvUpd fv_upd(const spdk_ctx* c, int t_id, double dt, int update) {
    FvUpd a;
    for (int d = 0; d < 3; d++) { a.act[d] = c->prm.active[d]; a.g[d] = c->prm.g[d]; }
    a.grav = (a.g[0] != 0.0 || a.g[1] != 0.0 || a.g[2] != 0.0);
    a.wdt = c->wt_host[t_id] * dt;
    a.update = update;
    return a;
}

__global__ void __launch_bounds__(BS) k_state_min_final(double* __restrict__ part, int n) {
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

extern "C" {

int spdk_fv_state_from_sd(spdk_ctx* c, const double* U_sp, spdk_sd_shape s, double* U_fv, spdk_fv_shape sc) {
    if (!c || !U_sp || !U_fv) { set_error("fv_state_from_sd: null argument"); return SPDK_E_ARG; }
    int D = 0;
    if (int rc = fv_state_tile_ok(c, s, sc, "fv_state_from_sd", D)) return rc;
    const int gx = c->ngh[0], gy = c->ngh[1], gz = c->ngh[2];
    const int ex = s.Nx - 2 * gx, ey = s.Ny - 2 * gy, ez = s.Nz - 2 * gz;
    const double* M = c->mat[SPDK_M_SP_TO_CV];
    const int ldm = c->mcols[SPDK_M_SP_TO_CV];
#define SPDK_FVS_IN(QQ, DD)                                                                                           \
    case QQ: {                                                                                                        \
        constexpr int EPB = ElTile<QQ, DD>::EPB;                                                                      \
        const int gpr = (ex + EPB - 1) / EPB;                                                                         \
        const uint32_t groups = (uint32_t)((size_t)gpr * ey * ez);                                                    \
        if (groups) hipLaunchKernelGGL((k_fv_state_in<QQ, DD>), dim3((groups + kXcd - 1) / kXcd * kXcd), dim3(BS), 0,  \
                                       c->stream, U_sp, SD(s), U_fv,                                                 \
                                       FV(sc), M, ldm, gx, gy, gz, c->fgh[0], c->fgh[1], c->fgh[2], ex, ey, gpr,     \
                                       groups);                                                                      \
        break;                                                                                                        \
    }
#define SPDK_FVS_Q(DD) \
    switch (s.nx) { SPDK_FVS_IN(1, DD) SPDK_FVS_IN(2, DD) SPDK_FVS_IN(3, DD) SPDK_FVS_IN(4, DD) SPDK_FVS_IN(5, DD) SPDK_FVS_IN(6, DD) }
    if (D == 1) { SPDK_FVS_Q(1) }
    else if (D == 2) { SPDK_FVS_Q(2) }
    else { SPDK_FVS_Q(3) }
#undef SPDK_FVS_Q
#undef SPDK_FVS_IN
    SPDK_LAUNCH_CHECK("fv_state_from_sd");
    return SPDK_OK;
}

int spdk_fv_state_update_prims(spdk_ctx* c, double* U_next, double* W_new, double* W_old, const double* U_fv,
                               const double* Fx, const double* Fy, const double* Fz, spdk_fv_shape sc, int t_id,
                               double dt, int w_new_vars) {
    if (!c || !U_next || !W_new || !W_old || !U_fv || !Fx) {
        set_error("fv_state_update_prims: null argument");
        return SPDK_E_ARG;
    }
    if (U_next == U_fv) { set_error("fv_state_update_prims: U_next and U_fv must differ"); return SPDK_E_ARG; }
    if (int rc = fv_state_shape_ok(c, sc, "fv_state_update_prims")) return rc;
    if (t_id < 0 || t_id >= c->T) { set_error("fv_state_update_prims: slice"); return SPDK_E_ARG; }
    const FvUpd a = fv_upd(c, t_id, dt, 0);
    const int Mx = sc.Nx - 2 * c->fgh[0], My = sc.Ny - 2 * c->fgh[1], Mz = sc.Nz - 2 * c->fgh[2];
    if (a.act[2]) {  // z columns of 8 cells per thread
        constexpr int KC = SPDK_UPD_KC;
        const uint32_t tz = (uint32_t)((size_t)Mx * My * ((Mz + KC - 1) / KC));
        hipLaunchKernelGGL(k_fv_state_update_prims_z<KC>, dim3(xcd_grid(tz)), dim3(BS), 0, c->stream, U_next, W_new,
                           W_old, U_fv, Fx, Fy ? Fy : Fx, Fz, FV(sc), fv_stag(sc, 0, a.act[0]), fv_stag(sc, 1, a.act[1]),
                           fv_stag(sc, 2, a.act[2]), geo_of(c), a, c->prm.gamma, c->fgh[0], c->fgh[1], c->fgh[2], Mx,
                           My, Mz, (unsigned)w_new_vars & 0x1fu, tz);
        SPDK_LAUNCH_CHECK("fv_state_update_prims");
        return SPDK_OK;
    }
    const uint32_t total = (uint32_t)((size_t)Mx * My * Mz);
    hipLaunchKernelGGL(k_fv_state_update_prims, dim3(xcd_grid(total)), dim3(BS), 0, c->stream, U_next, W_new, W_old, U_fv,
                       Fx, Fy ? Fy : Fx, Fz ? Fz : Fx, FV(sc), fv_stag(sc, 0, a.act[0]), fv_stag(sc, 1, a.act[1]),
                       fv_stag(sc, 2, a.act[2]), geo_of(c), a, c->prm.gamma, c->fgh[0], c->fgh[1], c->fgh[2], Mx, My,
                       (unsigned)w_new_vars & 0x1fu, total);
    SPDK_LAUNCH_CHECK("fv_state_update_prims");
    return SPDK_OK;
}

int spdk_fv_state_update_blended(spdk_ctx* c, double* U_next, const double* U_fv, const double* Fx, const double* Fy,
                                 const double* Fz, const double* theta, spdk_fv_shape sc, int t_id, double dt) {
    if (!c || !U_next || !U_fv || !Fx || !theta) { set_error("fv_state_update_blended: null argument"); return SPDK_E_ARG; }
    if (U_next == U_fv) { set_error("fv_state_update_blended: U_next and U_fv must differ"); return SPDK_E_ARG; }
    if (int rc = fv_state_shape_ok(c, sc, "fv_state_update_blended")) return rc;
    if (t_id < 0 || t_id >= c->T) { set_error("fv_state_update_blended: slice"); return SPDK_E_ARG; }
    const FvUpd a = fv_upd(c, t_id, dt, 1);
    const int Mx = sc.Nx - 2 * c->fgh[0], My = sc.Ny - 2 * c->fgh[1], Mz = sc.Nz - 2 * c->fgh[2];
    const uint32_t total = (uint32_t)((size_t)Mx * My * Mz);
    hipLaunchKernelGGL(k_fv_state_update_blended, dim3(xcd_grid(total)), dim3(BS), 0, c->stream, U_next, U_fv, Fx,
                       Fy ? Fy : Fx, Fz ? Fz : Fx, theta, FV(sc), fv_stag(sc, 0, a.act[0]), fv_stag(sc, 1, a.act[1]),
                       fv_stag(sc, 2, a.act[2]), geo_of(c), a, c->fgh[0], c->fgh[1], c->fgh[2], Mx, My, total);
    SPDK_LAUNCH_CHECK("fv_state_update_blended");
    return SPDK_OK;
}

int spdk_fv_state_to_sd(spdk_ctx* c, const double* U_fv, spdk_fv_shape sc, double* U_sp, spdk_sd_shape s, double* W_cv,
                        double* dt_next) {
    if (!c || !U_sp || !U_fv) { set_error("fv_state_to_sd: null argument"); return SPDK_E_ARG; }
    int D = 0;
    if (int rc = fv_state_tile_ok(c, s, sc, "fv_state_to_sd", D)) return rc;
    const spdk_params& P = c->prm;
    DtArgs dta;
    dta.gm = P.gamma;
    dta.cfl = P.cfl;
    dta.nu = P.nu;
    for (int d = 0; d < 3; d++) { dta.h[d] = P.h[d]; dta.act[d] = P.active[d]; }
    double* part = dt_next ? c->red : nullptr;
    const int ldm = c->mcols[SPDK_M_SP_TO_CV];
    int nb = 0;
#define SPDK_FVS_OUT(QQ, DD)                                                                                          \
    case QQ: {                                                                                                        \
        constexpr int EPB = ElTile<QQ, DD>::EPB;                                                                      \
        const int gpr = (s.Nx + EPB - 1) / EPB;                                                                       \
        const uint32_t groups = (uint32_t)((size_t)gpr * s.Ny * s.Nz);                                               \
        nb = (int)std::min<size_t>(groups, (size_t)c->red_blocks);                                                    \
        if (nb) hipLaunchKernelGGL((k_fv_state_out<QQ, DD>), dim3(nb), dim3(BS), 0, c->stream, U_fv, FV(sc), U_sp, W_cv, \
                                   SD(s), c->mat[SPDK_M_CV_TO_SP], c->mat[SPDK_M_SP_TO_CV], ldm, c->ngh[0], c->ngh[1], \
                                   c->ngh[2], c->fgh[0], c->fgh[1], c->fgh[2], gpr, groups, dta, part);             \
        break;                                                                                                        \
    }
#define SPDK_FVS_Q(DD)                                                                                                \
    switch (s.nx) {                                                                                                   \
        SPDK_FVS_OUT(1, DD) SPDK_FVS_OUT(2, DD) SPDK_FVS_OUT(3, DD) SPDK_FVS_OUT(4, DD) SPDK_FVS_OUT(5, DD)           \
        SPDK_FVS_OUT(6, DD)                                                                                           \
    }
    if (D == 1) { SPDK_FVS_Q(1) }
    else if (D == 2) { SPDK_FVS_Q(2) }
    else { SPDK_FVS_Q(3) }
#undef SPDK_FVS_Q
#undef SPDK_FVS_OUT
    SPDK_LAUNCH_CHECK("fv_state_to_sd");
    if (!dt_next) return SPDK_OK;
    hipLaunchKernelGGL(k_state_min_final, dim3(1), dim3(BS), 0, c->stream, c->red, nb);
    SPDK_LAUNCH_CHECK("fv_state_min_final");
    SPDK_CHECK(hipMemcpyAsync(c->red_host, c->red, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SPDK_CHECK(hipStreamSynchronize(c->stream));
    *dt_next = c->red_host[0];
    return SPDK_OK;
}

int spdk_fill_fv_ghosts(spdk_ctx* c, double* U, spdk_fv_shape s) {
    if (c->prm.active[0] && s.Nx >= 5 && (!c->prm.active[1] || s.Ny >= 5) && (!c->prm.active[2] || s.Nz >= 5)) {
        BcAxes b;
        for (int d = 0; d < 3; d++) { b.type[d] = c->prm.bc[d]; b.act[d] = c->prm.active[d]; }
        const uint32_t nzl = b.act[2] ? 4u * s.Ny * s.Nx : 0u;
        const uint32_t zin = b.act[2] ? (uint32_t)(s.Nz - 4) : (uint32_t)s.Nz;
        const uint32_t nyl = b.act[1] ? zin * 4u * s.Nx : 0u;
        const uint32_t yin = b.act[1] ? (uint32_t)(s.Ny - 4) : (uint32_t)s.Ny;
        const uint32_t nxl = zin * yin * 4u;
        const uint32_t total = nzl + nyl + nxl;
        hipLaunchKernelGGL(k_fv_bc_all, dim3(nblocks(total, BS)), dim3(BS), 0, c->stream, U, FV(s), b, nzl, nyl, total);
        SPDK_LAUNCH_CHECK("fill_fv_ghosts");
        return SPDK_OK;
    }
    for (int d = 0; d < 3; d++) {
        if (!c->prm.active[d]) continue;
        uint32_t total = (uint32_t)((size_t)(d == 2 ? 2 : s.Nz) * (d == 1 ? 2 : s.Ny) * (d == 0 ? 2 : s.Nx));
        hipLaunchKernelGGL(k_fv_bc, dim3(nblocks(total, BS)), dim3(BS), 0, c->stream, U, FV(s), d, c->prm.bc[d], total);
        SPDK_LAUNCH_CHECK("fill_fv_ghosts");
    }
    return SPDK_OK;
}

int spdk_fv_primitives(spdk_ctx* c, const double* U, double* W, spdk_fv_shape s) {
    uint32_t total = (uint32_t)FV(s).plane();
    hipLaunchKernelGGL(k_fv_prim, dim3(nblocks(total, BS)), dim3(BS), 0, c->stream, U, W, FV(s), c->prm.gamma, total);
    SPDK_LAUNCH_CHECK("fv_primitives");
    return SPDK_OK;
}

static int detect_troubles(spdk_ctx* c, const double* Wn, const double* Wo, double* tr, double* alx, double* aly,
                           double* alz, uint16_t* abits, spdk_fv_shape s, int pad, int mask) {
    const bool bits = abits != nullptr;
    Det a;
    for (int d = 0; d < 3; d++) a.act[d] = c->prm.active[d];
    a.moore = c->prm.nad_moore;
    a.delta = c->prm.nad_delta;
    a.use_sed = c->prm.sed && c->p > 1;
    a.pad = pad;
    a.mask = mask;
    a.tol = c->prm.nad_tol;
    uint32_t total = (uint32_t)FV(s).plane();
    if (bits && a.act[0] && a.act[1] && a.act[2] && a.moore && s.Nx >= 3 && s.Ny >= 3 && s.Nz >= 3) {
        constexpr int KC = SPDK_NAD_KC;
        const uint32_t tz = (uint32_t)(s.Nx - 2) * (s.Ny - 2) * ((s.Nz - 2 + KC - 1) / KC);
        hipLaunchKernelGGL((k_nad_sed_z<KC>), dim3(xcd_grid(tz)), dim3(BS), 0, c->stream, Wn, Wo, tr, abits, FV(s),
                           a, geo_of(c), tz);
        SPDK_LAUNCH_CHECK("nad_sed_z");
        total = 0;
    }
#define SPDK_NAD(X, Y, Z)                                                                                       \
    do {                                                                                                        \
        if (bits)                                                                                               \
            hipLaunchKernelGGL((k_nad_sed<X, Y, Z, true>), dim3(xcd_grid(total)), dim3(BS), 0, c->stream, Wn, Wo, tr, \
                               nullptr, nullptr, nullptr, abits, FV(s), a, geo_of(c), total);                     \
        else                                                                                                    \
            hipLaunchKernelGGL((k_nad_sed<X, Y, Z, false>), dim3(xcd_grid(total)), dim3(BS), 0, c->stream, Wn, Wo, tr, \
                               alx, aly ? aly : alx, alz ? alz : alx, nullptr, FV(s), a, geo_of(c), total);       \
    } while (0)
    if (total) switch ((a.act[0] ? 4 : 0) | (a.act[1] ? 2 : 0) | (a.act[2] ? 1 : 0)) {
        case 0: SPDK_NAD(0, 0, 0); break;
        case 1: SPDK_NAD(0, 0, 1); break;
        case 2: SPDK_NAD(0, 1, 0); break;
        case 3: SPDK_NAD(0, 1, 1); break;
        case 4: SPDK_NAD(1, 0, 0); break;
        case 5: SPDK_NAD(1, 0, 1); break;
        case 6: SPDK_NAD(1, 1, 0); break;
        default: SPDK_NAD(1, 1, 1); break;
    }
#undef SPDK_NAD
    SPDK_LAUNCH_CHECK("nad_sed");
    uint32_t t2 = (uint32_t)((size_t)(s.Nz - 2 * a.act[2]) * (s.Ny - 2 * a.act[1]) * (s.Nx - 2 * a.act[0]));
#define SPDK_AXES(KER, ...)                                                                                     \
    switch ((a3[0] ? 4 : 0) | (a3[1] ? 2 : 0) | (a3[2] ? 1 : 0)) {                                              \
        case 0: hipLaunchKernelGGL((KER<0, 0, 0>), __VA_ARGS__); break;                                         \
        case 1: hipLaunchKernelGGL((KER<0, 0, 1>), __VA_ARGS__); break;                                         \
        case 2: hipLaunchKernelGGL((KER<0, 1, 0>), __VA_ARGS__); break;                                         \
        case 3: hipLaunchKernelGGL((KER<0, 1, 1>), __VA_ARGS__); break;                                         \
        case 4: hipLaunchKernelGGL((KER<1, 0, 0>), __VA_ARGS__); break;                                         \
        case 5: hipLaunchKernelGGL((KER<1, 0, 1>), __VA_ARGS__); break;                                         \
        case 6: hipLaunchKernelGGL((KER<1, 1, 0>), __VA_ARGS__); break;                                         \
        default: hipLaunchKernelGGL((KER<1, 1, 1>), __VA_ARGS__); break;                                        \
    }
#define SPDK_AXES_B(KER, ...)                                                                                   \
    switch ((a3[0] ? 4 : 0) | (a3[1] ? 2 : 0) | (a3[2] ? 1 : 0)) {                                              \
        case 0: hipLaunchKernelGGL((KER<0, 0, 0, true>), __VA_ARGS__); break;                                   \
        case 1: hipLaunchKernelGGL((KER<0, 0, 1, true>), __VA_ARGS__); break;                                   \
        case 2: hipLaunchKernelGGL((KER<0, 1, 0, true>), __VA_ARGS__); break;                                   \
        case 3: hipLaunchKernelGGL((KER<0, 1, 1, true>), __VA_ARGS__); break;                                   \
        case 4: hipLaunchKernelGGL((KER<1, 0, 0, true>), __VA_ARGS__); break;                                   \
        case 5: hipLaunchKernelGGL((KER<1, 0, 1, true>), __VA_ARGS__); break;                                   \
        case 6: hipLaunchKernelGGL((KER<1, 1, 0, true>), __VA_ARGS__); break;                                   \
        default: hipLaunchKernelGGL((KER<1, 1, 1, true>), __VA_ARGS__); break;                                  \
    }
    {
        const int* a3 = a.act;
        if (bits) {
            SPDK_AXES_B(k_relax_pad, dim3(xcd_grid(t2)), dim3(BS), 0, c->stream, Wn, tr, nullptr, nullptr, nullptr,
                        (const uint16_t*)abits, FV(s), a, t2)
        } else {
            SPDK_AXES(k_relax_pad, dim3(xcd_grid(t2)), dim3(BS), 0, c->stream, Wn, tr, alx, aly ? aly : alx,
                      alz ? alz : alx, (const uint16_t*)nullptr, FV(s), a, t2)
        }
    }
#undef SPDK_AXES_B
    SPDK_LAUNCH_CHECK("relax_pad");
    return SPDK_OK;
}

int spdk_detect_troubles(spdk_ctx* c, const double* Wn, const double* Wo, double* tr, double* alx, double* aly, double* alz,
                         spdk_fv_shape s, int pad, int mask) {
    if (!c || !Wn || !Wo || !tr || !alx) { set_error("detect_troubles: null argument"); return SPDK_E_ARG; }
    return detect_troubles(c, Wn, Wo, tr, alx, aly, alz, nullptr, s, pad, mask);
}

int spdk_detect_troubles_bits(spdk_ctx* c, const double* Wn, const double* Wo, double* tr, uint16_t* abits,
                              spdk_fv_shape s, int pad, int mask) {
    if (!c || !Wn || !Wo || !tr || !abits) { set_error("detect_troubles_bits: null argument"); return SPDK_E_ARG; }
    if ((mask & ~0x1f) != 0) { set_error("detect_troubles_bits: mask covers variables 0..4"); return SPDK_E_ARG; }
    return detect_troubles(c, Wn, Wo, tr, nullptr, nullptr, nullptr, abits, s, pad, mask);
}

int spdk_blend_theta(spdk_ctx* c, const double* tr, double* th_tmp, double* th, spdk_fv_shape s) {
    int ax = c->prm.active[0], ay = c->prm.active[1], az = c->prm.active[2];
    if (c->prm.blending && ax && ay && az) {  // 3-D: both passes on cell tiles, theta_tmp not written
        const int bx = (s.Nx - 2 + ThTile::TX - 1) / ThTile::TX, by = (s.Ny - 2 + ThTile::TY - 1) / ThTile::TY;
        const int bz = (s.Nz - 2 + ThTile::TZ - 1) / ThTile::TZ;
        const size_t nb = (size_t)bx * by * bz;
        hipLaunchKernelGGL(k_blend_ring_tile, dim3((unsigned)((nb + kXcd - 1) / kXcd * kXcd)), dim3(BS), 0, c->stream,
                           tr, th, FV(s), bx, by, bz);
        SPDK_LAUNCH_CHECK("blend_ring");
        (void)th_tmp;
    } else if (c->prm.blending) {
        uint32_t t2 = (uint32_t)((size_t)(s.Nz - 2 * az) * (s.Ny - 2 * ay) * (s.Nx - 2 * ax));
        const int a3[3] = {ax, ay, az};
        SPDK_AXES(k_blend, dim3(xcd_grid(t2)), dim3(BS), 0, c->stream, tr, th_tmp, FV(s), t2)
        SPDK_LAUNCH_CHECK("blend");
        SPDK_AXES(k_ring, dim3(xcd_grid(t2)), dim3(BS), 0, c->stream, th_tmp, th, FV(s), t2)
        SPDK_LAUNCH_CHECK("ring");
    } else {
        uint32_t t = (uint32_t)FV(s).plane();
        hipLaunchKernelGGL(k_theta_raw, dim3(nblocks(t, BS)), dim3(BS), 0, c->stream, tr, th, FV(s), t);
        SPDK_LAUNCH_CHECK("theta_raw");
    }
    return SPDK_OK;
}

int spdk_muscl_hllc_blend(spdk_ctx* c, const double* W, const double* th, double* Fx, double* Fy, double* Fz,
                          spdk_fv_shape s, double dt_sub) {
    int ax = c->prm.active[0], ay = c->prm.active[1], az = c->prm.active[2];
    if (ax && (ay || !az) && c->fb_tile) {  // active axes x, x-y or x-y-z: the tiled kernel
        const int Mx = s.Nx - 3, My = ay ? s.Ny - 3 : 1, Mz = az ? s.Nz - 3 : 1;
        if (Mx <= 0 || My <= 0 || Mz <= 0) return SPDK_OK;
#define SPDK_FBT(AY_, AZ_)                                                                                     \
    do {                                                                                                       \
        using TL = FbTile<AY_, AZ_>;                                                                           \
        const int bx = (Mx + TL::TX - 1) / TL::TX, by = (My + TL::TY - 1) / TL::TY, bz = (Mz + TL::TZ - 1) / TL::TZ; \
        const unsigned nb = (unsigned)((size_t)bx * by * bz);                                                  \
        hipLaunchKernelGGL((k_fallback_tile<AY_, AZ_>), dim3((nb + kXcd - 1) / kXcd * kXcd), dim3(BS), 0, c->stream, \
                           W, th, Fx, Fy ? Fy : Fx, Fz ? Fz : Fx, FV(s), fv_stag(s, 0, ax), fv_stag(s, 1, ay),       \
                           fv_stag(s, 2, az), geo_of(c), dt_sub, c->prm.gamma, bx, by, bz);                    \
    } while (0)
        if (az) SPDK_FBT(1, 1);
        else if (ay) SPDK_FBT(1, 0);
        else SPDK_FBT(0, 0);
#undef SPDK_FBT
        SPDK_LAUNCH_CHECK("muscl_hllc_tile");
        return SPDK_OK;
    }
    uint32_t total = (uint32_t)((size_t)(s.Nz - 4 * az + az) * (s.Ny - 4 * ay + ay) * (s.Nx - 4 * ax + ax));
    hipLaunchKernelGGL(k_fallback, dim3(xcd_grid(total)), dim3(BS), 0, c->stream, W, th, Fx, Fy ? Fy : Fx, Fz ? Fz : Fx,
                       FV(s), fv_stag(s, 0, ax), fv_stag(s, 1, ay), fv_stag(s, 2, az), geo_of(c), dt_sub, ax, ay, az,
                       c->prm.gamma, total);
    SPDK_LAUNCH_CHECK("muscl_hllc");
    return SPDK_OK;
}

}  // extern "C"
