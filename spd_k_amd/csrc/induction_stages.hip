// Induction stage kernels: staggered face B, edge-centred E (src/induction.cpp).
// Each kernel restates the reference function cited at its launcher, with
// one thread per output (element, point).
#include <hip/hip_runtime.h>
#include <algorithm>
#include "ctx.hpp"
#include "physics.hpp"

using namespace spdk;
static constexpr int BS = 256;

__device__ __forceinline__ int pick(int d, int i, int j, int k) { return d == 0 ? i : (d == 1 ? j : k); }

// Index of axis `a` of (ii, jj, kk) replaced by `l`.
__device__ __forceinline__ void along(int a, int l, int& kk, int& jj, int& ii) {
    if (a == 0) ii = l;
    else if (a == 1) jj = l;
    else kk = l;
}

// rotational_a_to_b: src/induction.cpp:94-137 (B = curl A at face points).
// B_dim = d_{dim1} A2 / d1 - d_{dim2} A1 / d2, dim1 = dim+1, dim2 = dim+2
// (mod 3). Each derivative sums over the flux points of its own axis (p+2;
// the reference's count q is the same in 3-D); a derivative along an
// inactive axis is zero (2-D runs).
__global__ void __launch_bounds__(BS) k_curl(const double* __restrict__ A1, SD s1, const double* __restrict__ A2, SD s2,
                                             double* __restrict__ B, SD sb, const double* __restrict__ Dm, int ldd,
                                             double d1, double d2, int dim, int act, int nfp, uint32_t total) {
    uint32_t idx = blockIdx.x * BS + threadIdx.x;
    if (idx >= total) return;
    int k, j, i, kk, jj, ii;
    split6(idx, sb.Nz, sb.Ny, sb.Nx, sb.nz, sb.ny, sb.nx, k, j, i, kk, jj, ii);
    const int a1 = dim == 2 ? 0 : dim + 1, a2 = a1 == 2 ? 0 : a1 + 1;
    const int q1 = (act >> a1) & 1 ? nfp : 0, q2 = (act >> a2) & 1 ? nfp : 0;
    const int r1 = pick(a1, ii, jj, kk), r2 = pick(a2, ii, jj, kk);
    for (int t = 0; t < sb.nt; t++) {
        double d1a2 = 0, d2a1 = 0;
        for (int ll = 0; ll < q1; ll++) {
            int c = kk, b = jj, a = ii;
            along(a1, ll, c, b, a);
            d1a2 += A2[s2.off(t, 0, k, j, i, c, b, a)] * Dm[r1 * ldd + ll];
        }
        for (int ll = 0; ll < q2; ll++) {
            int c = kk, b = jj, a = ii;
            along(a2, ll, c, b, a);
            d2a1 += A1[s1.off(t, 0, k, j, i, c, b, a)] * Dm[r2 * ldd + ll];
        }
        B[sb.off(t, 0, k, j, i, kk, jj, ii)] = d1a2 / d1 - d2a1 / d2;
    }
}

// compute_E: src/induction.cpp:139-208. b1 is B1 interpolated sp->fp along
// dim2, b2 is B2 along dim1, each over the solution points of its axis
// (p+1, the reference's q in 3-D); along an inactive axis the value is
// taken as is.
__global__ void __launch_bounds__(BS) k_compute_E(double* __restrict__ E, SD se, const double* __restrict__ B1, SD s1,
                                                  const double* __restrict__ B2, SD s2, const double* __restrict__ S2F,
                                                  int lds, double v1, double v2, int dim, int act, int nsp, uint32_t total) {
    uint32_t idx = blockIdx.x * BS + threadIdx.x;
    if (idx >= total) return;
    int k, j, i, kk, jj, ii;
    split6(idx, se.Nz, se.Ny, se.Nx, se.nz, se.ny, se.nx, k, j, i, kk, jj, ii);
    const int a1 = dim == 2 ? 0 : dim + 1, a2 = a1 == 2 ? 0 : a1 + 1;
    const bool on1 = (act >> a1) & 1, on2 = (act >> a2) & 1;
    const int r1 = pick(a1, ii, jj, kk), r2 = pick(a2, ii, jj, kk);
    for (int t = 0; t < se.nt; t++) {
        double b1 = 0, b2 = 0;
        if (on2) {
            for (int ll = 0; ll < nsp; ll++) {
                int c = kk, b = jj, a = ii;
                along(a2, ll, c, b, a);
                b1 += B1[s1.off(t, 0, k, j, i, c, b, a)] * S2F[r2 * lds + ll];
            }
        } else {
            b1 = B1[s1.off(t, 0, k, j, i, kk, jj, ii)];
        }
        if (on1) {
            for (int ll = 0; ll < nsp; ll++) {
                int c = kk, b = jj, a = ii;
                along(a1, ll, c, b, a);
                b2 += B2[s2.off(t, 0, k, j, i, c, b, a)] * S2F[r1 * lds + ll];
            }
        } else {
            b2 = B2[s2.off(t, 0, k, j, i, kk, jj, ii)];
        }
        double vals[8] = {v1 * b2 - v2 * b1, b1, b2, v1, v2, 0.0, b1, b2};
        for (int v = 0; v < 8; v++) E[se.off(t, v, k, j, i, kk, jj, ii)] = vals[v];
    }
}

// upwind + E_riemann_solver: src/induction.cpp:502-560 (in place)
__device__ __forceinline__ double upwind(double l, double r, double v) {
    if (v == 0) return 0.5 * (l + r);
    return v > 0 ? l : r;
}
__global__ void __launch_bounds__(BS) k_E_riemann(double* __restrict__ E, SD s, int dim, int v1, uint32_t total) {
    uint32_t idx = blockIdx.x * BS + threadIdx.x;
    if (idx >= total) return;
    int Nx = s.Nx - (dim == 0), Ny = s.Ny - (dim == 1), Nz = s.Nz - (dim == 2);
    int px = dim == 0 ? 1 : s.nx, py = dim == 1 ? 1 : s.ny, pz = dim == 2 ? 1 : s.nz;
    int k, j, i, kk, jj, ii;
    split6(idx, Nz, Ny, Nx, pz, py, px, k, j, i, kk, jj, ii);
    int n = pick(dim, s.nx, s.ny, s.nz);
    size_t oL, oR;
    if (dim == 0) { oL = s.off(0, 0, k, j, i, kk, jj, n - 1); oR = s.off(0, 0, k, j, i + 1, kk, jj, 0); }
    else if (dim == 1) { oL = s.off(0, 0, k, j, i, kk, n - 1, ii); oR = s.off(0, 0, k, j + 1, i, kk, 0, ii); }
    else { oL = s.off(0, 0, k, j, i, n - 1, jj, ii); oR = s.off(0, 0, k + 1, j, i, 0, jj, ii); }
    size_t slab = s.slab();
    for (int t = 0; t < s.nt; t++) {
        size_t tb = (size_t)t * s.nv * slab;
        double v = smax(E[tb + v1 * slab + oL], E[tb + v1 * slab + oR]);
        for (int var = 0; var < s.nv - 3; var++) {
            double e = upwind(E[tb + var * slab + oL], E[tb + var * slab + oR], v);
            E[tb + var * slab + oL] = e;
            E[tb + var * slab + oR] = e;
        }
        for (int var = s.nv - 2; var < s.nv; var++) {
            double e = upwind(E[tb + var * slab + oL], E[tb + var * slab + oR], 0.0);
            E[tb + var * slab + oL] = e;
            E[tb + var * slab + oR] = e;
        }
    }
}

// dB/dt at one face point (src/induction.cpp:331-356 / :391-415):
// + d_{dim2} E1 / d2 - d_{dim1} E2 / d1, interleaved per flux point l as the
// reference accumulates them; a derivative along an inactive axis is zero.
__device__ __forceinline__ double dBdt_at(const double* E1, SD s1, const double* E2, SD s2, const double* Dm, int ldd,
                                          double d1, double d2, int dim, int t, int k, int j, int i, int kk, int jj, int ii,
                                          int nfp, int act) {
    const int a1 = dim == 2 ? 0 : dim + 1, a2 = a1 == 2 ? 0 : a1 + 1;
    const bool on1 = (act >> a1) & 1, on2 = (act >> a2) & 1;
    const int r1 = pick(a1, ii, jj, kk), r2 = pick(a2, ii, jj, kk);
    double dB = 0, E;
    for (int ll = 0; ll < nfp; ll++) {
        if (on2) {
            int c = kk, b = jj, a = ii;
            along(a2, ll, c, b, a);
            E = E1[s1.off(t, 0, k, j, i, c, b, a)];
            dB += E * Dm[r2 * ldd + ll] / d2;
        }
        if (on1) {
            int c = kk, b = jj, a = ii;
            along(a1, ll, c, b, a);
            E = E2[s2.off(t, 0, k, j, i, c, b, a)];
            dB -= E * Dm[r1 * ldd + ll] / d1;
        }
    }
    return dB;
}

// update_B_prediction (src/induction.cpp:303-365) / update_B_solution
// (:367-420), active elements only
__global__ void __launch_bounds__(BS) k_B_update(const double* __restrict__ B, double* __restrict__ Bout, SD sb, int nader,
                                                 const double* __restrict__ E1, SD s1, const double* __restrict__ E2, SD s2,
                                                 const double* __restrict__ Dm, int ldd, const double* __restrict__ inv,
                                                 const double* __restrict__ w, double d1, double d2, double dt, int dim,
                                                 int predict, int gx, int gy, int gz, int act, int nfp, uint32_t total) {
    uint32_t idx = blockIdx.x * BS + threadIdx.x;
    if (idx >= total) return;
    int k, j, i, kk, jj, ii;
    split6(idx, sb.Nz - 2 * gz, sb.Ny - 2 * gy, sb.Nx - 2 * gx, sb.nz, sb.ny, sb.nx, k, j, i, kk, jj, ii);
    k += gz; j += gy; i += gx;
    size_t o = sb.off(0, 0, k, j, i, kk, jj, ii);
    size_t slab = sb.slab();
    if (predict) {
        double dBdt[10];
        double B_old = B[o];
        for (int t = 0; t < nader; t++) dBdt[t] = dBdt_at(E1, s1, E2, s2, Dm, ldd, d1, d2, dim, t, k, j, i, kk, jj, ii, nfp, act);
        for (int t = 0; t < nader; t++) {
            double dB = 0;
            for (int l = 0; l < nader; l++) dB += dBdt[l] * inv[t * nader + l] * w[l] * dt;
            Bout[(size_t)t * slab + o] = B_old - dB;
        }
    } else {
        double dB = 0;
        for (int t = 0; t < nader; t++) dB += dBdt_at(E1, s1, E2, s2, Dm, ldd, d1, d2, dim, t, k, j, i, kk, jj, ii, nfp, act) * w[t] * dt;
        Bout[o] -= dB;
    }
}

// compute_B2_cv: src/induction.cpp:422-465 (all elements). Along an inactive
// axis the single point is taken with weight 1.
__global__ void __launch_bounds__(BS) k_B2_cv(double* __restrict__ B2, SD s, const double* __restrict__ Bx, SD sx,
                                              const double* __restrict__ By, SD sy, const double* __restrict__ Bz, SD sz,
                                              const double* __restrict__ F2C, int ldf, const double* __restrict__ S2C,
                                              int ldc, int act, uint32_t total) {
    uint32_t idx = blockIdx.x * BS + threadIdx.x;
    if (idx >= total) return;
    int k, j, i, kk, jj, ii;
    split6(idx, s.Nz, s.Ny, s.Nx, s.nz, s.ny, s.nx, k, j, i, kk, jj, ii);
    const bool ax = act & 1, ay = (act >> 1) & 1, az = (act >> 2) & 1;
    int px = s.nx, py = s.ny, pz = s.nz;
    int qx = sx.nx, qy = sy.ny, qz = sz.nz;  // flux points along x, y, z
    auto sc = [&](bool on, int r, int c) { return on ? S2C[r * ldc + c] : 1.0; };
    auto fc = [&](bool on, int r, int c) { return on ? F2C[r * ldf + c] : 1.0; };
    double bx = 0, by = 0, bz = 0;
    for (int nn = 0; nn < pz; nn++)
        for (int mm = 0; mm < py; mm++)
            for (int ll = 0; ll < qx; ll++)
                bx += Bx[sx.off(0, 0, k, j, i, nn, mm, ll)] * sc(az, kk, nn) * sc(ay, jj, mm) * fc(ax, ii, ll);
    for (int nn = 0; nn < pz; nn++)
        for (int mm = 0; mm < qy; mm++)
            for (int ll = 0; ll < px; ll++)
                by += By[sy.off(0, 0, k, j, i, nn, mm, ll)] * sc(az, kk, nn) * fc(ay, jj, mm) * sc(ax, ii, ll);
    for (int nn = 0; nn < qz; nn++)
        for (int mm = 0; mm < py; mm++)
            for (int ll = 0; ll < px; ll++)
                bz += Bz[sz.off(0, 0, k, j, i, nn, mm, ll)] * fc(az, kk, nn) * sc(ay, jj, mm) * sc(ax, ii, ll);
    B2[s.off(0, 0, k, j, i, kk, jj, ii)] = bx;
    B2[s.off(0, 1, k, j, i, kk, jj, ii)] = by;
    B2[s.off(0, 2, k, j, i, kk, jj, ii)] = bz;
    B2[s.off(0, 3, k, j, i, kk, jj, ii)] = bx * bx + by * by + bz * bz;
}

// compute_B_cv_from_cf: src/induction.cpp:467-500 (active elements, cv points)
__global__ void __launch_bounds__(BS) k_B_cv_from_cf(double* __restrict__ B, FV sf, const double* __restrict__ Bx, SD sx,
                                                     const double* __restrict__ By, SD sy, const double* __restrict__ Bz,
                                                     SD sz, const double* __restrict__ F2C, int ldf, int gx, int gy, int gz,
                                                     int fx, int fy, int fz, uint32_t total) {
    uint32_t idx = blockIdx.x * BS + threadIdx.x;
    if (idx >= total) return;
    int px = sx.nx - 1, py = sx.ny, pz = sx.nz, q = px + 1;
    int k, j, i, kk, jj, ii;
    split6(idx, sx.Nz - 2 * gz, sx.Ny - 2 * gy, sx.Nx - 2 * gx, pz, py, px, k, j, i, kk, jj, ii);
    k += gz; j += gy; i += gx;
    int I = ii + fx + (i - gx) * px, J = jj + fy + (j - gy) * py, K = kk + fz + (k - gz) * pz;
    double bx = 0, by = 0, bz = 0;
    for (int ll = 0; ll < q; ll++) {
        bx += Bx[sx.off(0, 0, k, j, i, kk, jj, ll)] * F2C[ii * ldf + ll];
        by += By[sy.off(0, 0, k, j, i, kk, ll, ii)] * F2C[jj * ldf + ll];
        bz += Bz[sz.off(0, 0, k, j, i, ll, jj, ii)] * F2C[kk * ldf + ll];
    }
    B[sf.off(0, K, J, I)] = bx;
    B[sf.off(1, K, J, I)] = by;
    B[sf.off(2, K, J, I)] = bz;
    B[sf.off(3, K, J, I)] = bx * bx + by * by + bz * bz;
}

// edge_integral: src/induction.cpp:602-654
__global__ void __launch_bounds__(BS) k_edge_integral(const double* __restrict__ Eep, SD s, double* __restrict__ E, FV sf,
                                                      const double* __restrict__ S2C, int ldc, int t_id, int dim, int gx,
                                                      int gy, int gz, int fx, int fy, int fz, uint32_t total) {
    uint32_t idx = blockIdx.x * BS + threadIdx.x;
    if (idx >= total) return;
    int Nx = s.Nx + (dim == 0 ? 0 : 1), Ny = s.Ny + (dim == 1 ? 0 : 1), Nz = s.Nz + (dim == 2 ? 0 : 1);
    int px = s.nx - (dim == 0 ? 0 : 1), py = s.ny - (dim == 1 ? 0 : 1), pz = s.nz - (dim == 2 ? 0 : 1);
    int k, j, i, kk, jj, ii;
    split6(idx, Nz - 2 * gz, Ny - 2 * gy, Nx - 2 * gx, pz, py, px, k, j, i, kk, jj, ii);
    k += gz; j += gy; i += gx;
    int I = ii + fx + (i - gx) * px, J = jj + fy + (j - gy) * py, K = kk + fz + (k - gz) * pz;
    if (!(K < sf.Nz && J < sf.Ny && I < sf.Nx)) return;
    int q = pick(dim, px, py, pz);
    int id = pick(dim, ii, jj, kk);
    double e = 0;
    for (int ll = 0; ll < q; ll++) {
        int a = dim == 0 ? ll : ii, b = dim == 1 ? ll : jj, c = dim == 2 ? ll : kk;
        double x = Eep[s.off(t_id, 0, k, j, i, c, b, a)];
        x *= S2C[id * ldc + ll];
        e += x;
    }
    E[sf.off(0, K, J, I)] = e;
}

// fv_update_B_solution: src/induction.cpp:656-720. The last face point of
// element i and the first of element i+1 map to one FV face; both threads
// write the same bits there.
__global__ void __launch_bounds__(BS) k_fv_update_B(double* __restrict__ Bn, double* __restrict__ Bo, FV sn,
                                                    double* __restrict__ B, SD sb, const double* __restrict__ E1, FV s1,
                                                    const double* __restrict__ E2, FV s2, const double* __restrict__ f1,
                                                    const double* __restrict__ f2, double w, double dt, int dim, int update, int gx,
                                                    int gy, int gz, int fx, int fy, int fz, uint32_t total) {
    uint32_t idx = blockIdx.x * BS + threadIdx.x;
    if (idx >= total) return;
    int k, j, i, kk, jj, ii;
    split6(idx, sb.Nz - 2 * gz, sb.Ny - 2 * gy, sb.Nx - 2 * gx, sb.nz, sb.ny, sb.nx, k, j, i, kk, jj, ii);
    k += gz; j += gy; i += gx;
    int qx = sb.nx - (dim == 0), qy = sb.ny - (dim == 1), qz = sb.nz - (dim == 2);
    int I = ii + fx + (i - gx) * qx, J = jj + fy + (j - gy) * qy, K = kk + fz + (k - gz) * qz;
    if (!(K < sn.Nz && J < sn.Ny && I < sn.Nx)) return;
    int dim1 = dim == 0 ? 1 : (dim == 1 ? 2 : 0), dim2 = dim1 == 0 ? 1 : (dim1 == 1 ? 2 : 0);
    size_t o = sb.off(0, 0, k, j, i, kk, jj, ii);
    double b_old = B[o];
    Bo[sn.off(0, K, J, I)] = b_old;
    double ep = E1[s1.off(0, K + (dim2 == 2), J + (dim2 == 1), I + (dim2 == 0))];
    int id = pick(dim2, I, J, K);
    double h = f2[id + 1] - f2[id];
    double dBdt = (ep - E1[s1.off(0, K, J, I)]) / h;
    ep = E2[s2.off(0, K + (dim1 == 2), J + (dim1 == 1), I + (dim1 == 0))];
    id = pick(dim1, I, J, K);
    h = f1[id + 1] - f1[id];
    dBdt -= (ep - E2[s2.off(0, K, J, I)]) / h;
    double b_new = b_old - dBdt * w * dt;  // (dBdt*w)*dt, as src/induction.cpp:712
    Bn[sn.off(0, K, J, I)] = b_new;
    if (update) B[o] = b_new;
}

// boundaries, FV with face alignment (src/boundary.cpp:123-190) along one axis;
// transverse extents are the array's own.
__global__ void __launch_bounds__(BS) k_fv_bc_aligned(double* __restrict__ U, FV s, int dim, int type, int shift,
                                                      uint32_t total) {
    uint32_t idx = blockIdx.x * BS + threadIdx.x;
    if (idx >= total) return;
    int Nx = dim == 0 ? 2 : s.Nx, Ny = dim == 1 ? 2 : s.Ny;
    int k, j, i;
    split3(idx, Ny, Nx, k, j, i);
    int l = pick(dim, i, j, k);
    int N = pick(dim, s.Nx, s.Ny, s.Nz) - shift;
    auto at = [&](int v, int ll) -> size_t {
        if (dim == 0) return s.off(v, k, j, ll);
        if (dim == 1) return s.off(v, k, ll, i);
        return s.off(v, ll, j, i);
    };
    for (int v = 0; v < s.nv; v++) {
        double BL, BR;
        if (type == SPDK_BC_PERIODIC) { BL = U[at(v, N - 4 + l - shift)]; BR = U[at(v, 2 + l + shift)]; }
        else if (type == SPDK_BC_GRADFREE) { BL = U[at(v, 2 + l + shift)]; BR = U[at(v, N - 4 + l - shift)]; }
        else {
            double sgn = (v == 1 + dim) ? -1.0 : 1.0;
            BL = sgn * U[at(v, 3 - l)];
            BR = sgn * U[at(v, N - 3 - l)];
        }
        U[at(v, l)] = BL;
        U[at(v, N - 2 + l)] = BR;
    }
}

// Not in the reference: per-block max |div B| over the solution points of the
// active elements (each face component differentiated fp->sp along its own
// axis, terms added x, y, z) and the partial sums of |B|^2 (B2 variable 3) x
// cv volume (products of FV face widths along the active axes). part[b] and
// part[nb + b] receive block b's max and sum.
__global__ void __launch_bounds__(BS) k_ind_diag(const double* __restrict__ Bx, SD sx, const double* __restrict__ By, SD sy,
                                                 const double* __restrict__ Bz, SD sz, const double* __restrict__ B2, SD s2,
                                                 const double* __restrict__ Dm, int ldd, const double* __restrict__ fx,
                                                 const double* __restrict__ fy, const double* __restrict__ fz, double hx,
                                                 double hy, double hz, int act, int nfp, int gx, int gy, int gz, int cx, int cy,
                                                 int cz, double* __restrict__ part, uint32_t total) {
    __shared__ double smax_[BS], ssum_[BS];
    double m = 0, e = 0;
    const double* Bd[3] = {Bx, By, Bz};
    const SD sd[3] = {sx, sy, sz};
    const double* fv[3] = {fx, fy, fz};
    const double hd[3] = {hx, hy, hz};
    const int cg[3] = {cx, cy, cz}, eg[3] = {gx, gy, gz};
    for (uint32_t idx = blockIdx.x * BS + threadIdx.x; idx < total; idx += gridDim.x * BS) {
        int k, j, i, kk, jj, ii;
        split6(idx, s2.Nz - 2 * gz, s2.Ny - 2 * gy, s2.Nx - 2 * gx, s2.nz, s2.ny, s2.nx, k, j, i, kk, jj, ii);
        k += gz; j += gy; i += gx;
        double div = 0;
        for (int d = 0; d < 3; d++) {
            if (!((act >> d) & 1)) continue;
            int r = pick(d, ii, jj, kk);
            double sum = 0;
            for (int l = 0; l < nfp; l++) {
                int c = kk, b = jj, a = ii;
                along(d, l, c, b, a);
                sum += Bd[d][sd[d].off(0, 0, k, j, i, c, b, a)] * Dm[r * ldd + l];
            }
            div += sum / hd[d];
        }
        m = nan_max(m, fabs(div));
        double vol = 1;
        const int el[3] = {i, j, k}, pt[3] = {ii, jj, kk}, np[3] = {s2.nx, s2.ny, s2.nz};
        for (int d = 0; d < 3; d++) {
            if (!((act >> d) & 1)) continue;
            int I = cg[d] + (el[d] - eg[d]) * np[d] + pt[d];
            vol *= fv[d][I + 1] - fv[d][I];
        }
        e += B2[s2.off(0, 3, k, j, i, kk, jj, ii)] * vol;
    }
    smax_[threadIdx.x] = m;
    ssum_[threadIdx.x] = e;
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            smax_[threadIdx.x] = nan_max(smax_[threadIdx.x], smax_[threadIdx.x + w]);
            ssum_[threadIdx.x] += ssum_[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = smax_[0];
        part[gridDim.x + blockIdx.x] = ssum_[0];
    }
}
__global__ void __launch_bounds__(BS) k_ind_diag_final(double* __restrict__ part, int nb) {
    __shared__ double smax_[BS], ssum_[BS];
    double m = 0, e = 0;
    for (int b = threadIdx.x; b < nb; b += BS) {
        m = nan_max(m, part[b]);
        e += part[nb + b];
    }
    smax_[threadIdx.x] = m;
    ssum_[threadIdx.x] = e;
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            smax_[threadIdx.x] = nan_max(smax_[threadIdx.x], smax_[threadIdx.x + w]);
            ssum_[threadIdx.x] += ssum_[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[0] = smax_[0];
        part[1] = ssum_[0];
    }
}

// ===========================================================================
extern "C" {

static int act_mask(const spdk_ctx* c) {
    return (c->prm.active[0] ? 1 : 0) | (c->prm.active[1] ? 2 : 0) | (c->prm.active[2] ? 4 : 0);
}

#define IND_TOTAL(s) ((uint32_t)((size_t)(s).Nz * (s).Ny * (s).Nx * (s).nz * (s).ny * (s).nx))

int spdk_B_curl(spdk_ctx* c, const double* A1, spdk_sd_shape s1, const double* A2, spdk_sd_shape s2, double* B,
                spdk_sd_shape sb, double d1, double d2, int dim) {
    uint32_t total = IND_TOTAL(sb);
    hipLaunchKernelGGL(k_curl, dim3(nblocks(total, BS)), dim3(BS), 0, c->stream, A1, SD(s1), A2, SD(s2), B, SD(sb),
                       c->mat[SPDK_M_DFP_TO_SP], c->mcols[SPDK_M_DFP_TO_SP], d1, d2, dim, act_mask(c), c->m, total);
    SPDK_LAUNCH_CHECK("B_curl");
    return SPDK_OK;
}

int spdk_induction_E(spdk_ctx* c, double* E, spdk_sd_shape se, const double* B1, spdk_sd_shape s1, const double* B2,
                     spdk_sd_shape s2, int dim) {
    if (se.nv != 8) { set_error("edge E arrays carry 8 variables"); return SPDK_E_SHAPE; }
    int dim1 = dim == 0 ? 1 : (dim == 1 ? 2 : 0), dim2 = dim1 == 0 ? 1 : (dim1 == 1 ? 2 : 0);
    uint32_t total = IND_TOTAL(se);
    hipLaunchKernelGGL(k_compute_E, dim3(nblocks(total, BS)), dim3(BS), 0, c->stream, E, SD(se), B1, SD(s1), B2, SD(s2),
                       c->mat[SPDK_M_SP_TO_FP], c->mcols[SPDK_M_SP_TO_FP], c->prm.vel[dim1], c->prm.vel[dim2], dim,
                       act_mask(c), c->n, total);
    SPDK_LAUNCH_CHECK("induction_E");
    return SPDK_OK;
}

int spdk_E_riemann(spdk_ctx* c, double* E, spdk_sd_shape s, int dim, int v1) {
    int Nx = s.Nx - (dim == 0), Ny = s.Ny - (dim == 1), Nz = s.Nz - (dim == 2);
    int px = dim == 0 ? 1 : s.nx, py = dim == 1 ? 1 : s.ny, pz = dim == 2 ? 1 : s.nz;
    uint32_t total = (uint32_t)((size_t)Nz * Ny * Nx * pz * py * px);
    hipLaunchKernelGGL(k_E_riemann, dim3(nblocks(total, BS)), dim3(BS), 0, c->stream, E, SD(s), dim, v1, total);
    SPDK_LAUNCH_CHECK("E_riemann");
    return SPDK_OK;
}

static int B_update(spdk_ctx* c, const double* B, double* Bout, spdk_sd_shape sb, int nader, const double* E1,
                    spdk_sd_shape s1, const double* E2, spdk_sd_shape s2, double d1, double d2, double dt, int dim,
                    int predict) {
    int gx = c->ngh[0], gy = c->ngh[1], gz = c->ngh[2];
    uint32_t total = (uint32_t)((size_t)(sb.Nz - 2 * gz) * (sb.Ny - 2 * gy) * (sb.Nx - 2 * gx) * sb.nz * sb.ny * sb.nx);
    hipLaunchKernelGGL(k_B_update, dim3(nblocks(total, BS)), dim3(BS), 0, c->stream, B, Bout, SD(sb), nader, E1, SD(s1), E2,
                       SD(s2), c->mat[SPDK_M_DFP_TO_SP], c->mcols[SPDK_M_DFP_TO_SP], c->mat[SPDK_M_INVADER], c->wt, d1, d2,
                       dt, dim, predict, gx, gy, gz, act_mask(c), c->m, total);
    SPDK_LAUNCH_CHECK("B_update");
    return SPDK_OK;
}
int spdk_B_predict(spdk_ctx* c, const double* B, double* B_ader, spdk_sd_shape s_ader, const double* E1, spdk_sd_shape s1,
                   const double* E2, spdk_sd_shape s2, double d1, double d2, double dt, int dim) {
    spdk_sd_shape sb = s_ader;
    sb.nt = 1;
    return B_update(c, B, B_ader, sb, s_ader.nt, E1, s1, E2, s2, d1, d2, dt, dim, 1);
}
int spdk_B_correct(spdk_ctx* c, double* B, spdk_sd_shape sb, const double* E1, spdk_sd_shape s1, const double* E2,
                   spdk_sd_shape s2, double d1, double d2, double dt, int dim) {
    return B_update(c, B, B, sb, s1.nt, E1, s1, E2, s2, d1, d2, dt, dim, 0);
}

int spdk_B2_cv(spdk_ctx* c, double* B2, spdk_sd_shape s, const double* Bx, spdk_sd_shape sx, const double* By,
               spdk_sd_shape sy, const double* Bz, spdk_sd_shape sz) {
    uint32_t total = IND_TOTAL(s);
    hipLaunchKernelGGL(k_B2_cv, dim3(nblocks(total, BS)), dim3(BS), 0, c->stream, B2, SD(s), Bx, SD(sx), By, SD(sy), Bz,
                       SD(sz), c->mat[SPDK_M_FP_TO_CV], c->mcols[SPDK_M_FP_TO_CV], c->mat[SPDK_M_SP_TO_CV],
                       c->mcols[SPDK_M_SP_TO_CV], act_mask(c), total);
    SPDK_LAUNCH_CHECK("B2_cv");
    return SPDK_OK;
}

int spdk_B_cv_from_cf(spdk_ctx* c, double* B, spdk_fv_shape sf, const double* Bx, spdk_sd_shape sx, const double* By,
                      spdk_sd_shape sy, const double* Bz, spdk_sd_shape sz) {
    int gx = c->ngh[0], gy = c->ngh[1], gz = c->ngh[2];
    uint32_t total = (uint32_t)((size_t)(sx.Nz - 2 * gz) * (sx.Ny - 2 * gy) * (sx.Nx - 2 * gx) * sx.nz * sx.ny * (sx.nx - 1));
    hipLaunchKernelGGL(k_B_cv_from_cf, dim3(nblocks(total, BS)), dim3(BS), 0, c->stream, B, FV(sf), Bx, SD(sx), By, SD(sy),
                       Bz, SD(sz), c->mat[SPDK_M_FP_TO_CV], c->mcols[SPDK_M_FP_TO_CV], gx, gy, gz, c->fgh[0], c->fgh[1],
                       c->fgh[2], total);
    SPDK_LAUNCH_CHECK("B_cv_from_cf");
    return SPDK_OK;
}

int spdk_edge_integral(spdk_ctx* c, const double* Eep, spdk_sd_shape s, double* E, spdk_fv_shape sf, int t_id, int dim) {
    int gx = c->ngh[0], gy = c->ngh[1], gz = c->ngh[2];
    int Nx = s.Nx + (dim == 0 ? 0 : 1), Ny = s.Ny + (dim == 1 ? 0 : 1), Nz = s.Nz + (dim == 2 ? 0 : 1);
    int px = s.nx - (dim == 0 ? 0 : 1), py = s.ny - (dim == 1 ? 0 : 1), pz = s.nz - (dim == 2 ? 0 : 1);
    uint32_t total = (uint32_t)((size_t)(Nz - 2 * gz) * (Ny - 2 * gy) * (Nx - 2 * gx) * pz * py * px);
    hipLaunchKernelGGL(k_edge_integral, dim3(nblocks(total, BS)), dim3(BS), 0, c->stream, Eep, SD(s), E, FV(sf),
                       c->mat[SPDK_M_SP_TO_CV], c->mcols[SPDK_M_SP_TO_CV], t_id, dim, gx, gy, gz, c->fgh[0], c->fgh[1],
                       c->fgh[2], total);
    SPDK_LAUNCH_CHECK("edge_integral");
    return SPDK_OK;
}

int spdk_fv_update_B(spdk_ctx* c, double* Bn, double* Bo, spdk_fv_shape sn, double* B, spdk_sd_shape sb, const double* E1,
                     spdk_fv_shape s1, const double* E2, spdk_fv_shape s2, int t_id, double dt, int dim, int update) {
    if (t_id < 0 || t_id >= c->T) { set_error("fv_update_B: slice"); return SPDK_E_ARG; }
    int dim1 = dim == 0 ? 1 : (dim == 1 ? 2 : 0), dim2 = dim1 == 0 ? 1 : (dim1 == 1 ? 2 : 0);
    int gx = c->ngh[0], gy = c->ngh[1], gz = c->ngh[2];
    uint32_t total = (uint32_t)((size_t)(sb.Nz - 2 * gz) * (sb.Ny - 2 * gy) * (sb.Nx - 2 * gx) * sb.nz * sb.ny * sb.nx);
    hipLaunchKernelGGL(k_fv_update_B, dim3(nblocks(total, BS)), dim3(BS), 0, c->stream, Bn, Bo, FV(sn), B, SD(sb), E1, FV(s1),
                       E2, FV(s2), c->fvf[dim1], c->fvf[dim2], c->wt_host[t_id], dt, dim, update, gx, gy, gz, c->fgh[0],
                       c->fgh[1], c->fgh[2], total);
    SPDK_LAUNCH_CHECK("fv_update_B");
    return SPDK_OK;
}

int spdk_fill_fv_ghosts_aligned(spdk_ctx* c, double* U, spdk_fv_shape s, int alignment, int a_dim) {
    for (int d = 0; d < 3; d++) {
        if (!c->prm.active[d]) continue;
        int shift = alignment * (a_dim == d);
        uint32_t total = (uint32_t)((size_t)(d == 2 ? 2 : s.Nz) * (d == 1 ? 2 : s.Ny) * (d == 0 ? 2 : s.Nx));
        hipLaunchKernelGGL(k_fv_bc_aligned, dim3(nblocks(total, BS)), dim3(BS), 0, c->stream, U, FV(s), d, c->prm.bc[d], shift,
                           total);
        SPDK_LAUNCH_CHECK("fill_fv_ghosts_aligned");
    }
    return SPDK_OK;
}

int spdk_induction_diagnostics(spdk_ctx* c, const double* Bx, spdk_sd_shape sx, const double* By, spdk_sd_shape sy,
                               const double* Bz, spdk_sd_shape sz, const double* B2, spdk_sd_shape s2, double* out) {
    if (!out || s2.nv != 4) { set_error("induction_diagnostics: B2 carries 4 variables, out 2 doubles"); return SPDK_E_ARG; }
    int gx = c->ngh[0], gy = c->ngh[1], gz = c->ngh[2];
    uint32_t total = (uint32_t)((size_t)(s2.Nz - 2 * gz) * (s2.Ny - 2 * gy) * (s2.Nx - 2 * gx) * s2.nz * s2.ny * s2.nx);
    // part[] holds 2 values per block inside the context's reduction buffer
    int nb = (int)std::min<size_t>(c->red_blocks / 2, nblocks(total, BS));
    if (nb < 1) nb = 1;
    const spdk_params& P = c->prm;
    hipLaunchKernelGGL(k_ind_diag, dim3(nb), dim3(BS), 0, c->stream, Bx, SD(sx), By, SD(sy), Bz, SD(sz), B2, SD(s2),
                       c->mat[SPDK_M_DFP_TO_SP], c->mcols[SPDK_M_DFP_TO_SP], c->fvf[0], c->fvf[1], c->fvf[2], P.h[0], P.h[1],
                       P.h[2], act_mask(c), c->m, gx, gy, gz, c->fgh[0], c->fgh[1], c->fgh[2], c->red, total);
    SPDK_LAUNCH_CHECK("ind_diag");
    hipLaunchKernelGGL(k_ind_diag_final, dim3(1), dim3(BS), 0, c->stream, c->red, nb);
    SPDK_LAUNCH_CHECK("ind_diag_final");
    SPDK_CHECK(hipMemcpyAsync(c->red_host, c->red, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SPDK_CHECK(hipStreamSynchronize(c->stream));
    out[0] = c->red_host[0];
    out[1] = c->red_host[1];
    return SPDK_OK;
}

}  // extern "C"
