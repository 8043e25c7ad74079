// Point-wise hydro physics shared by the stage and fused kernels.
//
// Every expression is written so that, compiled with -ffp-contract=off, each
// multiply, add, divide and square root rounds exactly where the reference's
// does (src/hydro.cpp); the FP64 kernels are then bit-identical to the
// reference arithmetic.
#pragma once
#include <hip/hip_runtime.h>

namespace spdk {

constexpr double kMinC2 = 1e-14;                                        // src/define.hpp:14
constexpr double kRhoMin = 1e-10, kRhoMax = 1e10, kPMin = 1e-10, kPMax = 1e10;  // define.hpp:16-19
constexpr int NV = 6;  // rho, m1, m2, m3, E, slot (define.hpp:35)
constexpr int NVP = 5; // physical variables

// first-argument-wins max/min, the semantics of std::max/std::min
__host__ __device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }
__host__ __device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }

// src/hydro.cpp:27-32
__device__ __forceinline__ double sound_speed(double rho, double p, double gm) {
    double c2 = gm * p / rho;
    c2 = smax(c2, kMinC2);
    return sqrt(c2);
}

// Primitive variables of one state (src/hydro.cpp:95-106). E_kin starts at
// +0.0 and accumulates in x, y, z order.
__device__ __forceinline__ void prim(const double* u, double* w, double gm) {
    double ek = 0;
    w[0] = u[0];
    w[1] = u[1] / u[0];
    ek += w[1] * u[1];
    w[2] = u[2] / u[0];
    ek += w[2] * u[2];
    w[3] = u[3] / u[0];
    ek += w[3] * u[3];
    w[4] = (u[4] - 0.5 * ek) * (gm - 1.);
}

// src/hydro.cpp:34-45
__device__ __forceinline__ void cons(const double* w, double* u, double gm) {
    double ek = 0;
    u[0] = w[0];
    u[1] = w[0] * w[1];
    ek += w[1] * u[1];
    u[2] = w[0] * w[2];
    ek += w[2] * u[2];
    u[3] = w[0] * w[3];
    ek += w[3] * u[3];
    u[4] = w[4] / (gm - 1.) + 0.5 * ek;
}

// Euler flux normal to axis D (src/hydro.cpp:156-162); V1 = 1 + D and the
// two transverse momenta follow cyclically.
template <int D>
__device__ __forceinline__ void flux(const double* u, const double* w, double* f) {
    constexpr int V1 = 1 + D, V2 = 1 + (D + 1) % 3, V3 = 1 + (D + 2) % 3;
    f[0] = u[0] * w[V1];
    f[V1] = u[V1] * w[V1] + w[4];
    f[V2] = u[V2] * w[V1];
    f[V3] = u[V3] * w[V1];
    f[4] = (u[4] + w[4]) * w[V1];
}

// Physical flux of a conservative state (src/hydro.cpp:179-192).
template <int D>
__device__ __forceinline__ void state_flux(const double* u, double* f, double gm) {
    double w[NVP];
    prim(u, w, gm);
    flux<D>(u, w, f);
}

// Local Lax-Friedrichs flux (src/hydro.cpp:270-289) for the physical
// variables; the slot flux is -0.5*c*(0-0) = +0 and is not carried.
template <int D>
__device__ __forceinline__ void riemann_llf(const double* uL, const double* uR, double* F, double gm) {
    constexpr int V1 = 1 + D;
    double wL[NVP], wR[NVP], fL[NVP], fR[NVP];
    prim(uL, wL, gm);
    prim(uR, wR, gm);
    flux<D>(uL, wL, fL);
    flux<D>(uR, wR, fR);
    double cL = sound_speed(wL[0], wL[4], gm) + fabs(wL[V1]);
    double cR = sound_speed(wR[0], wR[4], gm) + fabs(wR[V1]);
    double c = smax(cL, cR);
#pragma unroll
    for (int v = 0; v < NVP; v++) F[v] = 0.5 * (fR[v] + fL[v]) - 0.5 * c * (uR[v] - uL[v]);
}

// HLLC flux (src/hydro.cpp:291-363); the slot entry is zero.
template <int D>
__device__ __forceinline__ void riemann_hllc(const double* UL, const double* UR, double* F, double gm) {
    constexpr int V1 = 1 + D, V2 = 1 + (D + 1) % 3, V3 = 1 + (D + 2) % 3;
    double WL[NVP], WR[NVP];
    prim(UL, WL, gm);
    prim(UR, WR, gm);
    double cL = sound_speed(WL[0], WL[4], gm) + fabs(WL[V1]);
    double cR = sound_speed(WR[0], WR[4], gm) + fabs(WR[V1]);
    double c = smax(cL, cR);
    double vL = WL[V1], vR = WR[V1];
    double sL = smin(vL, vR) - c;
    double sR = smax(vL, vR) + c;
    double rcL = WL[0] * (vL - sL);
    double rcR = WR[0] * (sR - vR);
    double vs = (rcR * vR + rcL * vL + (WL[4] - WR[4])) / (rcR + rcL);
    double ps = (rcR * WL[4] + rcL * WR[4] + rcL * rcR * (vL - vR)) / (rcR + rcL);
    double rsL = WL[0] * (sL - vL) / (sL - vs);
    double esL = ((sL - vL) * UL[4] - WL[4] * vL + ps * vs) / (sL - vs);
    double rsR = WR[0] * (sR - vR) / (sR - vs);
    double esR = ((sR - vR) * UR[4] - WR[4] * vR + ps * vs) / (sR - vs);
    double rg, vg, pg, eg;
    if (sL > 0) { rg = WL[0]; vg = WL[V1]; pg = WL[4]; eg = UL[4]; }
    else if (vs > 0) { rg = rsL; vg = vs; pg = ps; eg = esL; }
    else if (sR > 0) { rg = rsR; vg = vs; pg = ps; eg = esR; }
    else { rg = WR[0]; vg = WR[V1]; pg = WR[4]; eg = UR[4]; }
    F[0] = rg * vg;
    F[V1] = rg * vg * vg + pg;
    F[V2] = rg * vg * (vs > 0 ? WL[V2] : WR[V2]);
    F[V3] = rg * vg * (vs > 0 ? WL[V3] : WR[V3]);
    F[4] = vg * (eg + pg);
}

}  // namespace spdk
