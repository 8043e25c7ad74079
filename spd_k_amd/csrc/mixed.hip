// Mixed-precision element-tile path for the hydro ADER step.
//
// Same sweep structure as the FP64 tile path (one kernel per Picard sweep
// over element tiles; only face traces cross elements, through HBM), with
// the state carried as FP32 deviations from each element's FP64 mean:
//
//     U(t, v, sp) = mu(v) + d(t, v, sp),   mu(v) = mean of U_sp(v) over the element
//
//   FP32: the space-time state d and its face traces, the sp->fp
//         interpolation, the flux-divergence contraction (applied to
//         F - F(mu) with dfp_to_sp / h premultiplied) and the predictor's
//         time contraction (dt * invader * w premultiplied);
//   FP64: the element means, every flux and Riemann evaluation (pressure,
//         sound speed, positivity), the corrector accumulation of U, the
//         cv transform of the new state and dt.
//
// Working on deviations keeps the FP32 rounding relative to the variation
// inside an element, not to the size of U; subtracting F(mu) before the
// divergence keeps it relative to the flux variation.
//
// Conservation. In exact arithmetic the quadrature of the flux divergence
// over an element equals its face-flux balance,
//     sum_sp w_sp dudt_sp = sum_d sum_q w'_q (F_q(1) - F_q(0)) / h_d,
// and the face fluxes are single-valued, so totals telescope. The FP32
// divergence breaks the identity at FP32 rounding. The corrector restores it:
// the balance I is accumulated in FP64 from the FP64 face fluxes (each face's
// flux is evaluated from the same inputs, with the same operations, by both
// elements that share it) and the update du of every point is shifted by
// (I - J) / sum w, J = sum_sp w_sp du_sp, so each element's total changes by
// exactly its face balance up to FP64 rounding. w are the quadrature
// weights of the solution points on [0, 1].
//
// A sweep maps work to threads in two ways. The first sweep from the
// zero-order guess (U_ader = U for every time node, the reference's
// copy_ader) evaluates one time slice, all axes at once; every later sweep
// evaluates all T slices one axis at a time, with exactly one pencil
// (slice, line of points along the axis) per thread (T * N^(d-1) = N^d).
//
// Picard control: the fixed count (p predictor sweeps, then the corrector,
// src/hydro_ader.hpp:291-305) or a residual test with the element residual
//     r_e = max |d_out - d_in| / max_v |mu(v)|   (ghost elements excluded).
// Residual form: the first sweep updates every element; later sweeps update
// only the active elements: those whose own r_e, or the r_e of an element
// whose traces they read, was >= tol in the previous sweep; the others keep
// their state and traces. Each element stops at convergence or at the cap
// (by default the fixed count, p sweeps: an element that never converges
// ends with exactly the fixed-count predictor from the zero-order guess,
// evaluated against its neighbours' latest traces). A corrector sweep over
// every element ends the step (its face fluxes come from one set of traces,
// so conservation holds). From an external or learned guess, the elements
// still unconverged at the cap restart from the zero-order guess and get
// the fixed count again (k_mrestart), the fallback of that form; or, with
// guess_keep, they keep the state of the cap as from the zero-order guess.
//
// Trace buffers. Each element writes its traces to the older of two buffers
// and records the sweep number (ver); readers take, per neighbour, the
// newest buffer written before the current sweep. A buffer being written is
// therefore never read, whether or not the element was active before.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include "ctx.hpp"
#include "physics.hpp"

#pragma clang fp contract(fast)

using namespace spdk;

namespace {

constexpr int kMaxTm = 8;
constexpr int kHistBins = 32;
constexpr int kGS = 12;          // per-element floats of the learned-guess scales
constexpr int kGExtra = 6;       // extra features per element
constexpr float kAmin = 1e-6f;   // floor of the element amplitude
constexpr float kAgate = 1e-5f;  // amplitude at or below which the guess is the zero-order one

// Minimum waves per SIMD the sweep kernel is compiled for (bounds its VGPRs);
// the 3-D p=3 predictor is held at 3 (168 VGPRs), where it fits without spills.
#ifndef SPDK_MSWEEP_WAVES
#define SPDK_MSWEEP_WAVES 1
#endif

template <int N, int DIM>
struct MT {
    static constexpr int T = N;  // ADER time nodes
    static constexpr int M = N + 1;
    static constexpr int NT = DIM == 1 ? 1 : (DIM == 2 ? N : N * N);
    static constexpr int TPE = N * NT;
    static constexpr int FSZ = M * NT;
    static constexpr int SLOTS = T > DIM ? T : DIM;
    static constexpr int ROW = N + 1;                // padded x-row of an LDS state plane
    static constexpr int PSZ = DIM == 1 ? N : TPE / N * ROW;  // padded state plane
    static constexpr int LDS_D = T * NVP * PSZ;      // floats: state slices
    static constexpr int LDS_F = SLOTS * NVP * FSZ;  // floats: flux slots
    static constexpr int LDS_X = 6 * NVP + 2;        // doubles: mu, F(mu) per axis, residual, shift, mean state
    // + 32 bytes: the element areas of a block start 8 banks apart, so reads
    // that cross elements at one in-element offset do not share a bank
    static constexpr int ELEM_BYTES = ((LDS_X * 8 + (LDS_D + LDS_F) * 4) + 15) / 16 * 16 + 32;
    static constexpr int EPB0 = TPE >= 256 ? 1 : 256 / TPE;
    static constexpr int EPB_LDS = 61440 / ELEM_BYTES;
    static constexpr int EPB = EPB0 < EPB_LDS ? EPB0 : EPB_LDS;
    static constexpr int BLOCK = ((EPB * TPE + 63) / 64) * 64;
    static constexpr int LDS_BYTES = EPB * ELEM_BYTES > BLOCK * 8 ? EPB * ELEM_BYTES : BLOCK * 8;
    static_assert(EPB >= 1 && BLOCK <= 256, "tile does not fit a workgroup");
    static_assert((LDS_D + LDS_F) * 4 >= 2 * NVP * TPE * 8, "cv ping-pong needs two FP64 state planes");
    static constexpr int XRES = 4 * NVP;      // X slot: element residual
    static constexpr int XACT = 4 * NVP + 1;  // X slot: element active this sweep
    static constexpr int XSHIFT = 4 * NVP + 2; // X slots: conservative shift per variable
    static constexpr int XMF = 4 * NVP + 2 + NVP + 1;  // 4 X slots = 8 floats: rho, v[3], m[3], H of mu
    static_assert(T <= kMaxTm, "time nodes");
};

struct MArgs {
    const double* U_sp;     // (1,6,E,TPE) solution at the step start (FP64)
    double* U_new;          // corrector: the new solution (U_sp itself)
    const double* mu;       // (5,E) element means of U_sp
    const float* d_old;     // (5,E,TPE) U_sp - mu
    float* D;               // (T,5,E,TPE) predictor state, read and rewritten in place
    float* trb[2][3];       // two trace buffers per axis, (T,5,E,2,NT) (slice 0 only for the zero-order guess)
    int* ver;               // (2,E): sweep number of each element's last write to each buffer
    int cur;                // this sweep's number
    const unsigned char* big_in;  // per element: r_e >= tol in the previous sweep (null: every element active)
    const int* blist;       // blocks to run (grid = its length; null: every block)
    unsigned char* big_out;       // per element, this sweep (null: not recorded)
    double* cnt_part;       // per-block count of active elements (predictor)
    double* W_cv;           // optional (1,6,E,TPE) output primitives at control volumes
    double* F_x;            // optional (T,6,E,NT,M) x flux-point fluxes
    double* part;           // per-block dt minima (corrector)
    double* res_part;       // per-block residual maxima (predictor)
    unsigned char* conv;    // optional per-element first sweep with r_e < tol
    const float* S2F;       // (M x N) FP32 sp_to_fp
    const float* Dh;        // (3 x N x M) FP32 dfp_to_sp / h_d
    const double* S2C;      // sp_to_cv (N x N) FP64
    double wq[kMaxTm];      // quadrature weights of the solution points on [0, 1]
    double wtot;            // sum of the element's point weights
    double ih[3];           // 1 / h_d
    float cdt[kMaxTm * kMaxTm];  // dt * invader(t,l) * w(l)
    double wdt[kMaxTm];          // dt * w(t)
    int Ne[3], act[3], bc[3];
    int E;
    double h[3];
    double gm, cfl;
    double tol;
    int sweep;
    double g[3];    // constant gravity (hydro/g1..g3)
    int grav;       // any component nonzero
    int path;       // 0: every slice is d_old (zero-order guess); 1: slices from D
    int tin_slices; // slices present in the readable trace buffers (1 or T)
    int out_pred;   // predictor (else corrector)
    // FV face-flux sweep: sub-cell face fluxes of slice t along axis d (FV
    // layout, FP64), the FV cell counts (ghost layers included) and ghost
    // layers per axis, S x S of sp_to_cv in FP32 (row o1*N + o2, column
    // ll*N + nn), and a switch to the VALU contraction
    double* Ffv[kMaxTm][3];
    int fvn[3], fgh[3];
    const float* SS;
    int fv_valu;
};

// ---- FP64 physics in the frame of the face normal: slot 1 is the normal
// momentum, slots 2 and 3 the transverse ones (cyclic, as V1, V2, V3 of
// src/hydro.cpp:156-162). One code path serves every axis; the caller maps
// slots to variables with vperm.
__device__ __forceinline__ int vperm(int v, int d) {
    if (v == 1) return d == 0 ? 1 : (d == 1 ? 2 : 3);
    if (v == 2) return d == 0 ? 2 : (d == 1 ? 3 : 1);
    if (v == 3) return d == 0 ? 3 : (d == 1 ? 1 : 2);
    return v;
}

// refined hardware reciprocal (one Newton step; not correctly rounded)
__device__ __forceinline__ double frcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
}

__device__ __forceinline__ void mflux(const double* u, double* f, double gm1) {
    double ri = frcp(u[0]);
    double p = (u[4] - 0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) * ri) * gm1;
    double vn = u[1] * ri;
    f[0] = u[1];
    f[1] = u[1] * vn + p;
    f[2] = u[2] * vn;
    f[3] = u[3] * vn;
    f[4] = (u[4] + p) * vn;
}

// F(mu + d) - F(mu) in FP32 for a deviation d (normal frame, slot 1 the
// normal momentum), written in differences so that every term is of the
// order of d: with m = mu_m + d_m, v = m / rho,
//   dv_i = (d_m_i - v_i(mu) d_rho) / rho,   d(m_i v_j) = m_i(mu) dv_j + d_m_i v_j,
//   dp = (gamma-1) (d_E - 1/2 sum_i d(m_i v_i)),  d(H v_n) = H(mu) dv_n + (d_E + dp) v_n.
// The FP32 rounding is therefore relative to the flux variation, as for the
// FP64 evaluation rounded to FP32 it replaces. c: rho, v_n, v_t1, v_t2,
// m_n, m_t1, m_t2, H = E + p of the element mean, in the frame.
__device__ __forceinline__ void mdflux(const float* d, const float* c, float* df, float gm1) {
    const float ri = __builtin_amdgcn_rcpf(c[0] + d[0]);
    const float dvn = (d[1] - c[1] * d[0]) * ri;
    const float dvt1 = (d[2] - c[2] * d[0]) * ri;
    const float dvt2 = (d[3] - c[3] * d[0]) * ri;
    const float vn = c[1] + dvn, vt1 = c[2] + dvt1, vt2 = c[3] + dvt2;
    const float dk = c[4] * dvn + d[1] * vn + c[5] * dvt1 + d[2] * vt1 + c[6] * dvt2 + d[3] * vt2;
    const float dp = gm1 * (d[4] - 0.5f * dk);
    df[0] = d[1];
    df[1] = c[4] * dvn + d[1] * vn + dp;
    df[2] = c[5] * dvn + d[2] * vn;
    df[3] = c[6] * dvn + d[3] * vn;
    df[4] = c[7] * dvn + (d[4] + dp) * vn;
}

// LLF (src/hydro.cpp:270-289): fluxes and the jump term in FP64; the wave
// speed, a dissipation coefficient evaluated identically on both sides of a
// face, from an FP32 square root.
__device__ __forceinline__ void mllf(const double* uL, const double* uR, double* F, double gm) {
    const double gm1 = gm - 1.0;
    double riL = frcp(uL[0]), riR = frcp(uR[0]);
    double pL = (uL[4] - 0.5 * (uL[1] * uL[1] + uL[2] * uL[2] + uL[3] * uL[3]) * riL) * gm1;
    double pR = (uR[4] - 0.5 * (uR[1] * uR[1] + uR[2] * uR[2] + uR[3] * uR[3]) * riR) * gm1;
    double vL = uL[1] * riL, vR = uR[1] * riR;
    float c2L = fmaxf((float)(gm * pL * riL), (float)kMinC2);
    float c2R = fmaxf((float)(gm * pR * riR), (float)kMinC2);
    double cL = (double)__builtin_sqrtf(c2L) + fabs(vL);
    double cR = (double)__builtin_sqrtf(c2R) + fabs(vR);
    double hc = 0.5 * fmax(cL, cR);
    double fL[NVP], fR[NVP];
    fL[0] = uL[1];
    fR[0] = uR[1];
    fL[1] = uL[1] * vL + pL;
    fR[1] = uR[1] * vR + pR;
    fL[2] = uL[2] * vL;
    fR[2] = uR[2] * vR;
    fL[3] = uL[3] * vL;
    fR[3] = uR[3] * vR;
    fL[4] = (uL[4] + pL) * vL;
    fR[4] = (uR[4] + pR) * vR;
#pragma unroll
    for (int v = 0; v < NVP; v++) F[v] = 0.5 * (fR[v] + fL[v]) - hc * (uR[v] - uL[v]);
}

// x[i] for a runtime i < K without indexing registers or kernel arguments
template <typename V, int K>
__device__ __forceinline__ V pick(const V (&x)[K], int i) {
    V r = x[0];
#pragma unroll
    for (int k = 1; k < K; k++) r = i == k ? x[k] : r;
    return r;
}

// LDS index of flux point ll of pencil q along axis d (x: q*M+ll, the
// reference's fp-x layout; y: (kk, ll, ii); z: (ll, jj, ii)).
template <int N, int DIM>
__device__ __forceinline__ int fidx(int d, int ll, int q) {
    constexpr int M = N + 1, NT = MT<N, DIM>::NT;
    if (d == 0) return q * M + ll;
    if (d == 1) return ((q / N) * M + ll) * N + (q % N);
    return ll * NT + q;
}
// Solution-point index of position ll along axis d of pencil q.
template <int N, int DIM>
__device__ __forceinline__ int spidx(int d, int ll, int q) {
    if (DIM == 1) return ll;
    if (DIM == 2) return d == 0 ? q * N + ll : ll * N + q;
    if (d == 0) return q * N + ll;
    if (d == 1) return ((q / N) * N + ll) * N + (q % N);
    return ll * N * N + q;
}

// LDS state planes pad every x-row of N points to N+1 floats, so the pencil
// reads along x, y and z (one pencil per lane) fall in distinct banks.
template <int N, int DIM>
__device__ __forceinline__ int lpad(int s) { return DIM == 1 ? s : s + s / N; }

// Pencil q along axis d: point l at state offset sb + l*ss (padded plane),
// flux point f at slot offset fb + f*fs (the fidx layout). Selects only: d
// may differ between lanes.
struct Geo { int sb, ss, fb, fs; };
template <int N, int DIM>
__device__ __forceinline__ Geo geo(int d, int q) {
    constexpr int M = N + 1, R = N + 1, NT = MT<N, DIM>::NT;
    Geo g;
    if (DIM == 1) {
        g.sb = 0; g.ss = 1; g.fb = 0; g.fs = 1;
    } else if (DIM == 2) {
        g.sb = d == 0 ? q * R : q;
        g.ss = d == 0 ? 1 : R;
        g.fb = d == 0 ? q * M : q;
        g.fs = d == 0 ? 1 : N;
    } else {
        const int qh = q / N, ql = q % N;
        g.sb = d == 0 ? q * R : (d == 1 ? qh * N * R + ql : q + qh);
        g.ss = d == 0 ? 1 : (d == 1 ? R : N * R);
        g.fb = d == 0 ? q * M : (d == 1 ? qh * M * N + ql : q);
        g.fs = d == 0 ? 1 : (d == 1 ? N : NT);
    }
    return g;
}

// Trace buffer of axis d holding element el's newest traces written before
// this sweep.
__device__ __forceinline__ const float* read_buffer(const MArgs& a, int d, size_t el) {
    const int v0 = a.ver[el], v1 = a.ver[(size_t)a.E + el];
    const int b = (v1 < a.cur && (v1 > v0 || v0 >= a.cur)) ? 1 : 0;
    const float* t0 = d == 0 ? a.trb[0][0] : (d == 1 ? a.trb[0][1] : a.trb[0][2]);
    const float* t1 = d == 0 ? a.trb[1][0] : (d == 1 ? a.trb[1][1] : a.trb[1][2]);
    return b ? t1 : t0;
}

// Elements whose traces element e reads across its two faces along axis d
// (ghost fills: periodic copies the opposite interior element, gradfree /
// reflective the adjacent one; src/boundary.cpp:48-113).
__device__ __forceinline__ void face_sources(const MArgs& a, int d, int ei, int ej, int ek, size_t e, size_t& el0,
                                             size_t& el1) {
    const int ed = d == 0 ? ei : (d == 1 ? ej : ek);
    const int Nd = d == 0 ? a.Ne[0] : (d == 1 ? a.Ne[1] : a.Ne[2]);
    const bool per = (d == 0 ? a.bc[0] : (d == 1 ? a.bc[1] : a.bc[2])) == SPDK_BC_PERIODIC;
    const size_t stride = d == 0 ? 1 : (d == 1 ? (size_t)a.Ne[0] : (size_t)a.Ne[0] * a.Ne[1]);
    const size_t ebase = e - (size_t)ed * stride;
    el0 = ebase + (size_t)(ed <= 1 ? (per ? Nd - 2 : 1) : ed - 1) * stride;
    el1 = ebase + (size_t)(ed >= Nd - 2 ? (per ? 1 : Nd - 2) : ed + 1) * stride;
}

// One pencil task: slice Dsl of the LDS state (floats [v][PSZ]) along axis
// d, pencil q. Writes F - F(mu) at the M flux points into the flux slot
// (floats [v][FSZ]) and, when asked, adds bw * (F(1) - F(0)) of the
// pencil to the face balance (want_bal). Face values follow src/boundary.cpp:48-113
// (ghost traces) and src/hydro.cpp:368-418 (LLF at element faces).
template <int N, int DIM>
__device__ __forceinline__ void pencil(const MArgs& a, const float* Dsl, float* Fslot, const double* X, int d, int q,
                                       int ts, int ei, int ej, int ek, size_t e, bool want_bal, double (&bal)[NVP],
                                       double bw, bool lower_face_only = false) {
    using G = MT<N, DIM>;
    constexpr int M = G::M, NT = G::NT;
    const size_t E = (size_t)a.E;
    const double gm1 = a.gm - 1.0;
    const Geo g = geo<N, DIM>(d, q);
    int pv[NVP];
#pragma unroll
    for (int v = 0; v < NVP; v++) pv[v] = vperm(v, d);
    // neighbour operands: n0 is the value left of face 0 (or the ghost fill
    // that replaces the own value at face M-1 of the left ghost), n1 the value
    // right of face M-1 (or the ghost fill at face 0 of the right ghost).
    // Ghost fills: periodic copies the opposite interior element, gradfree /
    // reflective the adjacent one (normal momentum negated). Issued first so
    // the loads overlap the interpolation and the interior fluxes.
    const int ed = d == 0 ? ei : (d == 1 ? ej : ek);
    const int Nd = d == 0 ? a.Ne[0] : (d == 1 ? a.Ne[1] : a.Ne[2]);
    const int type = d == 0 ? a.bc[0] : (d == 1 ? a.bc[1] : a.bc[2]);
    const size_t stride = d == 0 ? 1 : (d == 1 ? (size_t)a.Ne[0] : (size_t)a.Ne[0] * a.Ne[1]);
    const size_t ebase = e - (size_t)ed * stride;
    const bool per = type == SPDK_BC_PERIODIC;
    const size_t el0 = ebase + (size_t)(ed <= 1 ? (per ? Nd - 2 : 1) : ed - 1) * stride;
    const int sid0 = ed <= 1 ? (per ? 1 : 0) : 1;
    const size_t el1 = ebase + (size_t)(ed >= Nd - 2 ? (per ? 1 : Nd - 2) : ed + 1) * stride;
    const int sid1 = ed >= Nd - 2 ? (per ? 0 : 1) : 0;
    const bool flip0 = ed <= 1 && type == SPDK_BC_REFLECTIVE;
    const bool flip1 = ed >= Nd - 2 && type == SPDK_BC_REFLECTIVE;
    const float* tr0 = read_buffer(a, d, el0);
    const float* tr1 = read_buffer(a, d, el1);
    double n0[NVP], n1[NVP];
#pragma unroll
    for (int v = 0; v < NVP; v++) {
        const size_t pvE = (size_t)pv[v] * E;
        const size_t tb = ((size_t)ts * NVP + pv[v]) * E;
        n0[v] = a.mu[pvE + el0] + (double)tr0[((tb + el0) * 2 + sid0) * NT + q];
        n1[v] = a.mu[pvE + el1] + (double)tr1[((tb + el1) * 2 + sid1) * NT + q];
    }
    if (flip0) n0[1] = -n0[1];
    if (flip1) n1[1] = -n1[1];
    double mu[NVP], Fb[NVP];
#pragma unroll
    for (int v = 0; v < NVP; v++) {
        mu[v] = X[pv[v]];
        Fb[v] = X[NVP + d * NVP + pv[v]];
    }
    float uf[M][NVP];
#pragma unroll
    for (int v = 0; v < NVP; v++) {
        float p[N];
#pragma unroll
        for (int l = 0; l < N; l++) p[l] = Dsl[pv[v] * G::PSZ + g.sb + l * g.ss];
#pragma unroll
        for (int f = 0; f < M; f++) {
            float x = 0.f;
#pragma unroll
            for (int l = 0; l < N; l++) x += p[l] * a.S2F[f * N + l];
            uf[f][v] = x;
        }
    }
    // interior flux points: F(mu + d) - F(mu) in FP32 difference form
    {
        const float* XF = (const float*)(X + G::XMF);
        float cm[8];
        cm[0] = XF[0];
#pragma unroll
        for (int i = 1; i < 4; i++) {
            cm[i] = XF[pv[i]];
            cm[3 + i] = XF[3 + pv[i]];
        }
        cm[7] = XF[7];
#pragma unroll
        for (int f = 1; f < M - 1; f++) {
            float df[NVP];
            mdflux(uf[f], cm, df, (float)gm1);
#pragma unroll
            for (int v = 0; v < NVP; v++) Fslot[pv[v] * G::FSZ + g.fb + f * g.fs] = df[v];
        }
    }
    // faces: LLF at every face; at the outer face of a ghost element
    // (no neighbour) both states are the element's own, which gives its
    // physical flux exactly
#pragma unroll
    for (int side = 0; side < 2; side++) {
        if (side == 1 && lower_face_only) break;
        const int f = side ? M - 1 : 0;
        const bool raw = side ? (ed == Nd - 1) : (ed == 0);
        double own[NVP], fl[NVP], uL[NVP], uR[NVP];
#pragma unroll
        for (int v = 0; v < NVP; v++) {
            own[v] = mu[v] + (double)uf[f][v];
            if (side == 0) {
                uL[v] = raw ? own[v] : n0[v];
                uR[v] = ed == Nd - 1 ? n1[v] : own[v];
            } else {
                uL[v] = ed == 0 ? n0[v] : own[v];
                uR[v] = raw ? own[v] : n1[v];
            }
        }
        mllf(uL, uR, fl, a.gm);
#pragma unroll
        for (int v = 0; v < NVP; v++) Fslot[pv[v] * G::FSZ + g.fb + f * g.fs] = (float)(fl[v] - Fb[v]);
        if (want_bal) {
            // bal is indexed by variable: un-permute the momentum slots
            double c[NVP];
#pragma unroll
            for (int v = 0; v < NVP; v++) c[v] = side ? bw * fl[v] : -(bw * fl[v]);
            bal[0] += c[0];
            bal[4] += c[4];
            bal[1] += d == 0 ? c[1] : (d == 1 ? c[3] : c[2]);
            bal[2] += d == 0 ? c[2] : (d == 1 ? c[1] : c[3]);
            bal[3] += d == 0 ? c[3] : (d == 1 ? c[2] : c[1]);
        }
    }
}

// Transverse quadrature weight of pencil q (product of the 1-D weights of
// its transverse coordinates; the same formula for every axis).
template <int N, int DIM>
__device__ __forceinline__ double pencil_weight(const MArgs& a, int q) {
    if (DIM == 1) return 1.0;
    if (DIM == 2) return pick(a.wq, q);
    return pick(a.wq, q / N) * pick(a.wq, q % N);
}

// New state at the control volumes and the element's dt candidate (FP64):
// compute_primitives, sp->cv sweeps x, y, z (src/transforms.cpp:99-117) and
// src/hydro.cpp:226-256. `area` is the element's LDS (>= 2 planes of FP64).
template <int N, int DIM>
__device__ __forceinline__ double cv_and_dt(const double* un, double* area, bool live, int s, int ii, int jj, int kk,
                                            int ei, int ej, int ek, size_t e, double* W_cv, size_t slabE,
                                            const double* S2C, const int* Ne, const int* act, const double* h,
                                            double gm, double cfl) {
    using G = MT<N, DIM>;
    constexpr int TPE = G::TPE;
    double* wa = area;
    double* wb = area + NVP * TPE;
    if (live) {
        double wv[NVP];
        prim(un, wv, gm);
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
                    x += src[v * TPE + sl] * S2C[id * N + l];
                }
                dst[v * TPE + s] = x;
            }
        }
        __syncthreads();
        double* tmp = src;
        src = dst;
        dst = tmp;
    }
    double r = __builtin_inf();
    if (live) {
        double wc[NVP];
#pragma unroll
        for (int v = 0; v < NVP; v++) wc[v] = src[v * TPE + s];
        if (W_cv) {
#pragma unroll
            for (int v = 0; v < NVP; v++) W_cv[(size_t)v * slabE + e * TPE + s] = wc[v];
            W_cv[(size_t)NVP * slabE + e * TPE + s] = 0.0;
        }
        bool active = (!act[0] || (ei >= 1 && ei <= Ne[0] - 2)) && (!act[1] || (ej >= 1 && ej <= Ne[1] - 2)) &&
                      (!act[2] || (ek >= 1 && ek <= Ne[2] - 2));
        if (active) {
            double c_s = sound_speed(wc[0], wc[4], gm);
            double c_max = 0, dx_min = 1;
            if (act[0]) { c_max += (fabs(wc[1]) + c_s); dx_min = smin(dx_min, h[0]); }
            if (act[1]) { c_max += (fabs(wc[2]) + c_s); dx_min = smin(dx_min, h[1]); }
            if (act[2]) { c_max += (fabs(wc[3]) + c_s); dx_min = smin(dx_min, h[2]); }
            r = cfl * dx_min / c_max / N;
        }
    }
    return r;
}

// Block minimum of r into part[blockIdx.x]; `red` holds BLOCK doubles.
template <int BLOCK>
__device__ __forceinline__ void block_min(double r, double* red, double* part) {
    __syncthreads();
    red[threadIdx.x] = r;
    __syncthreads();
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = nan_min(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Conservative shift of the point updates du (see the header): red holds
// 2 * NVP * TPE doubles of the element's LDS, X_shift NVP doubles.
template <int N, int DIM>
__device__ __forceinline__ void cons_shift(double* du, const double* bal, double wsp, double wtot, double* red,
                                           double* X_shift, bool live, int s) {
    constexpr int TPE = MT<N, DIM>::TPE;
    if (live)
#pragma unroll
        for (int v = 0; v < NVP; v++) {
            red[v * TPE + s] = bal[v];
            red[(NVP + v) * TPE + s] = wsp * du[v];
        }
    __syncthreads();
    if (live)
        for (int v = s; v < NVP; v += TPE) {  // an element may have fewer points than variables
            double I = 0.0, J = 0.0;
            for (int j = 0; j < TPE; j++) {
                I += red[v * TPE + j];
                J += red[(NVP + v) * TPE + j];
            }
            X_shift[v] = (I - J) / wtot;
        }
    __syncthreads();
#pragma unroll
    for (int v = 0; v < NVP; v++) du[v] += X_shift[v];
}

// face_integral (src/transforms.cpp:378-481) of the round's axis d for every
// slice, from the flux deviations F - F(mu) in the flux slots (slot t =
// slice t): each element writes the sub-cell faces of its cells' lower sides
// along d (the right ghost along d the domain's last face), so every FV face
// is written once and the FV update built on them conserves exactly. The
// transverse contraction runs in FP32 on the deviations and F(mu) is added
// in FP64. 3-D p=3: on MFMA, one wave per element (v_mfma_f32_16x16x4_f32:
// A = S x S, 16 output pairs (o1, o2) by 16 transverse points (ll, nn);
// B = 16 panels (slice, variable, face) per tile, 5 tiles); otherwise (or
// with fv_valu) on the VALU, one output position per thread.
template <int N, int DIM>
__device__ __forceinline__ void mfv_faces(const MArgs& a, const float* Fs, const double* X, int d, int s, bool live,
                                          int ei, int ej, int ek) {
    using G = MT<N, DIM>;
    constexpr int M = G::M, FSZ = G::FSZ, T = G::T, TPE = G::TPE;
    const int ec[3] = {ei, ej, ek};
    const int Nd = d == 0 ? a.Ne[0] : (d == 1 ? a.Ne[1] : a.Ne[2]);
    // Along a periodic axis the domain's first and last faces are one face:
    // the first active element writes both copies, so they hold the same
    // bits (each element adds its own FP32-rounded deviations to its own mean
    // flux) and the FV update's face sums telescope exactly.
    const bool per = (d == 0 ? a.bc[0] : (d == 1 ? a.bc[1] : a.bc[2])) == SPDK_BC_PERIODIC;
    bool own = live;
#pragma unroll
    for (int x = 0; x < DIM; x++)
        own = own && ec[x] >= 1 && ec[x] <= (x == d ? (per ? a.Ne[x] - 2 : a.Ne[x] - 1) : a.Ne[x] - 2);
    const bool last = ec[d] == Nd - 1;
    const bool wrap = per && ec[d] == 1;  // also writes the last face
    // A reflective wall's face (face 0 of the first interior element or of
    // the upper ghost): the FP64 LLF of the mirrored pair is exactly zero in
    // every component but the normal momentum, and is written so (the mean
    // flux plus its FP32-rounded deviation would let mass and energy through
    // at that rounding).
    const bool refl = (d == 0 ? a.bc[0] : (d == 1 ? a.bc[1] : a.bc[2])) == SPDK_BC_REFLECTIVE;
    const bool wall_el = refl && (ec[d] == 1 || last);
    const int dim1 = d == 0 ? 1 : (d == 1 ? 2 : 0), dim2 = dim1 == 0 ? 1 : (dim1 == 1 ? 2 : 0);
    const int nx = a.fvn[0] + (d == 0), ny = a.fvn[1] + (d == 1), nz = a.fvn[2] + (d == 2);
    const size_t vstride = (size_t)nz * ny * nx;
    // offset of output position (ox, oy, oz) of this element, or (wrapped) of
    // the element at the end of the axis
    auto fv_off = [&](int ox, int oy, int oz, bool wrapped = false) -> size_t {
        const int ex = wrapped && d == 0 ? a.Ne[0] - 1 : ei, ey = wrapped && d == 1 ? a.Ne[1] - 1 : ej;
        const int ez = wrapped && d == 2 ? a.Ne[2] - 1 : ek;
        const int I = ox + a.fgh[0] + (ex - 1) * N;
        const int J = DIM >= 2 ? oy + a.fgh[1] + (ey - 1) * N : 0;
        const int K = DIM == 3 ? oz + a.fgh[2] + (ez - 1) * N : 0;
        return ((size_t)K * ny + J) * nx + I;
    };
    // pencil of axis d through transverse points ll (along dim1) and nn (along dim2)
    auto pq = [&](int ll, int nn) -> int { return DIM == 3 ? (d == 1 ? ll * N + nn : nn * N + ll) : (DIM == 2 ? (d == 0 ? ll : nn) : 0); };
    if (N == 4 && DIM == 3 && TPE == 64 && !a.fv_valu) {
        if (!own) return;  // one wave per element: the wave leaves together
        const int lane = s;  // = the lane in the element's wave
        const int li = lane & 15, lk = lane >> 4;
        float av[4];
#pragma unroll
        for (int ks = 0; ks < 4; ks++) av[ks] = a.SS[li * 16 + 4 * ks + lk];  // A[row li][k = 4 ks + lk]
        double* base[T];
#pragma unroll
        for (int t = 0; t < T; t++) base[t] = d == 0 ? a.Ffv[t][0] : (d == 1 ? a.Ffv[t][1] : a.Ffv[t][2]);
        constexpr int NCOL = T * NVP * N;  // panels (t, v, f), f fastest
#pragma unroll
        for (int ct = 0; ct < NCOL / 16; ct++) {
            const int col = ct * 16 + li;
            const int t = col / (NVP * N), v = (col / N) % NVP, f = col % N;
            const float* Fp = Fs + (t * NVP + v) * FSZ;
            typedef float f4 __attribute__((ext_vector_type(4)));
            f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < 4; ks++) {
                const int k = 4 * ks + lk, ll = k / N, nn = k % N;  // B[k][col]
                const float b = Fp[fidx<N, DIM>(d, f, pq(ll, nn))];
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av[ks], b, acc, 0, 0, 0);
            }
            if (!last || f == 0) {
                const double fm = X[NVP + d * NVP + v];
                double* ob = (t == 0 ? base[0] : (t == 1 ? base[1] : (t == 2 ? base[2] : base[3]))) + (size_t)v * vstride;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int row = 4 * lk + i, o1 = row / N, o2 = row % N;  // D[row][col]
                    int o[3];
                    o[d] = f;
                    o[dim1] = o1;
                    o[dim2] = o2;
                    const double val = wall_el && f == 0 && v != 1 + d ? 0.0 : fm + (double)acc[i];
                    ob[fv_off(o[0], o[1], o[2])] = val;
                    if (wrap && f == 0) ob[fv_off(o[0], o[1], o[2], true)] = val;
                }
            }
        }
        return;
    }
    // VALU: thread s -> output position (x fastest) of every slice and variable
    if (!own || s >= TPE) return;
    const int o[3] = {s % N, DIM >= 2 ? (s / N) % N : 0, DIM == 3 ? s / (N * N) : 0};
    const int f = o[d];
    if (last && f != 0) return;
    const int o1 = o[dim1], o2 = o[dim2];
    float w[N][N];  // S(o1, ll) S(o2, nn) of the active transverse axes
#pragma unroll
    for (int ll = 0; ll < N; ll++)
#pragma unroll
        for (int nn = 0; nn < N; nn++) {
            const float s1 = a.S2C ? (float)a.S2C[o1 * N + ll] : 0.f, s2 = (float)a.S2C[o2 * N + nn];
            w[ll][nn] = DIM == 3 ? s1 * s2 : 0.f;
        }
    const size_t off = fv_off(o[0], o[1], o[2]);
    const bool also_last = wrap && f == 0;
    const size_t off_last = also_last ? fv_off(o[0], o[1], o[2], true) : 0;
#pragma unroll
    for (int t = 0; t < T; t++) {
        double* ob = d == 0 ? a.Ffv[t][0] : (d == 1 ? a.Ffv[t][1] : a.Ffv[t][2]);
#pragma unroll
        for (int v = 0; v < NVP; v++) {
            const float* Fp = Fs + (t * NVP + v) * FSZ;
            float acc = 0.f;
            if (DIM == 3) {
#pragma unroll
                for (int ll = 0; ll < N; ll++)
#pragma unroll
                    for (int nn = 0; nn < N; nn++) acc += Fp[fidx<N, DIM>(d, f, pq(ll, nn))] * w[ll][nn];
            } else if (DIM == 2) {
                const int oa = d == 0 ? o1 : o2;  // the one active transverse axis (y for x faces, x for y faces)
#pragma unroll
                for (int l = 0; l < N; l++) acc += Fp[fidx<N, DIM>(d, f, l)] * (float)a.S2C[oa * N + l];
            } else {
                acc = Fp[fidx<N, DIM>(d, f, 0)];
            }
            const double val = wall_el && f == 0 && v != 1 + d ? 0.0 : X[NVP + d * NVP + v] + (double)acc;
            ob[(size_t)v * vstride + off] = val;
            if (also_last) ob[(size_t)v * vstride + off_last] = val;
        }
    }
    (void)M;
}

// 3-D p=3 MFMA face integrals of the round's axis d with the writes staged
// through LDS: each wave (element) forms its 5 tiles of 16 x 16 outputs, puts
// them over its own flux slots (read for the last time by its MFMAs), and after
// a block barrier every thread writes one value per (slice, variable) such
// that 16 consecutive threads write 16 consecutive faces along x (the
// block's elements are consecutive along x): whole 128-byte rows instead of
// one element's runs of 4. Same values and ownership rules as mfv_faces.
// Called by every thread of the block.
template <int N, int DIM>
__device__ __forceinline__ void mfv_faces_staged(const MArgs& a, unsigned char* lds_base, int blk, int d, int s,
                                                 int el, bool live) {
    using G = MT<N, DIM>;
    constexpr int FSZ = G::FSZ, T = G::T, TPE = G::TPE, EPB = G::EPB;
    static_assert(N == 4 && DIM == 3 && TPE == 64 && EPB == 4, "3-D p=3 tiles");
    constexpr int NCOL = T * NVP * N;  // panels (t, v, f), f fastest
    // staged results: one row of RS floats per (slice, variable); RS = 8 mod 32
    // banks puts the four variables a 16-lane group writes in distinct banks
    constexpr int RS = 72;
    static_assert(T * NVP * RS <= G::LDS_F, "staged results fit the flux slots");
    const int dim1 = d == 0 ? 1 : (d == 1 ? 2 : 0), dim2 = dim1 == 0 ? 1 : (dim1 == 1 ? 2 : 0);
    float* Fs = (float*)((double*)(lds_base + (size_t)(el < EPB ? el : 0) * G::ELEM_BYTES) + G::LDS_X) + G::LDS_D;
    // pencil of axis d through transverse points ll (along dim1) and nn (along dim2)
    auto pq = [&](int ll, int nn) -> int { return d == 1 ? ll * N + nn : nn * N + ll; };
    typedef float f4 __attribute__((ext_vector_type(4)));
    f4 acc[NCOL / 16];
    if (live) {  // one wave per element: uniform
        const int lane = s, li = lane & 15, lk = lane >> 4;
        float av[4];
#pragma unroll
        for (int ks = 0; ks < 4; ks++) av[ks] = a.SS[li * 16 + 4 * ks + lk];
#pragma unroll
        for (int ct = 0; ct < NCOL / 16; ct++) {
            const int col = ct * 16 + li;
            const int t = col / (NVP * N), v = (col / N) % NVP, f = col % N;
            const float* Fp = Fs + (t * NVP + v) * FSZ;
            acc[ct] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < 4; ks++) {
                const int k = 4 * ks + lk, ll = k / N, nn = k % N;
                acc[ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[ks], Fp[fidx<N, DIM>(d, f, pq(ll, nn))], acc[ct], 0, 0, 0);
            }
        }
        // results over the element's flux slots: res[(t, v)][position, x fastest]
#pragma unroll
        for (int ct = 0; ct < NCOL / 16; ct++) {
            const int col = ct * 16 + li;
            const int t = col / (NVP * N), v = (col / N) % NVP, f = col % N;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int row = 4 * lk + i, o1 = row / N, o2 = row % N;
                int o[3];
                o[d] = f;
                o[dim1] = o1;
                o[dim2] = o2;
                Fs[(t * NVP + v) * RS + o[0] + N * (o[1] + N * o[2])] = acc[ct][i];
            }
        }
    }
    __syncthreads();
    // thread -> (element e2 of the block, position): 16 faces along x per row
    const int tid = threadIdx.x, xx = tid % (EPB * N), row = tid / (EPB * N);
    const int e2l = xx / N, ox = xx % N, oy = row % N, oz = row / N;
    const int e2 = blk * EPB + e2l;
    if (e2 >= a.E) return;
    const int ei2 = e2 % a.Ne[0], ej2 = (e2 / a.Ne[0]) % a.Ne[1], ek2 = e2 / (a.Ne[0] * a.Ne[1]);
    const int ec[3] = {ei2, ej2, ek2};
    const int Nd = d == 0 ? a.Ne[0] : (d == 1 ? a.Ne[1] : a.Ne[2]);
    const bool per = (d == 0 ? a.bc[0] : (d == 1 ? a.bc[1] : a.bc[2])) == SPDK_BC_PERIODIC;
    bool own = true;
#pragma unroll
    for (int x = 0; x < DIM; x++)
        own = own && ec[x] >= 1 && ec[x] <= (x == d ? (per ? a.Ne[x] - 2 : a.Ne[x] - 1) : a.Ne[x] - 2);
    const int o[3] = {ox, oy, oz};
    const int f = o[d];
    if (!own || (ec[d] == Nd - 1 && f != 0)) return;
    const bool wrap = per && ec[d] == 1 && f == 0;
    // a reflective wall's face: zero in every component but the normal momentum (as mfv_faces)
    const bool wall = f == 0 && (ec[d] == 1 || ec[d] == Nd - 1) &&
                      (d == 0 ? a.bc[0] : (d == 1 ? a.bc[1] : a.bc[2])) == SPDK_BC_REFLECTIVE;
    const int nx = a.fvn[0] + (d == 0), ny = a.fvn[1] + (d == 1), nz = a.fvn[2] + (d == 2);
    const size_t vstride = (size_t)nz * ny * nx;
    auto off = [&](int ex, int ey, int ez) -> size_t {
        const int I = ox + a.fgh[0] + (ex - 1) * N, J = oy + a.fgh[1] + (ey - 1) * N, K = oz + a.fgh[2] + (ez - 1) * N;
        return ((size_t)K * ny + J) * nx + I;
    };
    const size_t o0 = off(ei2, ej2, ek2);
    const size_t ow = wrap ? off(d == 0 ? a.Ne[0] - 1 : ei2, d == 1 ? a.Ne[1] - 1 : ej2, d == 2 ? a.Ne[2] - 1 : ek2) : 0;
    const double* X2 = (const double*)(lds_base + (size_t)e2l * G::ELEM_BYTES);
    const float* res = (const float*)(X2 + G::LDS_X) + G::LDS_D;
    const int pos = ox + N * (oy + N * oz);
#pragma unroll
    for (int t = 0; t < T; t++) {
        double* ob = d == 0 ? a.Ffv[t][0] : (d == 1 ? a.Ffv[t][1] : a.Ffv[t][2]);
#pragma unroll
        for (int v = 0; v < NVP; v++) {
            const double val = wall && v != 1 + d ? 0.0 : X2[NVP + d * NVP + v] + (double)res[(t * NVP + v) * RS + pos];
            ob[(size_t)v * vstride + o0] = val;
            if (wrap) ob[(size_t)v * vstride + ow] = val;
        }
    }
}

// Barrier between sweep phases that touch only the element's own LDS area.
// With one element per wave (TPE = 64) the wave's own order is enough: a
// wave's LDS operations execute in order, and the wavefront-scope fences keep
// the compiler from moving its LDS accesses across the point. The waves of a
// block then run their phases independently (a block barrier makes each wait
// for the slowest). Otherwise, and in the face sweep (MODE 2, whose staged
// writes cross elements), the block barrier.
template <int TPE, int MODE>
__device__ __forceinline__ void elem_sync() {
    if constexpr (TPE == 64 && MODE != 2) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// MODE 0: a predictor sweep (out_pred set); MODE 1: the corrector (face
// balance, conservative update, cv state and dt); MODE 2: the flux
// evaluation of the predictor's slices with their face integrals written
// for the FV update (mfv_faces), nothing else. Compiled apart, the predictor
// needs fewer VGPRs (3 waves per SIMD at 3-D p=3 instead of 2).
template <int N, int DIM, int MODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MODE == 0 && N == 4 && DIM == 3 ? 3 : SPDK_MSWEEP_WAVES))) k_msweep(MArgs a) {
    constexpr bool CONS = MODE == 1;
    using G = MT<N, DIM>;
    constexpr int M = G::M, NT = G::NT, TPE = G::TPE, EPB = G::EPB, T = G::T, FSZ = G::FSZ;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int tid = threadIdx.x;
    const int el = tid / TPE, s = tid % TPE;
    // MODE 2: workgroups go to the 8 XCDs in turn; XCD x takes the x-th
    // contiguous eighth of the element blocks, so the face rows that
    // neighbouring blocks share (rows are not line-aligned) are written
    // through one L2. The grid is a multiple of 8; blocks past the end exit.
    int blk;
    if (a.blist) {
        blk = a.blist[blockIdx.x];
    } else if (MODE == 2) {
        blk = (int)((blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8);
        if (blk * EPB >= a.E) return;  // uniform per block
    } else {
        blk = (int)blockIdx.x;
    }
    const int e = blk * EPB + el;
    const bool live = el < EPB && e < a.E;
    unsigned char* base = lds_raw + (size_t)(el < EPB ? el : 0) * G::ELEM_BYTES;
    double* X = (double*)base;  // mu[5], F(mu) per axis [3][5], residual
    float* Ds = (float*)(X + G::LDS_X);
    float* Fs = Ds + G::LDS_D;
    const size_t E = (size_t)a.E;
    const size_t slabE = E * TPE;
    const size_t eo = (size_t)e * TPE + s;
    constexpr int PSZ = G::PSZ;
    const int so = lpad<N, DIM>(s);  // this point in a padded state plane

    int ei = 0, ej = 0, ek = 0;
    if (live) {
        ei = e % a.Ne[0];
        ej = (e / a.Ne[0]) % a.Ne[1];
        ek = e / (a.Ne[0] * a.Ne[1]);
    }
    const int ii = s % N, jj = DIM >= 2 ? (s / N) % N : 0, kk = DIM == 3 ? s / (N * N) : 0;
    // pencil through this point and the point's position along it, per axis
    const int qx = DIM == 3 ? kk * N + jj : jj, qy = DIM == 3 ? kk * N + ii : ii, qz = jj * N + ii;
    const double wsp = pick(a.wq, ii) * (DIM >= 2 ? pick(a.wq, jj) : 1.0) * (DIM == 3 ? pick(a.wq, kk) : 1.0);

    if (live)
        for (int v = s; v < NVP; v += TPE) X[v] = a.mu[(size_t)v * E + e];
    if (live && s == 0) {  // active: own residual or a trace source's was >= tol last sweep
        bool act = true;
        if (a.big_in) {
            act = a.big_in[e] != 0;
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                size_t el0, el1;
                face_sources(a, d, ei, ej, ek, (size_t)e, el0, el1);
                act = act || a.big_in[el0] || a.big_in[el1];
            }
        }
        X[G::XACT] = act ? 1.0 : 0.0;
    }
    elem_sync<TPE, MODE>();
    const bool act_e = live && X[G::XACT] != 0.0;
    if (!__syncthreads_or(act_e)) {  // nothing to do in this block
        if (live && s == 0 && a.big_out) a.big_out[e] = 0;
        if (tid == 0) {
            if (a.res_part) a.res_part[blockIdx.x] = 0.0;
            if (a.cnt_part) a.cnt_part[blockIdx.x] = 0.0;
        }
        return;
    }
    if (act_e && s < DIM) {  // F(mu) along axis s
        double um[NVP], fb[NVP];
#pragma unroll
        for (int v = 0; v < NVP; v++) um[v] = X[vperm(v, s)];
        mflux(um, fb, a.gm - 1.0);
#pragma unroll
        for (int v = 0; v < NVP; v++) X[NVP + s * NVP + vperm(v, s)] = fb[v];
    }
    if (act_e && s == 0) {  // the mean state in FP32 for the difference-form fluxes
        float* XF = (float*)(X + G::XMF);
        const double ri = 1.0 / X[0];
        const double vx = X[1] * ri, vy = X[2] * ri, vz = X[3] * ri;
        const double pm = (a.gm - 1.0) * (X[4] - 0.5 * (X[1] * vx + X[2] * vy + X[3] * vz));
        XF[0] = (float)X[0];
        XF[1] = (float)vx;
        XF[2] = (float)vy;
        XF[3] = (float)vz;
        XF[4] = (float)X[1];
        XF[5] = (float)X[2];
        XF[6] = (float)X[3];
        XF[7] = (float)(X[4] + pm);
    }
    float dold[NVP];
#pragma unroll
    for (int v = 0; v < NVP; v++) dold[v] = act_e && MODE != 2 ? a.d_old[(size_t)v * slabE + eo] : 0.f;
    if (act_e) {
        if (a.path) {
#pragma unroll
            for (int t = 0; t < T; t++)
#pragma unroll
                for (int v = 0; v < NVP; v++) Ds[(t * NVP + v) * PSZ + so] = a.D[((size_t)t * NVP + v) * slabE + eo];
        } else {
#pragma unroll
            for (int v = 0; v < NVP; v++) Ds[v * PSZ + so] = dold[v];
        }
    }
    elem_sync<TPE, MODE>();

    float dudt[T][NVP];
#pragma unroll
    for (int t = 0; t < T; t++)
#pragma unroll
        for (int v = 0; v < NVP; v++) dudt[t][v] = 0.f;
    // the corrector conserves: accumulate the face balance
    const bool cons = CONS;
    double bal[NVP] = {0.0, 0.0, 0.0, 0.0, 0.0};

    // Rounds of pencil tasks, then the flux divergence of the round's fluxes.
    // Zero-order guess (path 0): one round, one slice, task s = d * NT + q in
    // flux slot d. Otherwise DIM rounds, round d: task s = t * NT + q in flux
    // slot t. The round loop is not unrolled: one copy of its code.
    const int nrounds = a.path ? DIM : 1;
    double wsum = 0.0;
#pragma unroll
    for (int t = 0; t < T; t++) wsum += a.wdt[t];
#pragma unroll 1
    for (int r = 0; r < nrounds; r++) {
        // tasks of the round: TPE (= T * NT) or DIM * NT (more than TPE only for p = 1 in 3-D)
        const int ntask = a.path ? TPE : DIM * NT;
        for (int task = s; act_e && task < ntask; task += TPE) {
            const int d = a.path ? r : task / NT, t = a.path ? task / NT : 0, q = task % NT;
            const int slot = a.path ? t : d;
            const int ts = a.tin_slices > 1 ? t : 0;
            const double bw = (a.path ? pick(a.wdt, t) : wsum) * pencil_weight<N, DIM>(a, q) * pick(a.ih, d);
            // the face integrals (MODE 2) read faces 0..M-2 only, unless the
            // x flux points are written as well
            pencil<N, DIM>(a, Ds + t * NVP * PSZ, Fs + slot * NVP * FSZ, X, d, q, ts, ei, ej, ek, (size_t)e, cons,
                           bal, bw, MODE == 2 && !a.F_x);
        }
        elem_sync<TPE, MODE>();
        if (MODE == 2) {
            if constexpr (N == 4 && DIM == 3) {
                if (!a.F_x && !a.fv_valu) mfv_faces_staged<N, DIM>(a, lds_raw, blk, r, s, el, act_e);
                else mfv_faces<N, DIM>(a, Fs, X, r, s, act_e, ei, ej, ek);
            } else {
                mfv_faces<N, DIM>(a, Fs, X, r, s, act_e, ei, ej, ek);
            }
        }
        if (act_e) {
            const int d0 = a.path ? r : 0, d1 = MODE == 2 ? d0 : (a.path ? r + 1 : DIM);
#pragma unroll 1
            for (int d = d0; d < d1; d++) {
                const int qdd = d == 0 ? qx : (d == 1 ? qy : qz), idd = d == 0 ? ii : (d == 1 ? jj : kk);
                const Geo gd = geo<N, DIM>(d, qdd);
                float Dr[M];
#pragma unroll
                for (int l = 0; l < M; l++) Dr[l] = a.Dh[(d * N + idd) * M + l];
#pragma unroll
                for (int t = 0; t < T; t++) {
                    if (!a.path && t > 0) break;
                    const float* Ft = Fs + (a.path ? t : d) * NVP * FSZ;
#pragma unroll
                    for (int v = 0; v < NVP; v++) {
                        const float* F = Ft + v * FSZ;
                        float acc = 0.f;
#pragma unroll
                        for (int l = 0; l < M; l++) acc += F[gd.fb + l * gd.fs] * Dr[l];
                        dudt[t][v] += acc;
                    }
                }
            }
            if (a.F_x && (a.path ? r == 0 : true)) {  // x fluxes at the flux points (outputs)
                for (int idx = s; idx < FSZ; idx += TPE)
                    for (int t = 0; t < T; t++) {
                        size_t o = (((size_t)t * NV) * E + e) * FSZ + idx;
                        const float* Ft = Fs + (a.path ? t : 0) * NVP * FSZ;  // path 0: slot 0 is x
#pragma unroll
                        for (int v = 0; v < NVP; v++) a.F_x[o + (size_t)v * E * FSZ] = X[NVP + v] + (double)Ft[v * FSZ + idx];
                        a.F_x[o + (size_t)NVP * E * FSZ] = 0.0;
                    }
            }
        }
        elem_sync<TPE, MODE>();
    }
    if (MODE == 2) return;
    if (!a.path)  // the zero-order guess has identical slices
#pragma unroll
        for (int t = 1; t < T; t++)
#pragma unroll
            for (int v = 0; v < NVP; v++) dudt[t][v] = dudt[0][v];

    if (!CONS && a.grav && act_e) {
        // update_prediction's gravity source (src/transforms.cpp:17-27,
        // :309-313), dudt -= S, in the reference's in-place order: the
        // momenta's source takes the density just written, the energy's the
        // momenta just written (formed here as the update below forms them).
        // S = S(mean) + S(deviation): the mean part in FP64, rounded once.
        float rn[T];
#pragma unroll
        for (int t = 0; t < T; t++) {
            float acc = 0.f;
#pragma unroll
            for (int l = 0; l < T; l++) acc += a.cdt[t * kMaxTm + l] * dudt[l][0];
            rn[t] = dold[0] - acc;
        }
        float mn[T][3];
#pragma unroll
        for (int v = 1; v <= 3; v++) {
            const float gv = (float)a.g[v - 1], smu = (float)(X[0] * a.g[v - 1]);
#pragma unroll
            for (int t = 0; t < T; t++) dudt[t][v] -= smu + rn[t] * gv;
#pragma unroll
            for (int t = 0; t < T; t++) {
                float acc = 0.f;
#pragma unroll
                for (int l = 0; l < T; l++) acc += a.cdt[t * kMaxTm + l] * dudt[l][v];
                mn[t][v - 1] = dold[v] - acc;
            }
        }
        const float smu4 = (float)(X[1] * a.g[0] + X[2] * a.g[1] + X[3] * a.g[2]);
        const float g0 = (float)a.g[0], g1 = (float)a.g[1], g2 = (float)a.g[2];
#pragma unroll
        for (int t = 0; t < T; t++) dudt[t][4] -= smu4 + (mn[t][0] * g0 + mn[t][1] * g1 + mn[t][2] * g2);
    }
    if (!CONS) {
        // ---- predictor: d_new(t) = d_old - sum_l cdt(t,l) dudt(l); residual vs the input state
        float res = 0.f;
        if (act_e) {
#pragma unroll
            for (int t = 0; t < T; t++)
#pragma unroll
                for (int v = 0; v < NVP; v++) {
                    float acc = 0.f;
#pragma unroll
                    for (int l = 0; l < T; l++) acc += a.cdt[t * kMaxTm + l] * dudt[l][v];
                    float dn = dold[v] - acc;
                    // input state at this point: its own LDS slot (path 1) or d_old
                    float din = a.path ? Ds[(t * NVP + v) * PSZ + so] : dold[v];
                    const float df = fabsf(dn - din);
                    if (!(df <= res)) res = (df == df) ? df : __builtin_inff();  // a NaN never passes
                    a.D[((size_t)t * NVP + v) * slabE + eo] = dn;
                    Ds[(t * NVP + v) * PSZ + so] = dn;
                }
        }
        elem_sync<TPE, MODE>();
        // traces of the new state into this element's older buffer: task =
        // (d, t, q), q fastest: at 3-D p=3 each 64-lane pass is one axis
        const int wb = act_e ? (a.ver[(size_t)a.E + e] < a.ver[e] ? 1 : 0) : 0;
        if (act_e) {
            for (int task = s; task < T * DIM * NT; task += TPE) {
                const int d = task / (T * NT), t = (task / NT) % T, q = task % NT;
                const Geo gt = geo<N, DIM>(d, q);
#pragma unroll
                for (int v = 0; v < NVP; v++) {
                    float x0 = 0.f, x1 = 0.f;
#pragma unroll
                    for (int l = 0; l < N; l++) {
                        float p = Ds[(t * NVP + v) * PSZ + gt.sb + l * gt.ss];
                        x0 += p * a.S2F[l];
                        x1 += p * a.S2F[(M - 1) * N + l];
                    }
                    float* tb = wb ? (d == 0 ? a.trb[1][0] : (d == 1 ? a.trb[1][1] : a.trb[1][2]))
                                   : (d == 0 ? a.trb[0][0] : (d == 1 ? a.trb[0][1] : a.trb[0][2]));
                    float* to = tb + ((((size_t)t * NVP + v) * E + e) * 2) * NT + q;
                    to[0] = x0;
                    to[NT] = x1;
                }
            }
        }
        // element maximum of the point residuals (all >= +0; a NaN stays)
        float re = 0.f;
        if constexpr (TPE == 64) {  // one element per wave: across the lanes
            re = res;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) re = nan_maxf(re, __shfl_xor(re, off, 64));
        } else {
            __syncthreads();
            if (act_e) Fs[s] = res;
            __syncthreads();
            if (act_e && s == 0)
                for (int j = 0; j < TPE; j++) re = nan_maxf(re, Fs[j]);
        }
        if (act_e && s == 0) {
            double sc = 0.0;
#pragma unroll
            for (int v = 0; v < NVP; v++) sc = fmax(sc, fabs(X[v]));
            // Ghost elements feed no interior element (interior faces read
            // interior traces only, face_sources), so their residual is not
            // a convergence criterion: a ghost that drifts never blocks it.
            const bool ghost = (a.act[0] && (ei < 1 || ei > a.Ne[0] - 2)) ||
                               (a.act[1] && (ej < 1 || ej > a.Ne[1] - 2)) ||
                               (a.act[2] && (ek < 1 || ek > a.Ne[2] - 2));
            double rr = ghost ? 0.0 : (double)re / (sc > 0.0 ? sc : 1.0);
            X[G::XRES] = rr;
            if (a.conv && rr < a.tol && a.conv[e] == 0xff) a.conv[e] = (unsigned char)a.sweep;
            if (a.big_out) a.big_out[e] = rr < a.tol ? 0 : 1;  // NaN: not converged
            a.ver[(size_t)wb * E + e] = a.cur;
        }
        if (live && !act_e && s == 0) {  // kept state and traces
            X[G::XRES] = 0.0;
            if (a.big_out) a.big_out[e] = 0;
        }
        __syncthreads();
        if (tid == 0) {
            double rb = 0.0, cnt = 0.0;
            for (int j = 0; j < EPB; j++)
                if (blk * EPB + j < a.E) {
                    const double* Xj = (const double*)(lds_raw + (size_t)j * G::ELEM_BYTES);
                    rb = nan_max(rb, Xj[G::XRES]);
                    cnt += Xj[G::XACT];
                }
            a.res_part[blockIdx.x] = rb;
            if (a.cnt_part) a.cnt_part[blockIdx.x] = cnt;
        }
        return;
    }

    // ---- corrector (src/transforms.cpp:354-372), FP64 accumulation, conservative shift
    double un[NVP], du[NVP];
#pragma unroll
    for (int v = 0; v < NVP; v++) {
        du[v] = 0.0;
#pragma unroll
        for (int t = 0; t < T; t++) du[v] += a.wdt[t] * (double)dudt[t][v];
    }
    // update_solution's gravity source (src/transforms.cpp:364-367) from the
    // final predictor slices, in FP64: sum_t w_t dt S_t, added after the
    // conservative shift (which balances the fluxes only)
    double sg[NVP] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (a.grav && act_e) {
#pragma unroll
        for (int t = 0; t < T; t++) {
            const float* Dt = Ds + (a.path ? t : 0) * NVP * PSZ + so;  // path 0: one slice, d_old
            const double rho = X[0] + (double)Dt[0];
            const double m1 = X[1] + (double)Dt[PSZ], m2 = X[2] + (double)Dt[2 * PSZ], m3 = X[3] + (double)Dt[3 * PSZ];
            sg[1] += a.wdt[t] * (rho * a.g[0]);
            sg[2] += a.wdt[t] * (rho * a.g[1]);
            sg[3] += a.wdt[t] * (rho * a.g[2]);
            sg[4] += a.wdt[t] * (m1 * a.g[0] + m2 * a.g[1] + m3 * a.g[2]);
        }
    }
    __syncthreads();
    cons_shift<N, DIM>(du, bal, wsp, a.wtot, (double*)Ds, X + G::XSHIFT, act_e, s);
#pragma unroll
    for (int v = 1; v < NVP; v++) du[v] -= sg[v];
#pragma unroll
    for (int v = 0; v < NVP; v++) un[v] = act_e ? a.U_sp[(size_t)v * slabE + eo] - du[v] : 1.0;
    if (act_e)
#pragma unroll
        for (int v = 0; v < NVP; v++) a.U_new[(size_t)v * slabE + eo] = un[v];
    __syncthreads();
    double r = cv_and_dt<N, DIM>(un, (double*)Ds, act_e, s, ii, jj, kk, ei, ej, ek, (size_t)e, a.W_cv, slabE, a.S2C,
                                 a.Ne, a.act, a.h, a.gm, a.cfl);
    block_min<G::BLOCK>(r, (double*)lds_raw, a.part);
}

// Step start: element means, deviations and their face traces (one slice);
// resets the per-element convergence marks.
template <int N, int DIM>
__global__ void __launch_bounds__(256) k_mstart(const double* __restrict__ U_sp, double* __restrict__ mu,
                                                float* __restrict__ d_old, float* tr0, float* tr1, float* tr2,
                                                const float* __restrict__ S2F, unsigned char* conv, int* ver,
                                                int Eint) {
    using G = MT<N, DIM>;
    constexpr int M = G::M, NT = G::NT, TPE = G::TPE, EPB = G::EPB;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int tid = threadIdx.x, el = tid / TPE, s = tid % TPE;
    const int e = blockIdx.x * EPB + el;
    const bool live = el < EPB && e < Eint;
    unsigned char* base = lds_raw + (size_t)(el < EPB ? el : 0) * G::ELEM_BYTES;
    double* mloc = (double*)base;          // [v] element means
    double* us = mloc + 8;                 // [v][TPE] FP64
    float* ds = (float*)(us + NVP * TPE);  // [v][TPE] FP32
    const size_t E = (size_t)Eint, slabE = E * TPE;
    const size_t eo = (size_t)e * TPE + s;
    double u[NVP];
    if (live)
#pragma unroll
        for (int v = 0; v < NVP; v++) {
            u[v] = U_sp[(size_t)v * slabE + eo];
            us[v * TPE + s] = u[v];
        }
    __syncthreads();
    if (live)
        for (int v = s; v < NVP; v += TPE) {  // an element may have fewer points than variables
            double acc = 0.0;
            for (int j = 0; j < TPE; j++) acc += us[v * TPE + j];
            double m = acc * (1.0 / TPE);
            mloc[v] = m;
            mu[(size_t)v * E + e] = m;
        }
    if (live && s == 0) {  // traces of this state are in buffer 0, "written" before sweep 1
        if (conv) conv[e] = 0xff;
        ver[e] = 0;
        ver[(size_t)Eint + e] = -1;
    }
    __syncthreads();
    if (live)
#pragma unroll
        for (int v = 0; v < NVP; v++) {
            float d = (float)(u[v] - mloc[v]);
            d_old[(size_t)v * slabE + eo] = d;
            ds[v * TPE + s] = d;
        }
    __syncthreads();
    for (int task = s; live && task < DIM * NT; task += TPE) {
        const int d = task / NT, q = task % NT;
        float* trd = d == 0 ? tr0 : (d == 1 ? tr1 : tr2);
#pragma unroll
        for (int v = 0; v < NVP; v++) {
            float x0 = 0.f, x1 = 0.f;
#pragma unroll
            for (int l = 0; l < N; l++) {
                float p = ds[v * TPE + spidx<N, DIM>(d, l, q)];
                x0 += p * S2F[l];
                x1 += p * S2F[(M - 1) * N + l];
            }
            float* to = trd + (((size_t)v * E + e) * 2) * NT + q;
            to[0] = x0;
            to[NT] = x1;
        }
    }
}

// Face traces of a T-slice state (after an externally supplied guess).
template <int N, int DIM>
__global__ void __launch_bounds__(256) k_mtrace(const float* __restrict__ D, float* tr0, float* tr1, float* tr2,
                                                const float* __restrict__ S2F, int* ver, int Eint) {
    using G = MT<N, DIM>;
    constexpr int M = G::M, NT = G::NT, TPE = G::TPE, EPB = G::EPB, T = G::T;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int tid = threadIdx.x, el = tid / TPE, s = tid % TPE;
    const int e = blockIdx.x * EPB + el;
    const bool live = el < EPB && e < Eint;
    float* Ds = (float*)(lds_raw + (size_t)(el < EPB ? el : 0) * G::ELEM_BYTES);
    const size_t E = (size_t)Eint, slabE = E * TPE;
    if (live)
        for (int t = 0; t < T; t++)
#pragma unroll
            for (int v = 0; v < NVP; v++) Ds[(t * NVP + v) * TPE + s] = D[((size_t)t * NVP + v) * slabE + (size_t)e * TPE + s];
    __syncthreads();
    if (!live) return;
    if (s == 0) {
        ver[e] = 0;
        ver[(size_t)Eint + e] = -1;
    }
    for (int task = s; task < T * DIM * NT; task += TPE) {
        const int t = task / (DIM * NT), r = task % (DIM * NT), d = r / NT, q = r % NT;
        float* trd = d == 0 ? tr0 : (d == 1 ? tr1 : tr2);
#pragma unroll
        for (int v = 0; v < NVP; v++) {
            float x0 = 0.f, x1 = 0.f;
#pragma unroll
            for (int l = 0; l < N; l++) {
                float p = Ds[(t * NVP + v) * TPE + spidx<N, DIM>(d, l, q)];
                x0 += p * S2F[l];
                x1 += p * S2F[(M - 1) * N + l];
            }
            float* to = trd + ((((size_t)t * NVP + v) * E + e) * 2) * NT + q;
            to[0] = x0;
            to[NT] = x1;
        }
    }
}

// Restart from the zero-order guess: the elements flagged in `flags` get
// every time slice of the predictor set to d_old and the traces of d_old
// written, as version `cur`, into the older of their two buffers.
// next_flags marks them, so the next sweep updates them and the elements
// that read their traces.
template <int N, int DIM>
__global__ void __launch_bounds__(256) k_mrestart(const float* __restrict__ d_old, float* __restrict__ D, float* trA0,
                                                  float* trA1, float* trA2, float* trB0, float* trB1, float* trB2,
                                                  const float* __restrict__ S2F, int* ver,
                                                  const unsigned char* __restrict__ flags, unsigned char* next_flags,
                                                  int cur, int Eint) {
    using G = MT<N, DIM>;
    constexpr int M = G::M, NT = G::NT, TPE = G::TPE, EPB = G::EPB, T = G::T;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int tid = threadIdx.x, el = tid / TPE, s = tid % TPE;
    const int e = blockIdx.x * EPB + el;
    const bool live = el < EPB && e < Eint;
    const bool go = live && flags[e] != 0;
    float* ds = (float*)(lds_raw + (size_t)(el < EPB ? el : 0) * G::ELEM_BYTES);
    const size_t E = (size_t)Eint, slabE = E * TPE, eo = (size_t)e * TPE + s;
    int wb = 0;
    if (go) {
#pragma unroll
        for (int v = 0; v < NVP; v++) {
            const float x = d_old[(size_t)v * slabE + eo];
            ds[v * TPE + s] = x;
#pragma unroll
            for (int t = 0; t < T; t++) D[((size_t)t * NVP + v) * slabE + eo] = x;
        }
        wb = ver[E + e] < ver[e] ? 1 : 0;
    }
    __syncthreads();
    if (live && s == 0) next_flags[e] = go ? 1 : 0;
    for (int task = s; go && task < DIM * NT; task += TPE) {
        const int d = task / NT, q = task % NT;
        float* tb = wb ? (d == 0 ? trB0 : (d == 1 ? trB1 : trB2)) : (d == 0 ? trA0 : (d == 1 ? trA1 : trA2));
#pragma unroll
        for (int v = 0; v < NVP; v++) {
            float x0 = 0.f, x1 = 0.f;
#pragma unroll
            for (int l = 0; l < N; l++) {
                float p = ds[v * TPE + spidx<N, DIM>(d, l, q)];
                x0 += p * S2F[l];
                x1 += p * S2F[(M - 1) * N + l];
            }
#pragma unroll
            for (int t = 0; t < T; t++) {
                float* to = tb + ((((size_t)t * NVP + v) * E + e) * 2) * NT + q;
                to[0] = x0;
                to[NT] = x1;
            }
        }
    }
    if (go && s == 0) ver[(size_t)wb * E + e] = cur;
}

// The predictor buffer as FP64 slices in the stage layout (t, v, element,
// point): U = mu + d for the five physical variables, 0 in any further slot.
__global__ void k_mexport(const float* __restrict__ D, const double* __restrict__ mu, double* __restrict__ U, int nt,
                          int nv, size_t slabE, size_t tpe, size_t total) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const size_t r = i % slabE, tv = i / slabE;
        const int v = (int)(tv % nv), t = (int)(tv / nv);
        const size_t E = slabE / tpe;
        U[i] = v < NVP ? mu[(size_t)v * E + r / tpe] + (double)D[((size_t)t * NVP + v) * slabE + r] : 0.0;
    }
}

__global__ void k_fill_u8(unsigned char* p, size_t n, unsigned char v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// out[0] = min of a[0..n) (a may be null: +inf), out[1] = max of b[0..n),
// out[2] = sum of c[0..n) (b, c may be null: 0); zeroes *counter (if given)
// for the k_mblocks that follows
__global__ void k_mreduce(const double* a, const double* b, const double* c, int n, double* out, int* counter) {
    __shared__ double ra[1024], rb[1024], rc[1024];
    if (counter && threadIdx.x == 0) *counter = 0;
    double x = __builtin_inf(), y = 0.0, z = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        if (a) x = nan_min(x, a[i]);
        if (b) y = nan_max(y, b[i]);
        if (c) z += c[i];
    }
    ra[threadIdx.x] = x;
    rb[threadIdx.x] = y;
    rc[threadIdx.x] = z;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            ra[threadIdx.x] = nan_min(ra[threadIdx.x], ra[threadIdx.x + w]);
            rb[threadIdx.x] = nan_max(rb[threadIdx.x], rb[threadIdx.x + w]);
            rc[threadIdx.x] += rc[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = ra[0];
        out[1] = rb[0];
        out[2] = rc[0];
    }
}

// List of the blocks that hold an element the next sweep updates (its own
// flag or a trace source's flag set), appended in any order; *counter must
// be zero on entry. One thread per block.
template <int N, int DIM>
__global__ void __launch_bounds__(256) k_mblocks(MArgs a, const unsigned char* __restrict__ flags, int nblocks,
                                                 int* list, int* counter) {
    constexpr int EPB = MT<N, DIM>::EPB;
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblocks) return;
    bool any = false;
    for (int j = 0; j < EPB && !any; j++) {
        const int e = b * EPB + j;
        if (e >= a.E) break;
        const int ei = e % a.Ne[0], ej = (e / a.Ne[0]) % a.Ne[1], ek = e / (a.Ne[0] * a.Ne[1]);
        bool act = flags[e] != 0;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            size_t el0, el1;
            face_sources(a, d, ei, ej, ek, (size_t)e, el0, el1);
            act = act || flags[el0] || flags[el1];
        }
        any = act;
    }
    if (any) list[atomicAdd(counter, 1)] = b;
}

// Histogram of the per-element convergence sweep (bin k = k predictor
// sweeps needed; the last bin counts elements that never passed).
__global__ void k_mhist(const unsigned char* conv, int E, int nb, unsigned long long* hist) {
    __shared__ unsigned int loc[kHistBins];
    if (threadIdx.x < kHistBins) loc[threadIdx.x] = 0;
    __syncthreads();
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        int c = conv[e];
        int bin = c == 0xff ? nb - 1 : (c + 1 < nb - 1 ? c + 1 : nb - 1);
        atomicAdd(&loc[bin], 1u);
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)nb && loc[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)loc[threadIdx.x]);
}

// ---------------------------------------------------------------------------
// Learned guess. Each element is made dimensionless with its mean state:
// density by rho, momenta by rho c, energy by rho c^2 (c the sound speed of
// the mean state); its amplitude A = max |d / scale| (floored at kAmin) sets
// the unit of the features, and A times the acoustic Courant number
// kappa = dt c / h that of the targets:
//   features  x = d(v, sp) / (scale_v A)  (variables rho, the active momenta, E),
//             then v/c (3 components), kappa, log10 A, 1;
//   targets   y = (U*(t, v, sp) - U(v, sp)) / (scale_v kappa A),
// with U* the converged space-time predictor. The momentum of an inactive
// axis is not in the network (its guess is the zero-order one).
template <int DIM>
__device__ __forceinline__ int gvar(int vi) { return vi <= DIM ? vi : 4; }

// per element: scale_0..4, kappa, A, kappa A, v/c (3)
template <int N, int DIM>
__global__ void k_gscale(const double* __restrict__ mu, const float* __restrict__ d_old, int Eint, int e0, int nrows,
                         double dt, double hmin, double gm, float* __restrict__ S) {
    constexpr int TPE = MT<N, DIM>::TPE;
    const int r0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (r0 >= nrows) return;
    const int e = e0 + r0;
    const size_t E = (size_t)Eint;
    const double r = mu[e], ri = 1.0 / r;
    const double v1 = mu[E + e] * ri, v2 = mu[2 * E + e] * ri, v3 = mu[3 * E + e] * ri;
    const double p = (mu[4 * E + e] - 0.5 * r * (v1 * v1 + v2 * v2 + v3 * v3)) * (gm - 1.0);
    const double c = sqrt(fmax(gm * p * ri, kMinC2));
    const double sc[NVP] = {r, r * c, r * c, r * c, r * c * c};
    float a = 0.f;
    for (int vi = 0; vi < DIM + 2; vi++) {
        const int v = gvar<DIM>(vi);
        const float inv = (float)(1.0 / sc[v]);
        for (int s = 0; s < TPE; s++) a = fmaxf(a, fabsf(d_old[((size_t)v * E + e) * TPE + s] * inv));
    }
    a = fmaxf(a, kAmin);
    const double kap = dt * c / hmin;
    float* o = S + (size_t)kGS * e;
    for (int v = 0; v < NVP; v++) o[v] = (float)sc[v];
    o[5] = (float)kap;
    o[6] = a;
    o[7] = (float)(kap * a);
    o[8] = (float)(v1 / c);
    o[9] = (float)(v2 / c);
    o[10] = (float)(v3 / c);
    o[11] = 0.f;
}

template <int N, int DIM>
__global__ void k_gfeat(const float* __restrict__ d_old, const float* __restrict__ S, int Eint, int e0, int nrows,
                        float* __restrict__ X, long ldx, const int* __restrict__ list) {
    constexpr int TPE = MT<N, DIM>::TPE, NF = (DIM + 2) * TPE;
    const long total = (long)nrows * (NF + kGExtra);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const int row = (int)(i / (NF + kGExtra)), j = (int)(i % (NF + kGExtra));
        const size_t e = list ? (size_t)list[e0 + row] : (size_t)e0 + row;
        const float* sc = S + (size_t)kGS * e;
        float x;
        if (j < NF) {
            const int v = gvar<DIM>(j / TPE), s = j % TPE;
            x = d_old[((size_t)v * Eint + e) * TPE + s] / (sc[v] * sc[6]);
        } else {
            const int q = j - NF;
            x = q < 3 ? sc[8 + q] : (q == 3 ? sc[5] : (q == 4 ? log10f(sc[6]) : 1.0f));
        }
        X[(long)row * ldx + j] = x;
    }
}

// targets from the predictor state D (deviations) of the elements [e0, e0 + nrows)
template <int N, int DIM>
__global__ void k_gtarget(const float* __restrict__ D, const float* __restrict__ d_old, const float* __restrict__ S,
                          int Eint, int e0, int nrows, float* __restrict__ Y, long ldy) {
    constexpr int TPE = MT<N, DIM>::TPE, NV = DIM + 2, NO = MT<N, DIM>::T * NV * TPE;
    const size_t E = (size_t)Eint, slabE = E * TPE;
    const long total = (long)nrows * NO;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const int row = (int)(i / NO), j = (int)(i % NO);
        const int t = j / (NV * TPE), vi = (j / TPE) % NV, s = j % TPE;
        const int v = gvar<DIM>(vi);
        const size_t e = (size_t)e0 + row;
        const float* sc = S + (size_t)kGS * e;
        const float dn = D[((size_t)t * NVP + v) * slabE + e * TPE + s];
        const float d0 = d_old[(size_t)v * slabE + e * TPE + s];
        Y[(long)row * ldy + j] = (dn - d0) / (sc[v] * sc[7]);
    }
}

// guess from network outputs: D(t, v, e, sp) for the elements [e0, e0 + nrows)
template <int N, int DIM>
__global__ void k_gdecode(const float* __restrict__ Y, long ldy, const float* __restrict__ d_old,
                          const float* __restrict__ S, int Eint, int e0, int nrows, float* __restrict__ D,
                          const int* __restrict__ list) {
    constexpr int TPE = MT<N, DIM>::TPE, NV = DIM + 2, T = MT<N, DIM>::T;
    const size_t E = (size_t)Eint, slabE = E * TPE;
    const long total = (long)T * NVP * nrows * TPE;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const int s = (int)(i % TPE);
        const long r2 = i / TPE;
        const int row = (int)(r2 % nrows);
        const int tv = (int)(r2 / nrows), t = tv / NVP, v = tv % NVP;
        const size_t e = list ? (size_t)list[e0 + row] : (size_t)e0 + row;
        const float d0 = d_old[(size_t)v * slabE + e * TPE + s];
        float dn = d0;
        const int vi = v == 4 ? DIM + 1 : v;
        const float* sc = S + (size_t)kGS * e;
        // an element (nearly) uniform inside keeps the zero-order guess: alone
        // it is its own fixed point, and the network cannot see the
        // neighbours that would move it
        if ((v == 4 || v <= DIM) && sc[6] > kAgate)
            dn = d0 + Y[(long)row * ldy + (t * NV + vi) * TPE + s] * (sc[v] * sc[7]);
        D[((size_t)t * NVP + v) * slabE + e * TPE + s] = dn;
    }
}

// Elements whose amplitude is above the gate (the others keep the
// zero-order guess): appended to list in any order; *counter must be zero.
__global__ void k_glist(const float* __restrict__ S, int Eint, int* __restrict__ list, int* counter) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < Eint && S[(size_t)kGS * e + 6] > kAgate) list[atomicAdd(counter, 1)] = e;
}
__global__ void k_izero(int* p) { *p = 0; }
// The zero-order guess in every slice of the predictor buffer, for the
// elements at or below the amplitude gate (the others get the network's
// guess from k_gdecode: the same predicate as k_glist).
__global__ void k_gbroadcast(const float* __restrict__ d_old, float* __restrict__ D, int T, size_t slab,
                             const float* __restrict__ S, int tpe) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < slab; i += (size_t)gridDim.x * blockDim.x) {
        const size_t v = i / (slab / NVP), r = i % (slab / NVP);  // slab = NVP * E * TPE
        if (S[(size_t)kGS * (r / tpe) + 6] > kAgate) continue;
        const float x = d_old[i];
        for (int t = 0; t < T; t++) D[((size_t)t * NVP + v) * (slab / NVP) + r] = x;
    }
}

struct MLayout {
    size_t E = 0, tpe = 0, nt = 0;
    unsigned blocks = 0;
    // byte offsets
    size_t mu = 0, part = 0, resp = 0, cntp = 0, red2 = 0, blist = 0, hist = 0, dold = 0, D = 0, trA = 0, trB = 0,
           conv = 0, ver = 0, big = 0, gsc = 0, total = 0;
    size_t tr_axis = 0;  // bytes per axis per trace buffer
};

inline size_t al(size_t x) { return (x + 255) / 256 * 256; }

template <int N, int DIM>
MLayout mlayout(size_t E) {
    using G = MT<N, DIM>;
    MLayout L;
    L.E = E;
    L.tpe = G::TPE;
    L.nt = G::NT;
    L.blocks = (unsigned)((E + G::EPB - 1) / G::EPB);
    size_t o = 0;
    L.mu = o; o = al(o + sizeof(double) * NVP * E);
    L.part = o; o = al(o + sizeof(double) * (L.blocks + 8));
    L.resp = o; o = al(o + sizeof(double) * (L.blocks + 8));
    L.cntp = o; o = al(o + sizeof(double) * (L.blocks + 8));
    L.red2 = o; o = al(o + sizeof(double) * 8);  // [0..2] reductions, [4] block-list length (int)
    L.blist = o; o = al(o + sizeof(int) * (L.blocks + 8));
    L.hist = o; o = al(o + sizeof(unsigned long long) * kHistBins);
    L.dold = o; o = al(o + sizeof(float) * NVP * E * G::TPE);
    L.D = o; o = al(o + sizeof(float) * G::T * NVP * E * G::TPE);
    L.tr_axis = al(sizeof(float) * G::T * NVP * E * 2 * G::NT);
    L.trA = o; o += DIM * L.tr_axis;
    L.trB = o; o += DIM * L.tr_axis;
    L.conv = o; o = al(o + E);
    L.ver = o; o = al(o + sizeof(int) * 2 * E);
    L.big = o; o = al(o + 2 * E);
    L.gsc = o; o = al(o + sizeof(float) * kGS * E);
    L.total = o;
    return L;
}

// Quadrature weights of the N Chebyshev-Gauss solution points on [0, 1]
// (x_i = (1 - cos((2i+1) pi / 2N)) / 2, src/polynomials.cpp:89-93): the
// exact integrals of their Lagrange basis polynomials, expanded in powers of
// y = x - 1/2 (the points are symmetric about 1/2, which keeps the
// expansion well conditioned); odd powers integrate to zero.
void sp_weights(int n, double* w) {
    double y[kMaxTm];
    for (int i = 0; i < n; i++) y[i] = -0.5 * std::cos((2 * i + 1) / (2.0 * n) * 3.141592653589793);
    for (int i = 0; i < n; i++) {
        double c[kMaxTm + 1] = {1.0};  // coefficients of prod_{j != i} (y - y_j), ascending powers
        int deg = 0;
        double den = 1.0;
        for (int j = 0; j < n; j++) {
            if (j == i) continue;
            for (int k = deg + 1; k > 0; k--) c[k] = c[k - 1] - y[j] * c[k];
            c[0] = -y[j] * c[0];
            deg++;
            den *= y[i] - y[j];
        }
        double integral = 0.0;
        for (int k = 0; k <= deg; k += 2) integral += c[k] * 2.0 * std::pow(0.5, k + 1) / (k + 1);
        w[i] = integral / den;
    }
}

int ensure_mixed_mats(spdk_ctx* c) {
    if (c->mixed_mats) return SPDK_OK;
    const int n = c->n, m = c->m;
    // sp_to_fp, dfp_to_sp / h_d per axis, then S x S of sp_to_cv (row o1 n + o2, column ll n + nn)
    std::vector<float> h((size_t)m * n + 3 * (size_t)n * m + (size_t)n * n * n * n, 0.f);
    const std::vector<double>& s2f = c->mat_host[SPDK_M_SP_TO_FP];
    const std::vector<double>& dm = c->mat_host[SPDK_M_DFP_TO_SP];
    const std::vector<double>& s2c = c->mat_host[SPDK_M_SP_TO_CV];
    for (int i = 0; i < m * n; i++) h[i] = (float)s2f[i];
    for (int d = 0; d < 3; d++)
        for (int i = 0; i < n * m; i++) h[(size_t)m * n + d * n * m + i] = (float)(dm[i] / c->prm.h[d]);
    float* ss = h.data() + (size_t)m * n + 3 * (size_t)n * m;
    for (int o1 = 0; o1 < n; o1++)
        for (int o2 = 0; o2 < n; o2++)
            for (int ll = 0; ll < n; ll++)
                for (int nn = 0; nn < n; nn++)
                    ss[(o1 * n + o2) * n * n + ll * n + nn] = (float)(s2c[o1 * n + ll] * s2c[o2 * n + nn]);
    void* p = nullptr;
    SPDK_CHECK(hipMalloc(&p, h.size() * sizeof(float)));
    c->owned.push_back(p);
    SPDK_CHECK(hipMemcpy(p, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    c->mixed_mats = (float*)p;
    return SPDK_OK;
}

inline unsigned gblocks(long n) { long b = (n + 255) / 256; return (unsigned)(b > 262144 ? 262144 : b); }

double hmin_of(spdk_ctx* c) {
    double h = 0.0;
    for (int d = 0; d < 3; d++)
        if (c->prm.active[d]) h = (h == 0.0 || c->prm.h[d] < h) ? c->prm.h[d] : h;
    return h;
}

// features (X) and / or targets of the current predictor state (Y) for the
// elements [e0, e0 + nrows); the scales are refreshed for that range
template <int N, int DIM>
int guess_data(spdk_ctx* c, void* ws, size_t E, double dt, int e0, int nrows, float* X, long ldx, float* Y, long ldy) {
    MLayout L = mlayout<N, DIM>(E);
    char* W = (char*)ws;
    float* S = (float*)(W + L.gsc);
    const double* mu = (const double*)(W + L.mu);
    const float* d_old = (const float*)(W + L.dold);
    hipLaunchKernelGGL((k_gscale<N, DIM>), dim3((nrows + 255) / 256), dim3(256), 0, c->stream, mu, d_old, (int)E, e0,
                       nrows, dt, hmin_of(c), c->prm.gamma, S);
    SPDK_LAUNCH_CHECK("k_gscale");
    constexpr int TPE = MT<N, DIM>::TPE;
    if (X) {
        hipLaunchKernelGGL((k_gfeat<N, DIM>), dim3(gblocks((long)nrows * ((DIM + 2) * TPE + kGExtra))), dim3(256), 0,
                           c->stream, d_old, (const float*)S, (int)E, e0, nrows, X, ldx, (const int*)nullptr);
        SPDK_LAUNCH_CHECK("k_gfeat");
    }
    if (Y) {
        hipLaunchKernelGGL((k_gtarget<N, DIM>), dim3(gblocks((long)nrows * MT<N, DIM>::T * (DIM + 2) * TPE)),
                           dim3(256), 0, c->stream, (const float*)(W + L.D), d_old, (const float*)S, (int)E, e0, nrows,
                           Y, ldy);
        SPDK_LAUNCH_CHECK("k_gtarget");
    }
    return SPDK_OK;
}

// the attached network's guess for every element, written to the predictor buffer
template <int N, int DIM>
int mlp_guess(spdk_ctx* c, void* ws, size_t E, double dt) {
    constexpr int TPE = MT<N, DIM>::TPE, din = (DIM + 2) * TPE + kGExtra, dout = MT<N, DIM>::T * (DIM + 2) * TPE;
    spdk_mlp* net = c->guess_net;
    int pd[4];
    spdk_mlp_dims(net, pd);
    if (!c->guess_x) {
        int cap = 65536;
        while (cap > 1024 && (size_t)cap * pd[0] > (size_t)1 << 27) cap /= 2;
        c->guess_cap = cap;
        void *x = nullptr, *y = nullptr;
        SPDK_CHECK(hipMalloc(&x, (size_t)cap * din * sizeof(float)));
        SPDK_CHECK(hipMalloc(&y, (size_t)cap * dout * sizeof(float)));
        c->owned.push_back(x);
        c->owned.push_back(y);
        c->guess_x = (float*)x;
        c->guess_y = (float*)y;
    }
    if (c->guess_list_cap < E + 1) {
        void* l = nullptr;
        SPDK_CHECK(hipMalloc(&l, (E + 1) * sizeof(int)));
        c->owned.push_back(l);
        c->guess_list = (int*)l;
        c->guess_list_cap = E + 1;
    }
    MLayout L = mlayout<N, DIM>(E);
    char* W = (char*)ws;
    hipLaunchKernelGGL((k_gscale<N, DIM>), dim3((unsigned)((E + 255) / 256)), dim3(256), 0, c->stream,
                       (const double*)(W + L.mu), (const float*)(W + L.dold), (int)E, 0, (int)E, dt, hmin_of(c),
                       c->prm.gamma, (float*)(W + L.gsc));
    SPDK_LAUNCH_CHECK("k_gscale");
    // the network runs only for elements above the amplitude gate; every
    // element first gets the zero-order guess
    int* list = c->guess_list;
    int* counter = c->guess_list + E;
    hipLaunchKernelGGL(k_izero, dim3(1), dim3(1), 0, c->stream, counter);
    SPDK_LAUNCH_CHECK("k_izero");
    hipLaunchKernelGGL(k_glist, dim3((unsigned)((E + 255) / 256)), dim3(256), 0, c->stream, (const float*)(W + L.gsc),
                       (int)E, list, counter);
    SPDK_LAUNCH_CHECK("k_glist");
    const size_t slab = (size_t)NVP * E * TPE;
    hipLaunchKernelGGL(k_gbroadcast, dim3(gblocks((long)slab)), dim3(256), 0, c->stream, (const float*)(W + L.dold),
                       (float*)(W + L.D), MT<N, DIM>::T, slab, (const float*)(W + L.gsc), TPE);
    SPDK_LAUNCH_CHECK("k_gbroadcast");
    SPDK_CHECK(hipMemcpyAsync(c->red_host, counter, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    SPDK_CHECK(hipStreamSynchronize(c->stream));
    int nlist = 0;
    std::memcpy(&nlist, c->red_host, sizeof(int));
    for (int e0 = 0; e0 < nlist; e0 += c->guess_cap) {
        const int nrows = std::min(c->guess_cap, nlist - e0);
        hipLaunchKernelGGL((k_gfeat<N, DIM>), dim3(gblocks((long)nrows * din)), dim3(256), 0, c->stream,
                           (const float*)(W + L.dold), (const float*)(W + L.gsc), (int)E, e0, nrows, c->guess_x,
                           (long)din, (const int*)list);
        SPDK_LAUNCH_CHECK("k_gfeat");
        int rc = spdk_mlp_forward(net, c->guess_x, din, nrows, c->guess_y, dout);
        if (rc) return rc;
        hipLaunchKernelGGL((k_gdecode<N, DIM>), dim3(gblocks((long)MT<N, DIM>::T * NVP * nrows * TPE)), dim3(256), 0,
                           c->stream, (const float*)c->guess_y, (long)dout, (const float*)(W + L.dold),
                           (const float*)(W + L.gsc), (int)E, e0, nrows, (float*)(W + L.D), (const int*)list);
        SPDK_LAUNCH_CHECK("k_gdecode");
    }
    return SPDK_OK;
}

// Where the flux evaluation of the predictor writes its face integrals (the
// FV update follows; spdk_mixed_predict_fv_fluxes).
struct FvOut {
    double* const* Ffv;  // [t * 3 + d]
    int fvn[3], fgh[3];
    int valu;
};

template <int N, int DIM>
int run_mixed(spdk_ctx* c, double* U_sp, const int Ne[3], void* ws, double dt, double* W_cv, double* F_x,
              const spdk_picard* pc, spdk_picard_stats* st, double* dt_next, bool predict_only,
              const FvOut* fvo = nullptr) {
    using G = MT<N, DIM>;
    constexpr int T = G::T;
    if (c->T != T) { set_error("mixed path: ADER with T = p+1 time nodes only"); return SPDK_E_CONFIG; }
    int rc;
    if ((rc = ensure_mixed_mats(c))) return rc;
    const size_t E = (size_t)Ne[0] * Ne[1] * Ne[2];
    MLayout L = mlayout<N, DIM>(E);
    char* W = (char*)ws;
    float* trA[3] = {nullptr, nullptr, nullptr};
    float* trB[3] = {nullptr, nullptr, nullptr};
    for (int d = 0; d < DIM; d++) {
        trA[d] = (float*)(W + L.trA + d * L.tr_axis);
        trB[d] = (float*)(W + L.trB + d * L.tr_axis);
    }
    MArgs a;
    std::memset(&a, 0, sizeof(a));
    a.U_sp = U_sp;
    a.mu = (const double*)(W + L.mu);
    a.d_old = (const float*)(W + L.dold);
    a.D = (float*)(W + L.D);
    a.part = (double*)(W + L.part);
    a.res_part = (double*)(W + L.resp);
    a.S2F = c->mixed_mats;
    a.Dh = c->mixed_mats + (size_t)c->m * c->n;
    a.S2C = c->mat[SPDK_M_SP_TO_CV];
    const std::vector<double>& inv = c->mat_host[SPDK_M_INVADER];
    for (int t = 0; t < T; t++) {
        for (int l = 0; l < T; l++) a.cdt[t * kMaxTm + l] = (float)(dt * inv[t * T + l] * c->wt_host[l]);
        a.wdt[t] = dt * c->wt_host[t];
    }
    for (int d = 0; d < 3; d++) {
        a.Ne[d] = Ne[d];
        a.act[d] = c->prm.active[d];
        a.bc[d] = c->prm.bc[d];
        a.h[d] = c->prm.h[d];
        a.ih[d] = 1.0 / c->prm.h[d];
    }
    double wsum = 0.0;
    sp_weights(N, a.wq);
    for (int i = 0; i < N; i++) wsum += a.wq[i];
    a.wtot = wsum;
    for (int d = 1; d < DIM; d++) a.wtot *= wsum;
    a.E = (int)E;
    a.gm = c->prm.gamma;
    a.cfl = c->prm.cfl;
    for (int d = 0; d < 3; d++) a.g[d] = c->prm.g[d];
    a.grav = (a.g[0] != 0.0 || a.g[1] != 0.0 || a.g[2] != 0.0);
    a.tol = pc ? pc->tol : 0.0;
    unsigned char* conv = (unsigned char*)(W + L.conv);
    unsigned char* big[2] = {(unsigned char*)(W + L.big), (unsigned char*)(W + L.big) + E};
    double* red2 = (double*)(W + L.red2);
    a.ver = (int*)(W + L.ver);
    a.cnt_part = (double*)(W + L.cntp);
    for (int d = 0; d < 3; d++) { a.trb[0][d] = trA[d]; a.trb[1][d] = trB[d]; }
    dim3 grid(L.blocks), block(G::BLOCK);
    const size_t lds = G::LDS_BYTES;

    // nb blocks: every block, or the length of args.blist
    // timing covers the Picard sweeps (predictor and corrector), not the
    // face-integral flux evaluation
    auto sweep = [&](const MArgs& args, unsigned nb) -> int {
        bool rec = c->timing && c->ev_used < spdk_ctx::kEvPool && !args.Ffv[0][0];
        if (rec) SPDK_CHECK(hipEventRecord(c->evs[c->ev_used][0], c->stream));
        if (args.Ffv[0][0])
            hipLaunchKernelGGL((k_msweep<N, DIM, 2>), dim3(args.blist ? nb : (nb + 7) / 8 * 8), block, lds, c->stream,
                               args);
        else if (!args.out_pred || args.U_new)
            hipLaunchKernelGGL((k_msweep<N, DIM, 1>), dim3(nb), block, lds, c->stream, args);
        else
            hipLaunchKernelGGL((k_msweep<N, DIM, 0>), dim3(nb), block, lds, c->stream, args);
        SPDK_LAUNCH_CHECK("k_msweep");
        if (rec) SPDK_CHECK(hipEventRecord(c->evs[c->ev_used++][1], c->stream));
        return SPDK_OK;
    };
    auto start = [&]() -> int {
        hipLaunchKernelGGL((k_mstart<N, DIM>), grid, block, lds, c->stream, (const double*)U_sp, (double*)(W + L.mu),
                           (float*)(W + L.dold), trA[0], trA[1], trA[2], (const float*)a.S2F, conv, a.ver, (int)E);
        SPDK_LAUNCH_CHECK("k_mstart");
        return SPDK_OK;
    };
    // reductions of the nb per-block partials and, given the flags of the
    // last sweep, the block list of the next one; blocking read into
    // red_host[0..2] and the list length into nlist
    int* blist = (int*)(W + L.blist);
    int* counter = (int*)(red2 + 4);
    unsigned nlist = 0;
    auto reduce = [&](const double* mn, const double* mx, const double* sm, unsigned nb,
                      const unsigned char* flags) -> int {
        hipLaunchKernelGGL(k_mreduce, dim3(1), dim3(1024), 0, c->stream, mn, mx, sm, (int)nb, red2,
                           flags ? counter : nullptr);
        SPDK_LAUNCH_CHECK("k_mreduce");
        if (flags) {
            MArgs fa = a;
            hipLaunchKernelGGL((k_mblocks<N, DIM>), dim3((L.blocks + 255) / 256), dim3(256), 0, c->stream, fa, flags,
                               (int)L.blocks, blist, counter);
            SPDK_LAUNCH_CHECK("k_mblocks");
        }
        SPDK_CHECK(hipMemcpyAsync(c->red_host, red2, 5 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        SPDK_CHECK(hipStreamSynchronize(c->stream));
        int n = 0;
        std::memcpy(&n, c->red_host + 4, sizeof(int));
        nlist = flags ? (unsigned)n : 0;
        return SPDK_OK;
    };
    // fixed count: p predictor sweeps from the zero-order guess, then the
    // corrector (left to the caller when predict_only)
    int cur = 0;  // newest trace version
    auto fixed_count = [&]() -> int {
        for (int k = 0; k < (predict_only ? T - 1 : T); k++) {
            MArgs s = a;
            const bool last = !predict_only && k == T - 1;
            s.cur = k + 1;
            s.path = k > 0;
            s.tin_slices = k == 0 ? 1 : T;
            s.out_pred = !last;
            s.U_new = last ? U_sp : nullptr;
            s.W_cv = last ? W_cv : nullptr;
            s.F_x = last ? F_x : nullptr;
            s.sweep = k;
            cur = k + 1;
            int r;
            if ((r = sweep(s, L.blocks))) return r;
        }
        return SPDK_OK;
    };

    const bool learned = pc && pc->guess == SPDK_GUESS_MLP;
    const bool external = pc && (pc->guess == SPDK_GUESS_EXTERNAL || learned);
    if (!(external && !learned) && (rc = start())) return rc;  // an external guess follows spdk_mixed_prepare
    if (learned && (rc = mlp_guess<N, DIM>(c, ws, E, dt))) return rc;
    double sweeps = 0;  // element-weighted: a sweep over a fraction f of the elements counts f
    if (pc && pc->residual) {
        const int cap_fixed = T - 1;  // the fixed count of predictor sweeps
        int cap = pc->max_sweeps > 0 ? std::min(pc->max_sweeps, 250) : cap_fixed;
        if (external) {
            // the caller's guess is in the predictor buffer: its traces, version 0
            hipLaunchKernelGGL((k_mtrace<N, DIM>), grid, block, lds, c->stream, (const float*)a.D, trA[0], trA[1],
                               trA[2], (const float*)a.S2F, a.ver, (int)E);
            SPDK_LAUNCH_CHECK("k_mtrace");
        }
        bool guess_phase = external;
        int ksw = 0;                           // sweeps this step (histogram index)
        int kph = 0;                           // sweeps in this phase
        const unsigned char* flags = nullptr;  // unconverged after the last sweep (null: all)
        bool capped = false;
        unsigned nb = L.blocks;                // blocks of the next sweep: all, or the list
        const int* list = nullptr;
        for (;;) {
            MArgs sa = a;
            sa.cur = ++cur;
            sa.path = (ksw > 0 || external) ? 1 : 0;
            sa.tin_slices = sa.path ? T : 1;
            sa.out_pred = 1;
            sa.conv = conv;
            sa.sweep = ksw;
            sa.big_in = flags;
            sa.big_out = big[ksw % 2];
            sa.blist = list;
            if (list) {  // elements of blocks left out keep no stale flag
                hipLaunchKernelGGL(k_fill_u8, dim3(256), dim3(256), 0, c->stream, big[ksw % 2], E, (unsigned char)0);
                SPDK_LAUNCH_CHECK("k_fill_u8");
            }
            if ((rc = sweep(sa, nb))) return rc;
            if ((rc = reduce(nullptr, a.res_part, a.cnt_part, nb, big[ksw % 2]))) return rc;
            sweeps += c->red_host[2] / (double)E;
            if (st) st->last_residual = c->red_host[1];
            flags = big[ksw % 2];
            ksw++;
            kph++;
            if (c->red_host[1] < pc->tol || nlist == 0) break;
            if (kph < cap && ksw < 250) {
                list = blist;
                nb = nlist;
                continue;
            }
            if (guess_phase && !pc->guess_keep) {
                // unconverged from the guess: restart those elements from the
                // zero-order guess with the fixed count
                hipLaunchKernelGGL((k_mrestart<N, DIM>), grid, block, lds, c->stream, (const float*)a.d_old, a.D,
                                   trA[0], trA[1], trA[2], trB[0], trB[1], trB[2], (const float*)a.S2F, a.ver, flags,
                                   big[ksw % 2], ++cur, (int)E);
                SPDK_LAUNCH_CHECK("k_mrestart");
                flags = big[ksw % 2];
                ksw++;
                if ((rc = reduce(nullptr, nullptr, nullptr, 0, flags))) return rc;
                if (nlist == 0) break;
                list = blist;
                nb = nlist;
                guess_phase = false;
                kph = 0;
                cap = cap_fixed;
                capped = true;
                continue;
            }
            capped = true;  // the rest keep the result of the cap
            break;
        }
        if (capped && st) st->fallbacks++;
        hipLaunchKernelGGL(k_mhist, dim3(256), dim3(256), 0, c->stream, (const unsigned char*)conv, (int)E,
                           kHistBins, (unsigned long long*)(W + L.hist));
        SPDK_LAUNCH_CHECK("k_mhist");
        if (!predict_only) {
            MArgs sa = a;  // the corrector over every element
            sa.cur = ++cur;
            sa.path = 1;
            sa.tin_slices = T;
            sa.out_pred = 0;
            sa.U_new = U_sp;
            sa.W_cv = W_cv;
            sa.F_x = F_x;
            sa.sweep = ksw;
            if ((rc = sweep(sa, L.blocks))) return rc;
        }
        sweeps += 1.0;  // the corrector, or the caller's flux evaluation of the predictor
    } else {
        if ((rc = fixed_count())) return rc;
        sweeps += T;
    }
    if (fvo) {  // the flux evaluation of the final slices with their face integrals, every element
        MArgs sa = a;
        sa.cur = ++cur;
        sa.path = 1;
        sa.tin_slices = T;
        sa.out_pred = 0;
        sa.F_x = F_x;
        for (int t = 0; t < T; t++)
            for (int d = 0; d < 3; d++) sa.Ffv[t][d] = fvo->Ffv[t * 3 + d];
        for (int d = 0; d < 3; d++) { sa.fvn[d] = fvo->fvn[d]; sa.fgh[d] = fvo->fgh[d]; }
        sa.SS = c->mixed_mats + (size_t)c->m * c->n + 3 * (size_t)c->n * c->m;
        sa.fv_valu = fvo->valu;
        // no residual or activity partials: the grid is padded to a multiple
        // of 8 workgroups (XCD order), beyond the partial arrays' length
        sa.res_part = nullptr;
        sa.cnt_part = nullptr;
        sa.big_in = nullptr;
        sa.big_out = nullptr;
        sa.blist = nullptr;
        if ((rc = sweep(sa, L.blocks))) return rc;
    }
    if (!predict_only) {
        if ((rc = reduce(a.part, nullptr, nullptr, L.blocks, nullptr))) return rc;
        if (dt_next) *dt_next = c->red_host[0];
    }
    if (st) {
        st->steps++;
        st->sweeps += sweeps;
    }
    if (c->ev_used > 0) {  // harvest the sweep timings (a predictor-only call has not synchronised yet)
        SPDK_CHECK(hipEventSynchronize(c->evs[c->ev_used - 1][1]));
        for (int i = 0; i < c->ev_used; i++) {
            float ms = 0;
            SPDK_CHECK(hipEventElapsedTime(&ms, c->evs[i][0], c->evs[i][1]));
            c->kernel_ms += ms;
            c->kernel_launches++;
        }
        c->ev_used = 0;
    }
    return SPDK_OK;
}

template <int N, int DIM>
int run_prepare(spdk_ctx* c, const double* U_sp, const int Ne[3], void* ws) {
    using G = MT<N, DIM>;
    int rc;
    if ((rc = ensure_mixed_mats(c))) return rc;
    const size_t E = (size_t)Ne[0] * Ne[1] * Ne[2];
    MLayout L = mlayout<N, DIM>(E);
    char* W = (char*)ws;
    float* tr[3] = {nullptr, nullptr, nullptr};
    for (int d = 0; d < DIM; d++) tr[d] = (float*)(W + L.trA + d * L.tr_axis);
    hipLaunchKernelGGL((k_mstart<N, DIM>), dim3(L.blocks), dim3(G::BLOCK), G::LDS_BYTES, c->stream, U_sp,
                       (double*)(W + L.mu), (float*)(W + L.dold), tr[0], tr[1], tr[2], (const float*)c->mixed_mats,
                       (unsigned char*)(W + L.conv), (int*)(W + L.ver), (int)E);
    SPDK_LAUNCH_CHECK("k_mstart");
    return SPDK_OK;
}

// p' predictor sweeps from the zero-order guess (no corrector): the
// predictor buffer then holds the p'-th Picard iterate.
template <int N, int DIM>
int run_iterate(spdk_ctx* c, double* U_sp, const int Ne[3], void* ws, double dt, int nsweeps) {
    spdk_picard pc{0, SPDK_GUESS_ZERO, 0.0, nsweeps};
    (void)pc;
    using G = MT<N, DIM>;
    int rc;
    if ((rc = run_prepare<N, DIM>(c, U_sp, Ne, ws))) return rc;
    const size_t E = (size_t)Ne[0] * Ne[1] * Ne[2];
    MLayout L = mlayout<N, DIM>(E);
    char* W = (char*)ws;
    MArgs a;
    std::memset(&a, 0, sizeof(a));
    a.U_sp = U_sp;
    a.mu = (const double*)(W + L.mu);
    a.d_old = (const float*)(W + L.dold);
    a.D = (float*)(W + L.D);
    a.part = (double*)(W + L.part);
    a.res_part = (double*)(W + L.resp);
    a.S2F = c->mixed_mats;
    a.Dh = c->mixed_mats + (size_t)c->m * c->n;
    a.S2C = c->mat[SPDK_M_SP_TO_CV];
    a.ver = (int*)(W + L.ver);
    const std::vector<double>& inv = c->mat_host[SPDK_M_INVADER];
    for (int t = 0; t < G::T; t++) {
        for (int l = 0; l < G::T; l++) a.cdt[t * kMaxTm + l] = (float)(dt * inv[t * G::T + l] * c->wt_host[l]);
        a.wdt[t] = dt * c->wt_host[t];
    }
    for (int d = 0; d < 3; d++) {
        a.Ne[d] = Ne[d];
        a.act[d] = c->prm.active[d];
        a.bc[d] = c->prm.bc[d];
        a.h[d] = c->prm.h[d];
        a.ih[d] = 1.0 / c->prm.h[d];
        a.trb[0][d] = DIM > d ? (float*)(W + L.trA + d * L.tr_axis) : nullptr;
        a.trb[1][d] = DIM > d ? (float*)(W + L.trB + d * L.tr_axis) : nullptr;
    }
    double wsum = 0.0;
    sp_weights(N, a.wq);
    for (int i = 0; i < N; i++) wsum += a.wq[i];
    a.wtot = wsum;
    for (int d = 1; d < DIM; d++) a.wtot *= wsum;
    a.E = (int)E;
    a.gm = c->prm.gamma;
    a.cfl = c->prm.cfl;
    for (int d = 0; d < 3; d++) a.g[d] = c->prm.g[d];
    a.grav = (a.g[0] != 0.0 || a.g[1] != 0.0 || a.g[2] != 0.0);
    for (int k = 0; k < nsweeps; k++) {
        MArgs s = a;
        s.cur = k + 1;
        s.path = k > 0;
        s.tin_slices = k == 0 ? 1 : G::T;
        s.out_pred = 1;
        s.sweep = k;
        hipLaunchKernelGGL((k_msweep<N, DIM, 0>), dim3(L.blocks), dim3(G::BLOCK), G::LDS_BYTES, c->stream, s);
        SPDK_LAUNCH_CHECK("k_msweep");
    }
    return SPDK_OK;
}

typedef int (*mstep_fn)(spdk_ctx*, double*, const int*, void*, double, double*, double*, const spdk_picard*,
                        spdk_picard_stats*, double*, bool, const FvOut*);
typedef MLayout (*mlayout_fn)(size_t);
typedef int (*mprep_fn)(spdk_ctx*, const double*, const int*, void*);
typedef int (*miter_fn)(spdk_ctx*, double*, const int*, void*, double, int);
typedef int (*mdata_fn)(spdk_ctx*, void*, size_t, double, int, int, float*, long, float*, long);
struct MVariant {
    int n, dim;
    mstep_fn step;
    mlayout_fn layout;
    mprep_fn prepare;
    miter_fn iterate;
    mdata_fn data;
};
#define SPDK_MVARIANT(Nn, D) \
    {Nn, D, run_mixed<Nn, D>, mlayout<Nn, D>, run_prepare<Nn, D>, run_iterate<Nn, D>, guess_data<Nn, D>}
const MVariant kMVariants[] = {SPDK_MVARIANT(2, 1), SPDK_MVARIANT(3, 1), SPDK_MVARIANT(4, 1), SPDK_MVARIANT(5, 1),
                               SPDK_MVARIANT(2, 2), SPDK_MVARIANT(3, 2), SPDK_MVARIANT(4, 2), SPDK_MVARIANT(5, 2),
                               SPDK_MVARIANT(2, 3), SPDK_MVARIANT(3, 3), SPDK_MVARIANT(4, 3)};

const MVariant* find_mvariant(spdk_ctx* c) {
    int dim = c->prm.active[0] + c->prm.active[1] + c->prm.active[2];
    if (!c->prm.active[0] || (c->prm.active[2] && !c->prm.active[1])) return nullptr;
    if (c->prm.rk_order != 0 || c->T != c->n) return nullptr;
    for (const MVariant& v : kMVariants)
        if (v.n == c->n && v.dim == dim) return &v;
    return nullptr;
}

}  // namespace

extern "C" {

size_t spdk_mixed_workspace_bytes(spdk_ctx* c, const int N[3]) {
    const MVariant* v = c ? find_mvariant(c) : nullptr;
    if (!v) {
        set_error("mixed path: ADER hydro with a tile variant for this degree/dimensionality only");
        return 0;
    }
    if (c->prm.nu != 0.0) {
        set_error("mixed path covers the inviscid update (gravity included)");
        return 0;
    }
    return v->layout((size_t)N[0] * N[1] * N[2]).total;
}

int spdk_mixed_predictor_buffer(spdk_ctx* c, void* ws, const int N[3], float** D, double** mu, float** d_old) {
    const MVariant* v = c ? find_mvariant(c) : nullptr;
    if (!v || !ws) { set_error("mixed path: no variant or workspace"); return SPDK_E_CONFIG; }
    MLayout L = v->layout((size_t)N[0] * N[1] * N[2]);
    if (D) *D = (float*)((char*)ws + L.D);
    if (mu) *mu = (double*)((char*)ws + L.mu);
    if (d_old) *d_old = (float*)((char*)ws + L.dold);
    return SPDK_OK;
}

static int check_mixed(spdk_ctx* c, const void* U_sp, const spdk_sd_shape& s, const void* ws, const MVariant** out);

int spdk_mixed_export_predictor(spdk_ctx* c, void* ws, double* U_ader, spdk_sd_shape s) {
    const MVariant* v = c ? find_mvariant(c) : nullptr;
    if (!v || !ws || !U_ader) { set_error("mixed export: no variant, workspace or output"); return SPDK_E_ARG; }
    const int tpe = v->dim == 1 ? v->n : (v->dim == 2 ? v->n * v->n : v->n * v->n * v->n);
    if (s.nt != c->T || s.nv < NVP || (size_t)s.nz * s.ny * s.nx != (size_t)tpe) {
        set_error("mixed export: U_ader shape (T, >= 5 variables, solution points)");
        return SPDK_E_SHAPE;
    }
    const size_t E = (size_t)s.Nx * s.Ny * s.Nz;
    MLayout L = v->layout(E);
    const size_t total = (size_t)s.nt * s.nv * E * tpe;
    hipLaunchKernelGGL(k_mexport, dim3((unsigned)std::min<size_t>((total + 255) / 256, 65536)), dim3(256), 0, c->stream,
                       (const float*)((char*)ws + L.D), (const double*)((char*)ws + L.mu), U_ader, s.nt, s.nv,
                       E * tpe, (size_t)tpe, total);
    SPDK_LAUNCH_CHECK("k_mexport");
    return SPDK_OK;
}

int spdk_mixed_guess_dims(spdk_ctx* c, int* din, int* dout) {
    const MVariant* v = c ? find_mvariant(c) : nullptr;
    if (!v) { set_error("mixed path: no variant"); return SPDK_E_CONFIG; }
    const int dim = v->dim, tpe = dim == 1 ? v->n : (dim == 2 ? v->n * v->n : v->n * v->n * v->n);
    if (din) *din = (dim + 2) * tpe + kGExtra;
    if (dout) *dout = v->n * (dim + 2) * tpe;
    return SPDK_OK;
}

int spdk_mixed_attach_mlp(spdk_ctx* c, spdk_mlp* net) {
    if (!c) return SPDK_E_ARG;
    if (net) {
        int din = 0, dout = 0, rc = spdk_mixed_guess_dims(c, &din, &dout);
        if (rc) return rc;
        int pd[4];
        spdk_mlp_dims(net, pd);
        if (pd[0] != (din + 63) / 64 * 64 || pd[3] != (dout + 63) / 64 * 64) {
            set_error("mixed: the network's input/output sizes do not match this degree and dimension");
            return SPDK_E_SHAPE;
        }
    }
    if (net) {
        int rc = spdk_mlp_set_stream(net, c->stream);  // ordered with the step's kernels
        if (rc) return rc;
    }
    c->guess_net = net;
    return SPDK_OK;
}

int spdk_mixed_iterate(spdk_ctx* c, double* U_sp, spdk_sd_shape s, void* ws, double dt, int nsweeps) {
    const MVariant* v = nullptr;
    int rc = check_mixed(c, U_sp, s, ws, &v);
    if (rc) return rc;
    if (nsweeps < 1 || nsweeps > 1000) { set_error("mixed: sweeps outside 1..1000"); return SPDK_E_ARG; }
    int Ne[3] = {s.Nx, s.Ny, s.Nz};
    return v->iterate(c, U_sp, Ne, ws, dt, nsweeps);
}

int spdk_mixed_guess_data(spdk_ctx* c, void* ws, const int N[3], double dt, int e0, int nrows, float* X, long ldx,
                          float* Y, long ldy) {
    const MVariant* v = c ? find_mvariant(c) : nullptr;
    if (!v || !ws || !N) { set_error("mixed guess data: bad argument"); return SPDK_E_ARG; }
    const size_t E = (size_t)N[0] * N[1] * N[2];
    if (e0 < 0 || nrows < 0 || (size_t)e0 + nrows > E) { set_error("mixed guess data: rows outside the mesh"); return SPDK_E_ARG; }
    return v->data(c, ws, E, dt, e0, nrows, X, ldx, Y, ldy);
}

int spdk_mixed_histogram(spdk_ctx* c, void* ws, const int N[3], unsigned long long* hist, int nbins, int reset) {
    const MVariant* v = c ? find_mvariant(c) : nullptr;
    if (!v || !ws || nbins < 0) { set_error("mixed histogram: bad argument"); return SPDK_E_ARG; }
    MLayout L = v->layout((size_t)N[0] * N[1] * N[2]);
    unsigned long long* d = (unsigned long long*)((char*)ws + L.hist);
    if (hist && nbins > 0) {
        unsigned long long tmp[kHistBins] = {};
        SPDK_CHECK(hipMemcpyAsync(tmp, d, sizeof(tmp), hipMemcpyDeviceToHost, c->stream));
        SPDK_CHECK(hipStreamSynchronize(c->stream));
        for (int i = 0; i < nbins; i++) hist[i] = i < kHistBins ? tmp[i] : 0;
    }
    if (reset) SPDK_CHECK(hipMemsetAsync(d, 0, sizeof(unsigned long long) * kHistBins, c->stream));
    return SPDK_OK;
}

static int check_mixed(spdk_ctx* c, const void* U_sp, const spdk_sd_shape& s, const void* ws, const MVariant** out) {
    if (!c || !U_sp || !ws) { set_error("mixed: null argument"); return SPDK_E_ARG; }
    if (c->prm.nu != 0.0) {
        set_error("mixed path covers the inviscid update (gravity included)");
        return SPDK_E_CONFIG;
    }
    const MVariant* v = find_mvariant(c);
    if (!v) { set_error("mixed path: ADER hydro with a tile variant for this degree/dimensionality only"); return SPDK_E_CONFIG; }
    if (s.nv != NV || s.nt != 1 || s.nx != c->n || s.ny != (c->prm.active[1] ? c->n : 1) ||
        s.nz != (c->prm.active[2] ? c->n : 1)) {
        set_error("mixed: U_sp shape");
        return SPDK_E_SHAPE;
    }
    for (int d = 0; d < 3; d++) {
        int Nd = d == 0 ? s.Nx : (d == 1 ? s.Ny : s.Nz);
        if (c->prm.active[d] && Nd < 3) { set_error("mixed: axis needs ghost elements"); return SPDK_E_SHAPE; }
    }
    if ((uint64_t)s.Nx * s.Ny * s.Nz >= (1ull << 31)) { set_error("mixed: too many elements"); return SPDK_E_SHAPE; }
    *out = v;
    return SPDK_OK;
}

int spdk_mixed_prepare(spdk_ctx* c, const double* U_sp, spdk_sd_shape s, void* ws) {
    const MVariant* v = nullptr;
    int rc = check_mixed(c, U_sp, s, ws, &v);
    if (rc) return rc;
    int Ne[3] = {s.Nx, s.Ny, s.Nz};
    return v->prepare(c, U_sp, Ne, ws);
}

static int check_picard(spdk_ctx* c, const spdk_picard* pc) {
    if (pc && pc->residual && !(pc->tol > 0.0)) { set_error("mixed: residual tolerance must be positive"); return SPDK_E_ARG; }
    if (pc && pc->guess != SPDK_GUESS_ZERO && pc->guess != SPDK_GUESS_EXTERNAL && pc->guess != SPDK_GUESS_MLP) {
        set_error("mixed: unknown guess");
        return SPDK_E_ARG;
    }
    if (pc && pc->guess != SPDK_GUESS_ZERO && !pc->residual) {
        set_error("mixed: a guess other than the zero-order one needs the residual test");
        return SPDK_E_ARG;
    }
    if (pc && pc->guess == SPDK_GUESS_MLP && !c->guess_net) {
        set_error("mixed: no network attached (spdk_mixed_attach_mlp)");
        return SPDK_E_CONFIG;
    }
    return SPDK_OK;
}

int spdk_hydro_step_mixed(spdk_ctx* c, double* U_sp, spdk_sd_shape s, void* ws, double dt, double* W_cv, double* F_x,
                          const spdk_picard* pc, spdk_picard_stats* stats, double* dt_next) {
    const MVariant* v = nullptr;
    int rc0 = check_mixed(c, U_sp, s, ws, &v);
    if (rc0 || (rc0 = check_picard(c, pc))) return rc0;
    int Ne[3] = {s.Nx, s.Ny, s.Nz};
    std::strncpy(c->kernel_name, "k_msweep", sizeof(c->kernel_name) - 1);
    return v->step(c, U_sp, Ne, ws, dt, W_cv, F_x, pc, stats, dt_next, false, nullptr);
}

int spdk_mixed_predict_fv_fluxes(spdk_ctx* c, double* U_sp, spdk_sd_shape s, void* ws, double dt,
                                 const spdk_picard* pc, spdk_picard_stats* stats, double* const* F_fv,
                                 spdk_fv_shape sc, double* F_x, int valu) {
    const MVariant* v = nullptr;
    int rc0 = check_mixed(c, U_sp, s, ws, &v);
    if (rc0 || (rc0 = check_picard(c, pc))) return rc0;
    if (!F_fv) { set_error("mixed fv fluxes: null face arrays"); return SPDK_E_ARG; }
    const int Ne[3] = {s.Nx, s.Ny, s.Nz}, nc[3] = {sc.Nx, sc.Ny, sc.Nz};
    FvOut fo;
    fo.Ffv = F_fv;
    fo.valu = valu;
    for (int d = 0; d < 3; d++) {
        const bool act = c->prm.active[d] != 0;
        const bool ok = act ? (c->ngh[d] == 1 && nc[d] == (Ne[d] - 2) * c->n + 2 * c->fgh[d]) : nc[d] == 1;
        if (!ok || sc.nv != NV) {
            set_error("mixed fv fluxes: FV cell shape disagrees with U_sp along axis " + std::to_string(d));
            return SPDK_E_SHAPE;
        }
        for (int t = 0; t < c->T; t++)
            if (act && !F_fv[t * 3 + d]) { set_error("mixed fv fluxes: null face array"); return SPDK_E_ARG; }
        fo.fvn[d] = nc[d];
        fo.fgh[d] = c->fgh[d];
    }
    std::strncpy(c->kernel_name, "k_msweep", sizeof(c->kernel_name) - 1);
    return v->step(c, U_sp, Ne, ws, dt, nullptr, F_x, pc, stats, nullptr, true, &fo);
}

int spdk_mixed_predict(spdk_ctx* c, double* U_sp, spdk_sd_shape s, void* ws, double dt, const spdk_picard* pc,
                       spdk_picard_stats* stats) {
    const MVariant* v = nullptr;
    int rc0 = check_mixed(c, U_sp, s, ws, &v);
    if (rc0 || (rc0 = check_picard(c, pc))) return rc0;
    int Ne[3] = {s.Nx, s.Ny, s.Nz};
    std::strncpy(c->kernel_name, "k_msweep", sizeof(c->kernel_name) - 1);
    return v->step(c, U_sp, Ne, ws, dt, nullptr, nullptr, pc, stats, nullptr, true, nullptr);
}

}  // extern "C"
