// Fused ADER step of the induction solver without the FV update: the stage
// sequence of src/induction_ader.hpp:343-359 (copy_ader, p+1 x Solve_E with
// p predictor updates, then update_B_solution) in p+2 launches, bit-exact
// with the stage kernels.
//
// What crosses HBM between launches is the raw edge field E of each element
// (compute_E's first variable, src/induction.cpp:139-208, before any face
// solve), FP64, for all time slices. A sweep launch works on a tile of
// consecutive elements (along x), slice by slice:
//   A. the tile's raw E of the slice goes to LDS; every edge point gets its
//      face-solved value from its own raw E and its neighbours' raw E at the
//      shared faces (from LDS inside the tile, else HBM), upwinded along the
//      first face axis and then the second, exactly as the in-place
//      E_riemann_solver calls do (src/induction.cpp:509-560,
//      src/induction_ader.hpp:437-448); boundary ghosts are resolved by index
//      (periodic wrap, gradfree mirror) as the SD trace fill makes them;
//   B. dB/dt of the slice at every face-B point (the derivative sums of
//      src/induction.cpp:331-356), kept in registers for all slices;
// then either the corrector into B (last launch, update_B_solution) or, slice
// by slice, the predictor slice in LDS (update_B_prediction) and its raw E,
// written for the next launch.
// The zero-order guess makes every slice equal, so the first launch only
// computes the raw E of B once per point and writes it to every slice.
//
// The same kernels with R = float are the mixed-precision form: raw E, the
// face solves, the sp->fp interpolations and the predictor slices in FP32
// (with dfp_to_sp/h and invader*w*dt premultiplied), while the corrector forms
// dB/dt in FP64 from the FP32 edge values and accumulates B in FP64. Since
// every face-B update is a discrete curl of one set of edge values, div B
// stays at FP64 round-off.
//
// Only the upwind velocities of the face solves matter: they are the uniform
// advection velocity of the run (compute_E stores it per point, the face
// solve takes max(v_L, v_R) = v), and the other variables the stage path
// carries never reach B.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstring>
#include <type_traits>
#include "ctx.hpp"

// Element tiles of the induction sweep in XCD-contiguous order (0: launch order).
#ifndef SPDK_ISWEEP_XCD
#define SPDK_ISWEEP_XCD 1
#endif

using namespace spdk;

namespace {

constexpr int BS = 256;
constexpr int kMaxT = 5;
constexpr int kLdsBudget = 40 * 1024;  // bytes of element tiles per workgroup

// Point counts of one configuration: N solution points per active axis; D = 3,
// or D = 2 with z inactive (one point along z).
template <int N, int D>
struct IG {
    static constexpr int M = N + 1;
    __host__ __device__ static constexpr int sp(int a) { return a < D ? N : 1; }
    __host__ __device__ static constexpr int fp(int a) { return a < D ? M : 1; }
    // face-B component d: flux points along d, solution points elsewhere
    __host__ __device__ static constexpr int bcnt(int d, int a) { return a == d ? fp(a) : sp(a); }
    // edge-E component e: solution points along e, flux points elsewhere
    __host__ __device__ static constexpr int ecnt(int e, int a) { return a == e ? sp(a) : fp(a); }
    __host__ __device__ static constexpr int bpts(int d) { return bcnt(d, 0) * bcnt(d, 1) * bcnt(d, 2); }
    __host__ __device__ static constexpr int epts(int e) { return ecnt(e, 0) * ecnt(e, 1) * ecnt(e, 2); }
    static constexpr int BP = bpts(0) + bpts(1) + bpts(2);
    static constexpr int EP = epts(0) + epts(1) + epts(2);
    __host__ __device__ static constexpr int boff(int d) { return d == 0 ? 0 : (d == 1 ? bpts(0) : bpts(0) + bpts(1)); }
    __host__ __device__ static constexpr int eoff(int e) { return e == 0 ? 0 : (e == 1 ? epts(0) : epts(0) + epts(1)); }
};

template <int N, int D, int T, class R>
struct Tile {
    using G = IG<N, D>;
    // per element: raw and face-solved E of every slice (one predictor slice
    // reuses the raw buffer)
    static constexpr int PER = 2 * T * G::EP;
    // LDS bound, and at most two face-B points per thread (their dB/dt of
    // every slice stay in registers)
    static constexpr int EPB_LDS = kLdsBudget / (int)(PER * sizeof(R));
    static constexpr int EPB_REG = 2 * BS / G::BP;
    static constexpr int EPB_RAW = EPB_LDS < EPB_REG ? EPB_LDS : EPB_REG;
    static constexpr int EPB = EPB_RAW < 1 ? 1 : (EPB_RAW > 16 ? 16 : EPB_RAW);
    static constexpr int NB = (EPB * G::BP + BS - 1) / BS;  // face-B points per thread
};

// Edge points of one element (all components) ordered by how many faces
// they lie on (none, the first face axis only, the second only, both) and by
// component, so that the lanes of a wave mostly take the same path in the
// face solves.
template <int N, int D>
struct PermT {
    int p[IG<N, D>::EP];
    constexpr PermT() : p() {
        using G = IG<N, D>;
        int k = 0;
        for (int key = 0; key < 12; key++) {
            const int cc = key / 3, e = key % 3;
            int ax[2] = {-1, -1}, nax = 0;
            for (int c = 0; c < 3; c++)
                if (c != e && c < D) ax[nax++] = c;
            for (int lp = 0; lp < G::epts(e); lp++) {
                const int pt[3] = {lp % G::ecnt(e, 0), (lp / G::ecnt(e, 0)) % G::ecnt(e, 1),
                                   lp / (G::ecnt(e, 0) * G::ecnt(e, 1))};
                int code = 0;
                for (int i = 0; i < nax; i++)
                    if (pt[ax[i]] == 0 || pt[ax[i]] == G::ecnt(e, ax[i]) - 1) code |= 1 << i;
                if (code == cc) p[k++] = G::eoff(e) + lp;
            }
        }
    }
};
template <int N, int D>
__constant__ PermT<N, D> kPerm = PermT<N, D>();
template <int N, int D>
constexpr bool perm_is_permutation() {
    constexpr int EP = IG<N, D>::EP;
    PermT<N, D> P;
    bool seen[EP] = {};
    for (int i = 0; i < EP; i++) {
        if (P.p[i] < 0 || P.p[i] >= EP || seen[P.p[i]]) return false;
        seen[P.p[i]] = true;
    }
    return true;
}

struct IArgs {
    void* Eout;           // raw E written by this launch (T, 3 components, elements, points)
    const void* Ein;      // raw E of the state this launch reads (element type R)
    double* B[3];         // face B, ghost-inclusive layout (1, 1, Nz, Ny, Nx, pts)
    const double* S2F;    // sp_to_fp (M x N)
    const double* Dm;     // dfp_to_sp (N x M)
    const double* inv;    // invader (T x T)
    const double* w;      // time weights (T)
    double vel[3];        // advection velocity
    double h[3];          // element widths
    double dt;
    int n[3];             // active elements per axis (1 on an inactive axis)
    int G[3];             // ghost-inclusive element counts
    int g[3];             // ghost elements per side (1 on active axes)
    int periodic[3];      // 1: periodic wrap, 0: gradfree mirror
    int nel;              // active elements
};

template <class R>
__device__ __forceinline__ R upwind(R l, R r, double v) {
    if (v == 0) return R(0.5) * (l + r);
    return v > 0 ? l : r;
}

// Raw-E slab of component e (all slices): T x nel x epts(e), components
// back to back.
template <int N, int D, int T>
__device__ __forceinline__ size_t eslab(int e, int nel) {
    using G = IG<N, D>;
    size_t o = 0;
    for (int c = 0; c < e; c++) o += (size_t)T * nel * G::epts(c);
    return o;
}

// Element and point that hold the raw E seen at element el, point pt of
// component e, where el may be one past either end of an active axis: that is
// the ghost the SD trace fill (src/boundary.cpp:48-113) makes, i.e. periodic,
// the element at the other end, same point; gradfree, the element itself with
// the point mirrored across its face. Returns the linear element index.
template <int N, int D>
__device__ __forceinline__ int resolve(const IArgs& a, int e, const int el[3], int r[3]) {
    using G = IG<N, D>;
    int q[3] = {el[0], el[1], el[2]};
#pragma unroll
    for (int c = 0; c < D; c++) {
        if (q[c] < 0) {
            if (a.periodic[c]) q[c] = a.n[c] - 1;
            else { q[c] = 0; r[c] = G::ecnt(e, c) - 1 - r[c]; }
        } else if (q[c] >= a.n[c]) {
            if (a.periodic[c]) q[c] = 0;
            else { q[c] = a.n[c] - 1; r[c] = G::ecnt(e, c) - 1 - r[c]; }
        }
    }
    return (q[2] * a.n[1] + q[1]) * a.n[0] + q[0];
}

// E after the face solve along axis `ax` only (or the raw value off that
// axis's faces); raw(e, el, pt) reads a raw value.
template <int N, int D, class R, class Raw>
__device__ __forceinline__ R solved_1(const IArgs& a, int e, const int el[3], const int pt[3], int ax, Raw raw) {
    using G = IG<N, D>;
    const int last = G::ecnt(e, ax) - 1;
    const R own = raw(e, el, pt);
    if (pt[ax] != 0 && pt[ax] != last) return own;
    int el2[3] = {el[0], el[1], el[2]}, pt2[3] = {pt[0], pt[1], pt[2]};
    if (pt[ax] == 0) {
        el2[ax] -= 1;
        pt2[ax] = last;
        return upwind<R>(raw(e, el2, pt2), own, a.vel[ax]);
    }
    el2[ax] += 1;
    pt2[ax] = 0;
    return upwind<R>(own, raw(e, el2, pt2), a.vel[ax]);
}

// Face-solved E of component e: along its first active face axis, then its
// second (x before y before z, the order of the stage calls).
template <int N, int D, class R, class Raw>
__device__ __forceinline__ R solved(const IArgs& a, int e, const int el[3], const int pt[3], Raw raw) {
    using G = IG<N, D>;
    int ax[2] = {0, 0}, nax = 0;
#pragma unroll
    for (int c = 0; c < 3; c++)
        if (c != e && c < D) ax[nax++] = c;
    if (nax == 1) return solved_1<N, D, R>(a, e, el, pt, ax[0], raw);
    const int b = ax[1], last = G::ecnt(e, b) - 1;
    const R own = solved_1<N, D, R>(a, e, el, pt, ax[0], raw);
    if (pt[b] != 0 && pt[b] != last) return own;
    int el2[3] = {el[0], el[1], el[2]}, pt2[3] = {pt[0], pt[1], pt[2]};
    if (pt[b] == 0) {
        el2[b] -= 1;
        pt2[b] = last;
        return upwind<R>(solved_1<N, D, R>(a, e, el2, pt2, ax[0], raw), own, a.vel[b]);
    }
    el2[b] += 1;
    pt2[b] = 0;
    return upwind<R>(own, solved_1<N, D, R>(a, e, el2, pt2, ax[0], raw), a.vel[b]);
}

__device__ __forceinline__ void split_pt(int p, int nx, int ny, int pt[3]) {
    pt[0] = p % nx;
    p /= nx;
    pt[1] = p % ny;
    pt[2] = p / ny;
}

// Raw E (compute_E's first variable) of component e at point pt from B
// given by an accessor bget(d, point[3]): v1*b2 - v2*b1 with b1 = B_{e+1}
// interpolated sp->fp along e+2 and b2 = B_{e+2} along e+1 (an inactive axis
// takes the value as is).
template <int N, int D, class R, class BGet>
__device__ __forceinline__ R edge_e(int e, const int pt[3], const R* S2F, const R vel[3], BGet bget) {
    const int a1 = e == 2 ? 0 : e + 1, a2 = a1 == 2 ? 0 : a1 + 1;
    R b1 = 0, b2 = 0;
    int q[3] = {pt[0], pt[1], pt[2]};
    if (a2 < D) {
        for (int l = 0; l < N; l++) {
            q[a2] = l;
            b1 += bget(a1, q) * S2F[pt[a2] * N + l];
        }
        q[a2] = pt[a2];
    } else {
        b1 = bget(a1, q);
    }
    if (a1 < D) {
        for (int l = 0; l < N; l++) {
            q[a1] = l;
            b2 += bget(a2, q) * S2F[pt[a1] * N + l];
        }
    } else {
        b2 = bget(a2, q);
    }
    return vel[a1] * b2 - vel[a2] * b1;
}

// Zero-order guess: the raw E of B, written to every slice.
template <int N, int D, int T, class R>
__global__ void __launch_bounds__(BS) k_istart(IArgs a) {
    using G = IG<N, D>;
    constexpr int M = G::M;
    __shared__ R sS2F[M * N];
    for (int i = threadIdx.x; i < M * N; i += BS) sS2F[i] = (R)a.S2F[i];
    __syncthreads();
    const R vel[3] = {(R)a.vel[0], (R)a.vel[1], (R)a.vel[2]};
    R* out = (R*)a.Eout;
    const size_t total = (size_t)a.nel * G::EP;
    for (size_t idx = (size_t)blockIdx.x * BS + threadIdx.x; idx < total; idx += (size_t)gridDim.x * BS) {
        int lin = (int)(idx / G::EP), p = (int)(idx % G::EP);
        int e = p < G::eoff(1) ? 0 : (p < G::eoff(2) ? 1 : 2);
        p -= G::eoff(e);
        int el[3] = {lin % a.n[0], (lin / a.n[0]) % a.n[1], lin / (a.n[0] * a.n[1])};
        int pt[3];
        split_pt(p, G::ecnt(e, 0), G::ecnt(e, 1), pt);
        size_t gb = ((size_t)(el[2] + a.g[2]) * a.G[1] + (el[1] + a.g[1])) * a.G[0] + (el[0] + a.g[0]);
        auto bget = [&](int d, const int q[3]) -> R {
            return (R)a.B[d][gb * G::bpts(d) + (q[2] * G::bcnt(d, 1) + q[1]) * G::bcnt(d, 0) + q[0]];
        };
        R v = edge_e<N, D, R>(e, pt, sS2F, vel, bget);
        size_t o = eslab<N, D, T>(e, a.nel) + (size_t)lin * G::epts(e) + p;
        for (int t = 0; t < T; t++) out[o + (size_t)t * a.nel * G::epts(e)] = v;
    }
}

// One Picard sweep (CORR = false: predictor slices and their raw E; CORR =
// true: the corrector into B). The tile's raw E of every slice is staged in
// LDS (face partners inside the tile are read there, others from HBM),
// face-solved into a second LDS buffer and differentiated into the dB/dt of
// the thread's face-B points, kept in registers for all slices; the
// predictor slices then go through the raw buffer one at a time.
template <int N, int D, int T, bool CORR, class R>
__global__ void __launch_bounds__(BS) k_isweep(IArgs a) {
    using G = IG<N, D>;
    using TL = Tile<N, D, T, R>;
    constexpr bool PLAIN = std::is_same<R, double>::value;
    using Acc = typename std::conditional<PLAIN || CORR, double, float>::type;  // dB/dt
    constexpr int EPB = TL::EPB, M = G::M, NB = TL::NB, EPT = T * G::EP;
    static_assert(G::BP <= EPT, "a predictor slice fits in the raw-E buffer");
    __shared__ R sE[EPB * EPT], sS[EPB * EPT];  // (element, slice, point)
    R* const sB = sE;                           // predictor slice, after the face solves
    __shared__ R sS2F[M * N];
    __shared__ double sDm[N * M], sInv[T * T], sW[T];
    __shared__ Acc sDmH[PLAIN ? 1 : D][PLAIN ? 1 : N * M];  // dfp_to_sp / h_axis
    __shared__ float sP[PLAIN ? 1 : T * T];                  // invader * w * dt
    __shared__ int sEl[EPB][3];
    __shared__ int sPerm[G::EP];
    // workgroups go to the 8 XCDs in turn; XCD x takes the x-th contiguous
    // eighth of the element tiles, so a tile's y (and z) neighbours, whose
    // face traces it reads, were mostly written through the same L2. The grid
    // is a multiple of 8; tiles past the end exit (uniform per block).
    const int blk = SPDK_ISWEEP_XCD ? (int)((blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8) : (int)blockIdx.x;
    if (blk * EPB >= a.nel) return;
    for (int i = threadIdx.x; i < G::EP; i += BS) sPerm[i] = kPerm<N, D>.p[i];
    for (int i = threadIdx.x; i < N * M; i += BS) {
        sDm[i] = a.Dm[i];
        sS2F[i] = (R)a.S2F[i];
        if (!PLAIN)
            for (int c = 0; c < D; c++) sDmH[c][i] = (Acc)(a.Dm[i] / a.h[c]);
    }
    if (threadIdx.x < T * T) {
        sInv[threadIdx.x] = a.inv[threadIdx.x];
        if (!PLAIN) sP[threadIdx.x] = (float)(a.inv[threadIdx.x] * a.w[threadIdx.x % T] * a.dt);
    }
    if (threadIdx.x < T) sW[threadIdx.x] = a.w[threadIdx.x];
    const int e0 = blk * EPB;
    const int ne = min(EPB, a.nel - e0);
    if ((int)threadIdx.x < ne) {
        int lin = e0 + threadIdx.x;
        sEl[threadIdx.x][0] = lin % a.n[0];
        sEl[threadIdx.x][1] = (lin / a.n[0]) % a.n[1];
        sEl[threadIdx.x][2] = lin / (a.n[0] * a.n[1]);
    }
    const R* Ein = (const R*)a.Ein;

    // A1. the tile's raw E of every slice (contiguous per component and slice)
#pragma unroll
    for (int e = 0; e < 3; e++)
        for (int t = 0; t < T; t++) {
            const R* src = Ein + eslab<N, D, T>(e, a.nel) + ((size_t)t * a.nel + e0) * G::epts(e);
            for (int i = threadIdx.x; i < ne * G::epts(e); i += BS) {
                int q = i / G::epts(e), p = i % G::epts(e);
                sE[(q * T + t) * G::EP + G::eoff(e) + p] = src[i];
            }
        }
    __syncthreads();
    // A2. face solves of every slice; items run point-major over the tile's
    // elements, points in class order within a slice
    for (int idx = threadIdx.x; idx < ne * EPT; idx += BS) {
        int rank = idx / ne, q = idx - rank * ne;
        int t = rank / G::EP;
        int p = sPerm[rank - t * G::EP];
        int e = p < G::eoff(1) ? 0 : (p < G::eoff(2) ? 1 : 2);
        int pt[3];
        split_pt(p - G::eoff(e), G::ecnt(e, 0), G::ecnt(e, 1), pt);
        auto raw = [&](int ec, const int el[3], const int pc[3]) -> R {
            int r[3] = {pc[0], pc[1], pc[2]};
            int lin = resolve<N, D>(a, ec, el, r);
            int pp = (r[2] * G::ecnt(ec, 1) + r[1]) * G::ecnt(ec, 0) + r[0];
            int qq = lin - e0;
            if (qq >= 0 && qq < ne) return sE[(qq * T + t) * G::EP + G::eoff(ec) + pp];
            return Ein[eslab<N, D, T>(ec, a.nel) + ((size_t)t * a.nel + lin) * G::epts(ec) + pp];
        };
        sS[(q * T + t) * G::EP + p] = solved<N, D, R>(a, e, sEl[q], pt, raw);
    }
    __syncthreads();
    // B. dB/dt of every slice at the thread's face-B points
    Acc dBdt[NB][T];
#pragma unroll
    for (int j = 0; j < NB; j++) {
        int idx = threadIdx.x + j * BS;
#pragma unroll
        for (int t = 0; t < T; t++) dBdt[j][t] = 0;
        if (idx < ne * G::BP) {
            // point-major over the tile's elements: a wave's lanes share the
            // component
            int p = idx / ne, q = idx - p * ne;
            int d = p < G::boff(1) ? 0 : (p < G::boff(2) ? 1 : 2);
            int pt[3];
            split_pt(p - G::boff(d), G::bcnt(d, 0), G::bcnt(d, 1), pt);
            const int a1 = d == 2 ? 0 : d + 1, a2 = a1 == 2 ? 0 : a1 + 1;
#pragma unroll
            for (int t = 0; t < T; t++) {
                // E1 = E_{d+1} differentiated along d+2, E2 = E_{d+2} along d+1
                const R* E1 = sS + (q * T + t) * G::EP + G::eoff(a1);
                const R* E2 = sS + (q * T + t) * G::EP + G::eoff(a2);
                Acc dB = 0, E;
                for (int l = 0; l < M; l++) {
                    if (a2 < D) {
                        int r[3] = {pt[0], pt[1], pt[2]};
                        r[a2] = l;
                        E = (Acc)E1[(r[2] * G::ecnt(a1, 1) + r[1]) * G::ecnt(a1, 0) + r[0]];
                        if (PLAIN) dB += E * sDm[pt[a2] * M + l] / a.h[a2];
                        else dB += E * sDmH[PLAIN ? 0 : a2][pt[a2] * M + l];
                    }
                    if (a1 < D) {
                        int r[3] = {pt[0], pt[1], pt[2]};
                        r[a1] = l;
                        E = (Acc)E2[(r[2] * G::ecnt(a2, 1) + r[1]) * G::ecnt(a2, 0) + r[0]];
                        if (PLAIN) dB -= E * sDm[pt[a1] * M + l] / a.h[a1];
                        else dB -= E * sDmH[PLAIN ? 0 : a1][pt[a1] * M + l];
                    }
                }
                dBdt[j][t] = dB;
            }
        }
    }

    // the thread's face-B points: old values, then the corrector or the
    // predictor slices
    double B_old[NB];
    double* Bp[NB];
#pragma unroll
    for (int j = 0; j < NB; j++) {
        int idx = threadIdx.x + j * BS;
        Bp[j] = nullptr;
        B_old[j] = 0;
        if (idx < ne * G::BP) {
            int p = idx / ne, q = idx - p * ne;
            int d = p < G::boff(1) ? 0 : (p < G::boff(2) ? 1 : 2);
            size_t gb = ((size_t)(sEl[q][2] + a.g[2]) * a.G[1] + (sEl[q][1] + a.g[1])) * a.G[0] + (sEl[q][0] + a.g[0]);
            Bp[j] = a.B[d] + gb * G::bpts(d) + (p - G::boff(d));
            B_old[j] = *Bp[j];
        }
    }
    if (CORR) {
#pragma unroll
        for (int j = 0; j < NB; j++) {
            if (!Bp[j]) continue;
            double dB = 0;
#pragma unroll
            for (int t = 0; t < T; t++) dB += (double)dBdt[j][t] * sW[t] * a.dt;
            *Bp[j] = B_old[j] - dB;
        }
        return;
    }
    R* Eout = (R*)a.Eout;
    const R vel[3] = {(R)a.vel[0], (R)a.vel[1], (R)a.vel[2]};
    for (int t = 0; t < T; t++) {
        __syncthreads();  // the raw buffer is free (face solves done / previous slice's E written)
        // C1. predictor slice t
#pragma unroll
        for (int j = 0; j < NB; j++) {
            int idx = threadIdx.x + j * BS;
            if (idx < ne * G::BP) {
                Acc dB = 0;
                if (PLAIN) {
#pragma unroll
                    for (int l = 0; l < T; l++) dB += dBdt[j][l] * sInv[t * T + l] * sW[l] * a.dt;
                } else {
#pragma unroll
                    for (int l = 0; l < T; l++) dB += dBdt[j][l] * (Acc)sP[t * T + l];
                }
                sB[(idx - (idx / ne) * ne) * G::BP + idx / ne] = (R)B_old[j] - (R)dB;
            }
        }
        __syncthreads();
        // C2. its raw E
#pragma unroll
        for (int e = 0; e < 3; e++) {
            R* dst = Eout + eslab<N, D, T>(e, a.nel) + ((size_t)t * a.nel + e0) * G::epts(e);
            for (int i = threadIdx.x; i < ne * G::epts(e); i += BS) {
                int q = i / G::epts(e), pe = i % G::epts(e);
                int pt[3];
                split_pt(pe, G::ecnt(e, 0), G::ecnt(e, 1), pt);
                const R* Bs = sB + q * G::BP;
                auto bget = [&](int d, const int r3[3]) -> R {
                    return Bs[G::boff(d) + (r3[2] * G::bcnt(d, 1) + r3[1]) * G::bcnt(d, 0) + r3[0]];
                };
                dst[i] = edge_e<N, D, R>(e, pt, sS2F, vel, bget);
            }
        }
    }
}

template <int N, int D, int T, class R>
size_t ws_bytes(int nel) {
    return 2 * (size_t)T * nel * IG<N, D>::EP * sizeof(R);
}

template <int N, int D, int T, class R>
int run_step(spdk_ctx* c, IArgs a, void* ws) {
    using TL = Tile<N, D, T, R>;
    static_assert(perm_is_permutation<N, D>(), "edge-point order must be a permutation");
    const size_t half = (size_t)T * a.nel * IG<N, D>::EP;
    R* buf[2] = {(R*)ws, (R*)ws + half};
    unsigned nb = (unsigned)((a.nel + TL::EPB - 1) / TL::EPB);
    if (SPDK_ISWEEP_XCD) nb = (nb + 7) / 8 * 8;
    const unsigned nstart = (unsigned)std::min<size_t>(((size_t)a.nel * IG<N, D>::EP + BS - 1) / BS, 65536);
    a.Eout = buf[0];
    a.Ein = nullptr;
    hipLaunchKernelGGL((k_istart<N, D, T, R>), dim3(nstart), dim3(BS), 0, c->stream, a);
    SPDK_LAUNCH_CHECK("induction istart");
    for (int s = 0; s < T; s++) {
        a.Ein = buf[s & 1];
        a.Eout = buf[(s + 1) & 1];
        bool rec = c->timing && c->ev_used < spdk_ctx::kEvPool;
        if (rec) SPDK_CHECK(hipEventRecord(c->evs[c->ev_used][0], c->stream));
        if (s < T - 1) hipLaunchKernelGGL((k_isweep<N, D, T, false, R>), dim3(nb), dim3(BS), 0, c->stream, a);
        else hipLaunchKernelGGL((k_isweep<N, D, T, true, R>), dim3(nb), dim3(BS), 0, c->stream, a);
        SPDK_LAUNCH_CHECK("induction isweep");
        if (rec) SPDK_CHECK(hipEventRecord(c->evs[c->ev_used++][1], c->stream));
    }
    if (c->timing && c->ev_used > 0) {  // harvest the sweep timings (syncs the step)
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

typedef int (*istep_fn)(spdk_ctx*, IArgs, void*);
typedef size_t (*iws_fn)(int);
struct IVariant {
    int n, dim, T;
    istep_fn step[2];  // FP64, mixed
    iws_fn ws[2];
};
#define SPDK_IVARIANT(Nn, D)                                                                  \
    {                                                                                         \
        Nn, D, Nn, {run_step<Nn, D, Nn, double>, run_step<Nn, D, Nn, float>},                 \
            {ws_bytes<Nn, D, Nn, double>, ws_bytes<Nn, D, Nn, float>}                        \
    }
const IVariant kIVariants[] = {SPDK_IVARIANT(2, 2), SPDK_IVARIANT(3, 2), SPDK_IVARIANT(4, 2), SPDK_IVARIANT(5, 2),
                               SPDK_IVARIANT(2, 3), SPDK_IVARIANT(3, 3), SPDK_IVARIANT(4, 3)};

const IVariant* find_ivariant(spdk_ctx* c) {
    if (!c->prm.active[0] || !c->prm.active[1]) return nullptr;
    int dim = c->prm.active[2] ? 3 : 2;
    if (c->prm.rk_order != 0) return nullptr;  // ADER only
    for (const IVariant& v : kIVariants)
        if (v.n == c->n && v.dim == dim && v.T == c->T) return &v;
    return nullptr;
}

bool bc_supported(spdk_ctx* c) {
    for (int d = 0; d < 3; d++)
        if (c->prm.active[d] && c->prm.bc[d] != SPDK_BC_PERIODIC && c->prm.bc[d] != SPDK_BC_GRADFREE) return false;
    return true;
}

size_t workspace_bytes(spdk_ctx* c, const int N[3], int mixed) {
    const IVariant* v = c ? find_ivariant(c) : nullptr;
    if (!v || !bc_supported(c)) {
        set_error("fused induction path: ADER, x-y or x-y-z active, p = 1..4 (1..3 in 3-D), periodic or gradfree "
                  "boundaries");
        return 0;
    }
    size_t nel = 1;
    for (int d = 0; d < 3; d++) nel *= (size_t)(c->prm.active[d] ? N[d] - 2 * c->ngh[d] : 1);
    if (nel == 0 || nel >= (1ull << 31)) { set_error("fused induction: element count"); return 0; }
    return v->ws[mixed]((int)nel);
}

int step(spdk_ctx* c, double* Bx, spdk_sd_shape sx, double* By, spdk_sd_shape sy, double* Bz, spdk_sd_shape sz, void* ws,
         double dt, int mixed) {
    if (!c || !Bx || !By || !Bz || !ws) { set_error("fused induction: null argument"); return SPDK_E_ARG; }
    const IVariant* v = find_ivariant(c);
    if (!v || !bc_supported(c)) { set_error("fused induction: unsupported configuration"); return SPDK_E_CONFIG; }
    const spdk_sd_shape* s[3] = {&sx, &sy, &sz};
    const int n = c->n, m = c->m;
    for (int d = 0; d < 3; d++) {
        const spdk_sd_shape& q = *s[d];
        int want[3];
        for (int a = 0; a < 3; a++) want[a] = c->prm.active[a] ? (a == d ? m : n) : 1;
        if (q.nt != 1 || q.nv != 1 || q.nx != want[0] || q.ny != want[1] || q.nz != want[2] || q.Nx != sx.Nx ||
            q.Ny != sx.Ny || q.Nz != sx.Nz) {
            set_error("fused induction: face-B shape");
            return SPDK_E_SHAPE;
        }
    }
    IArgs a;
    std::memset(&a, 0, sizeof(a));
    a.B[0] = Bx;
    a.B[1] = By;
    a.B[2] = Bz;
    a.S2F = c->mat[SPDK_M_SP_TO_FP];
    a.Dm = c->mat[SPDK_M_DFP_TO_SP];
    a.inv = c->mat[SPDK_M_INVADER];
    a.w = c->wt;
    a.dt = dt;
    const int Gd[3] = {sx.Nx, sx.Ny, sx.Nz};
    size_t nel = 1;
    for (int d = 0; d < 3; d++) {
        a.vel[d] = c->prm.vel[d];
        a.h[d] = c->prm.h[d];
        a.G[d] = Gd[d];
        a.g[d] = c->ngh[d];
        a.n[d] = Gd[d] - 2 * c->ngh[d];
        a.periodic[d] = c->prm.bc[d] == SPDK_BC_PERIODIC;
        if (a.n[d] < 1) { set_error("fused induction: axis needs ghost elements"); return SPDK_E_SHAPE; }
        nel *= (size_t)a.n[d];
    }
    if (nel >= (1ull << 31)) { set_error("fused induction: too many elements"); return SPDK_E_SHAPE; }
    a.nel = (int)nel;
    std::strncpy(c->kernel_name, "k_isweep", sizeof(c->kernel_name) - 1);
    return v->step[mixed](c, a, ws);
}

}  // namespace

extern "C" {

size_t spdk_induction_fused_workspace_bytes(spdk_ctx* c, const int N[3]) { return workspace_bytes(c, N, 0); }

int spdk_induction_step_fused(spdk_ctx* c, double* Bx, spdk_sd_shape sx, double* By, spdk_sd_shape sy, double* Bz,
                              spdk_sd_shape sz, void* ws, double dt) {
    return step(c, Bx, sx, By, sy, Bz, sz, ws, dt, 0);
}

size_t spdk_induction_mixed_workspace_bytes(spdk_ctx* c, const int N[3]) { return workspace_bytes(c, N, 1); }

int spdk_induction_step_mixed(spdk_ctx* c, double* Bx, spdk_sd_shape sx, double* By, spdk_sd_shape sy, double* Bz,
                              spdk_sd_shape sz, void* ws, double dt) {
    return step(c, Bx, sx, By, sy, Bz, sz, ws, dt, 1);
}

}  // extern "C"
