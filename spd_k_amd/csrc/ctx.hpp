// Library-private context and launch helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>
#include "../../include/spdk.h"

struct spdk_ctx {
    spdk_params prm;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int p = 0, n = 1, m = 2, T = 1;
    int ngh[3] = {0, 0, 0};  // ghost elements per axis
    int fgh[3] = {0, 0, 0};  // ghost FV cells per axis
    double* mat[SPDK_M_COUNT] = {};
    std::vector<double> mat_host[SPDK_M_COUNT];  // host copies, row-major
    int mrows[SPDK_M_COUNT] = {}, mcols[SPDK_M_COUNT] = {};
    double* wt = nullptr;
    double wt_host[10] = {};  // host copy of the time weights
    double* fvf[3] = {};
    double* fvc[3] = {};
    int fvnf[3] = {}, fvn[3] = {};
    double* red = nullptr;       // reduction partials (device)
    double* red_host = nullptr;  // pinned host landing slot
    int red_blocks = 2048;
    // optional timing of the dominant kernel
    bool timing = false;
    double kernel_ms = 0;
    int kernel_launches = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    char kernel_name[64] = "";
    static constexpr int kEvPool = 64;
    hipEvent_t evs[kEvPool][2] = {};
    int ev_used = 0;  // event pairs recorded since the last harvest
    bool fb_tile = true;  // MUSCL-HLLC fallback on cell tiles (else one thread per cell)
    // device buffers created on first use by a path and freed with the context
    float* mixed_mats = nullptr;  // FP32 sp_to_fp and dfp_to_sp / h_d
    spdk_mlp* guess_net = nullptr;  // learned Picard guess (not owned)
    float *guess_x = nullptr, *guess_y = nullptr;  // feature / output chunks
    int guess_cap = 0;
    int* guess_list = nullptr;  // elements the network guesses for (list, then the counter)
    size_t guess_list_cap = 0;
    std::vector<void*> owned;
};

namespace spdk {

void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define SPDK_CHECK(call)                                   \
    do {                                                   \
        hipError_t e_ = (call);                            \
        if (e_ != hipSuccess) return spdk::hip_fail(e_, #call); \
    } while (0)

#define SPDK_LAUNCH_CHECK(name)                                        \
    do {                                                               \
        hipError_t e_ = hipGetLastError();                             \
        if (e_ != hipSuccess) return spdk::hip_fail(e_, "launch " name); \
    } while (0)

// Strides of an SD array (t, v, k, j, i, kk, jj, ii), 64-bit offsets.
struct SD {
    int nt, nv, Nz, Ny, Nx, nz, ny, nx;
    __host__ __device__ SD() {}
    __host__ SD(const spdk_sd_shape& s) : nt(s.nt), nv(s.nv), Nz(s.Nz), Ny(s.Ny), Nx(s.Nx), nz(s.nz), ny(s.ny), nx(s.nx) {}
    __host__ __device__ __forceinline__ size_t pts() const { return (size_t)nz * ny * nx; }
    __host__ __device__ __forceinline__ size_t slab() const { return (size_t)Nz * Ny * Nx * pts(); }  // one (t,v)
    __host__ __device__ __forceinline__ size_t off(int t, int v, int k, int j, int i, int kk, int jj, int ii) const {
        return ((size_t)t * nv + v) * slab() + ((((size_t)k * Ny + j) * Nx + i) * pts()) + ((size_t)kk * ny + jj) * nx + ii;
    }
    __host__ __device__ __forceinline__ size_t total_cells() const { return (size_t)Nz * Ny * Nx * pts(); }
};
struct FV {
    int nv, Nz, Ny, Nx;
    __host__ __device__ FV() {}
    __host__ FV(const spdk_fv_shape& s) : nv(s.nv), Nz(s.Nz), Ny(s.Ny), Nx(s.Nx) {}
    __host__ __device__ __forceinline__ size_t plane() const { return (size_t)Nz * Ny * Nx; }
    __host__ __device__ __forceinline__ size_t off(int v, int k, int j, int i) const {
        return (size_t)v * plane() + ((size_t)k * Ny + j) * Nx + i;
    }
};

// Flat (element, point) decomposition, point index fastest (src/define.hpp:143-162).
__device__ __forceinline__ void split6(uint32_t idx, int Mz, int My, int Mx, int nz, int ny, int nx, int& k, int& j,
                                       int& i, int& kk, int& jj, int& ii) {
    ii = idx % nx; idx /= nx;
    jj = idx % ny; idx /= ny;
    kk = idx % nz; idx /= nz;
    i = idx % Mx; idx /= Mx;
    j = idx % My;
    k = idx / My;
    (void)Mz;
}
__device__ __forceinline__ void split3(uint32_t idx, int My, int Mx, int& k, int& j, int& i) {
    i = idx % Mx; idx /= Mx;
    j = idx % My;
    k = idx / My;
}

// Reductions that keep a NaN: once either operand is NaN the result is NaN,
// so a diverged state reaches the host (dt, residual, diagnostics) instead of
// being folded away. For ordered operands they equal min / max.
__host__ __device__ __forceinline__ double nan_min(double a, double b) { return (b < a || b != b) ? b : a; }
__host__ __device__ __forceinline__ double nan_max(double a, double b) { return (b > a || b != b) ? b : a; }
__host__ __device__ __forceinline__ float nan_maxf(float a, float b) { return (b > a || b != b) ? b : a; }

inline unsigned nblocks(size_t total, int bs) { return (unsigned)((total + bs - 1) / bs); }

}  // namespace spdk
