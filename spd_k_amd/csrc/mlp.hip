// Per-element MLP for the Picard initial guess: BF16 MFMA forward and
// backward kernels with FP32 accumulation, and Adam with FP32 master weights.
//
// Network: x (din) -> GELU(W1 x + b1) -> GELU(W2 h1 + b2) -> W3 h2 + b3 (dout),
// GELU(z) = z/2 (1 + erf(z / sqrt 2)). Weights are stored "out x in"
// (row n holds the weights of output n); every feature dimension is padded
// to a multiple of 64 with zero weights.
//
// One GEMM kernel serves every product: C(m, n) = sum_k A(m, k) B(k, n) with
// A and B addressed through element strides (so transposed operands need no
// copies), 64 x 64 output tiles, K in steps of 32 staged through LDS, and
// four waves each computing a 16 x 64 strip with v_mfma_f32_16x16x32_bf16.
// Lane layout of that instruction (checked against a CPU product): lane l
// holds A[l%16][8(l/16)+j], B[8(l/16)+j][l%16] (j < 8) and C[4(l/16)+i][l%16]
// (i < 4).
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <vector>
#include "ctx.hpp"

using namespace spdk;

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

struct spdk_mlp {
    int dims[4] = {};  // padded: in, h1, h2, out
    int din = 0, dout = 0;
    int cap = 0;       // batch capacity of the activation buffers
    hipStream_t stream = nullptr;
    float *W[3] = {}, *b[3] = {};             // FP32 master parameters
    bf16 *Wb[3] = {}, *WbT[3] = {};           // BF16 copies, out x in and in x out
    float *mW[3] = {}, *vW[3] = {}, *mb[3] = {}, *vb[3] = {};  // Adam moments
    float *gW[3] = {}, *gb[3] = {};           // gradients
    bf16* X = nullptr;                        // (cap, d0)
    float* Z[3] = {};                         // pre-activations (cap, d_{l+1})
    bf16* H[2] = {};                          // activations (cap, d1), (cap, d2)
    float* dZ = nullptr;                      // (cap, max d) gradient scratch
    bf16 *dZb[2] = {};                        // BF16 copies of layer gradients (ping-pong)
    float* loss = nullptr;                    // device scalar
    long long step = 0;
    std::vector<void*> owned;
};

namespace {

enum { EPI_STORE = 0, EPI_BIAS_GELU = 1, EPI_BIAS = 2, EPI_GELU_BWD = 3, EPI_ATOMIC = 4 };

struct GemmP {
    const bf16* A;
    long sam, sak;  // A(m, k) = A[m * sam + k * sak]
    const bf16* B;
    long sbk, sbn;  // B(k, n) = B[k * sbk + n * sbn]
    int M, N, K;
    int kchunk;     // split K: blockIdx.z covers [z * kchunk, (z + 1) * kchunk)
    int epi;
    const float* bias;
    float* C;
    long ldc;
    bf16* Cb;
    long ldcb;
    const float* Zp;  // pre-activation for the GELU derivative
    long ldz;
};

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

constexpr int TM = 64, TN = 64, TK = 32, LDK = TK + 8;  // LDS rows padded to 80 bytes

__global__ void __launch_bounds__(256) k_gemm(GemmP p) {
    __shared__ __align__(16) bf16 As[TM * LDK];
    __shared__ __align__(16) bf16 Bs[TN * LDK];  // B staged transposed: Bs[n][k]
    const int tid = threadIdx.x, w = tid / 64, l = tid % 64;
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const int k_begin = blockIdx.z * p.kchunk;
    const int k_end = min(p.K, k_begin + p.kchunk);
    floatx4 acc[4];
#pragma unroll
    for (int c = 0; c < 4; c++) acc[c] = floatx4{0.f, 0.f, 0.f, 0.f};
    const int lr = tid / 4, lk = (tid % 4) * 8;  // staging: row lr, k offset lk (8 values)
    for (int k0 = k_begin; k0 < k_end; k0 += TK) {
        {   // A tile
            const int m = m0 + lr;
            bf16x8 v;
            if (m < p.M && k0 + lk + 7 < k_end && p.sak == 1) {
                v = *(const bf16x8*)(p.A + (long)m * p.sam + (k0 + lk));
            } else {
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const int k = k0 + lk + j;
                    v[j] = (m < p.M && k < k_end) ? p.A[(long)m * p.sam + (long)k * p.sak] : (bf16)0.0f;
                }
            }
            *(bf16x8*)(As + lr * LDK + lk) = v;
        }
        {   // B tile, transposed into Bs[n][k]
            const int n = n0 + lr;
            bf16x8 v;
            if (n < p.N && k0 + lk + 7 < k_end && p.sbk == 1) {
                v = *(const bf16x8*)(p.B + (long)n * p.sbn + (k0 + lk));
            } else {
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const int k = k0 + lk + j;
                    v[j] = (n < p.N && k < k_end) ? p.B[(long)k * p.sbk + (long)n * p.sbn] : (bf16)0.0f;
                }
            }
            *(bf16x8*)(Bs + lr * LDK + lk) = v;
        }
        __syncthreads();
        const bf16x8 a = *(const bf16x8*)(As + (16 * w + l % 16) * LDK + 8 * (l / 16));
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const bf16x8 b = *(const bf16x8*)(Bs + (16 * c + l % 16) * LDK + 8 * (l / 16));
            acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[c], 0, 0, 0);
        }
        __syncthreads();
    }
    // epilogue: lane l holds C[16w + 4(l/16) + i][16c + l%16]
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const int n = n0 + 16 * c + l % 16;
        if (n >= p.N) continue;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int m = m0 + 16 * w + 4 * (l / 16) + i;
            if (m >= p.M) continue;
            float v = acc[c][i];
            if (p.epi == EPI_STORE) {
                p.C[(long)m * p.ldc + n] = v;
            } else if (p.epi == EPI_ATOMIC) {
                atomicAdd(&p.C[(long)m * p.ldc + n], v);
            } else if (p.epi == EPI_BIAS_GELU) {
                v += p.bias[n];
                p.C[(long)m * p.ldc + n] = v;
                p.Cb[(long)m * p.ldcb + n] = (bf16)gelu(v);
            } else if (p.epi == EPI_BIAS) {
                v += p.bias[n];
                p.C[(long)m * p.ldc + n] = v;
            } else {  // EPI_GELU_BWD
                v *= gelu_d(p.Zp[(long)m * p.ldz + n]);
                p.C[(long)m * p.ldc + n] = v;
                p.Cb[(long)m * p.ldcb + n] = (bf16)v;
            }
        }
    }
}

// loss = mean (Y - T)^2 over the first dout columns; dY = 2 (Y - T) / count
// (FP32 and BF16). Padded columns get zero gradient.
__global__ void k_mse(const float* Y, long ldy, const float* T, long ldt, int M, int dout, int dpad, float inv_count,
                      float* dY, bf16* dYb, float* loss) {
    __shared__ float red[256];
    float acc = 0.f;
    const long total = (long)M * dpad;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const int m = (int)(i / dpad), n = (int)(i % dpad);
        float g = 0.f;
        if (n < dout) {
            const float d = Y[(long)m * ldy + n] - T[(long)m * ldt + n];
            acc += d * d;
            g = 2.f * d * inv_count;
        }
        dY[(long)m * dpad + n] = g;
        dYb[(long)m * dpad + n] = (bf16)g;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0] * inv_count);
}

// bias gradient: gb[n] = sum_m dZ[m][n]
__global__ void k_colsum(const float* dZ, int M, int N, float* gb) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.f;
    for (int m = 0; m < M; m++) s += dZ[(long)m * N + n];
    gb[n] = s;
}

__global__ void k_fill(float* p, long n, float v) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}

// Adam (bias-corrected) on FP32 masters; refreshes the BF16 copies of a
// weight matrix in both layouts (rows x cols = out x in).
__global__ void k_adam(float* w, float* m, float* v, const float* g, long n, int cols, float lr, float b1, float b2,
                       float eps, float c1, float c2, bf16* wb, bf16* wbT, int rows) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
        w[i] = wi;
        if (wb) {
            wb[i] = (bf16)wi;
            const long r = i / cols, cc = i % cols;
            wbT[cc * rows + r] = (bf16)wi;
        }
    }
}

__global__ void k_to_bf16(const float* src, bf16* dst, bf16* dstT, long n, int cols, int rows) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        dst[i] = (bf16)src[i];
        if (dstT) dstT[(i % cols) * rows + i / cols] = (bf16)src[i];
    }
}

// FP32 rows (M x din, leading dimension ld) into the padded BF16 input.
__global__ void k_pack_input(const float* Xf, long ld, int M, int din, int dpad, bf16* X) {
    const long total = (long)M * dpad;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const int m = (int)(i / dpad), n = (int)(i % dpad);
        X[i] = (bf16)(n < din ? Xf[(long)m * ld + n] : 0.0f);
    }
}

// dst[r][:] = src[idx[r]][:] for r < n (cols values per row)
__global__ void k_gather(const float* src, long ld, const int* idx, int n, int cols, float* dst) {
    const long total = (long)n * cols;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const int r = (int)(i / cols), c = (int)(i % cols);
        dst[i] = src[(long)idx[r] * ld + c];
    }
}

// flags[r] = 1 if every value of row r is finite with magnitude < maxabs
__global__ void k_rows_ok(const float* Y, long ld, int rows, int cols, float maxabs, unsigned char* flags) {
    const int r = blockIdx.x;
    if (r >= rows) return;
    __shared__ int bad;
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    for (int c = threadIdx.x; c < cols; c += blockDim.x) {
        const float v = Y[(long)r * ld + c];
        if (!(fabsf(v) < maxabs)) bad = 1;
    }
    __syncthreads();
    if (threadIdx.x == 0) flags[r] = bad ? 0 : 1;
}

inline unsigned nb(long n, int bs) { long b = (n + bs - 1) / bs; return (unsigned)(b > 65535 * 4 ? 65535 * 4 : b); }

int gemm(hipStream_t st, const GemmP& p, int ksplit = 1) {
    GemmP q = p;
    q.kchunk = ((p.K + ksplit - 1) / ksplit + TK - 1) / TK * TK;
    const int nz = (p.K + q.kchunk - 1) / q.kchunk;
    dim3 grid((p.N + TN - 1) / TN, (p.M + TM - 1) / TM, nz);
    if (grid.y > 65535 * 16) { set_error("mlp: batch too large for one launch"); return SPDK_E_ARG; }
    hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, st, q);
    SPDK_LAUNCH_CHECK("k_gemm");
    return SPDK_OK;
}

int alloc(spdk_mlp* net, void** p, size_t bytes) {
    SPDK_CHECK(hipMalloc(p, bytes));
    net->owned.push_back(*p);
    SPDK_CHECK(hipMemsetAsync(*p, 0, bytes, net->stream));
    return SPDK_OK;
}

// forward over B rows already packed in net->X; output pre-activation Z[2] (B x d3)
int forward(spdk_mlp* net, int B) {
    const int* d = net->dims;
    const bf16* in = net->X;
    for (int l = 0; l < 3; l++) {
        GemmP p{};
        p.A = in;
        p.sam = d[l];
        p.sak = 1;
        p.B = net->Wb[l];  // B(k, n) = W[n][k]
        p.sbk = 1;
        p.sbn = d[l];
        p.M = B;
        p.N = d[l + 1];
        p.K = d[l];
        p.epi = l < 2 ? EPI_BIAS_GELU : EPI_BIAS;
        p.bias = net->b[l];
        p.C = net->Z[l];
        p.ldc = d[l + 1];
        p.Cb = l < 2 ? net->H[l] : nullptr;
        p.ldcb = d[l + 1];
        int rc = gemm(net->stream, p);
        if (rc) return rc;
        if (l < 2) in = net->H[l];
    }
    return SPDK_OK;
}

// gradients of the MSE loss for B rows (targets T: B x dout, leading dim ldt)
int backward(spdk_mlp* net, int B, const float* T, long ldt) {
    const int* d = net->dims;
    hipStream_t st = net->stream;
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(1), 0, st, net->loss, 1L, 0.f);
    const float inv = 1.0f / ((float)B * (float)net->dout);
    hipLaunchKernelGGL(k_mse, dim3(nb((long)B * d[3], 256)), dim3(256), 0, st, net->Z[2], (long)d[3], T, ldt, B,
                       net->dout, d[3], inv, net->dZ, net->dZb[0], net->loss);
    SPDK_LAUNCH_CHECK("k_mse");
    int cur = 0;
    for (int l = 2; l >= 0; l--) {
        const bf16* act = l > 0 ? net->H[l - 1] : net->X;  // layer input (B x d[l])
        // gW[l][n][k] = sum_m dZ[m][n] act[m][k]: M' = d[l+1], N' = d[l], K' = B
        long nw = (long)d[l + 1] * d[l];
        hipLaunchKernelGGL(k_fill, dim3(nb(nw, 256)), dim3(256), 0, st, net->gW[l], nw, 0.f);
        GemmP p{};
        p.A = net->dZb[cur];
        p.sam = 1;
        p.sak = d[l + 1];
        p.B = act;
        p.sbk = d[l];
        p.sbn = 1;
        p.M = d[l + 1];
        p.N = d[l];
        p.K = B;
        p.epi = EPI_ATOMIC;
        p.C = net->gW[l];
        p.ldc = d[l];
        int rc = gemm(st, p, (B + 511) / 512);
        if (rc) return rc;
        hipLaunchKernelGGL(k_colsum, dim3((d[l + 1] + 255) / 256), dim3(256), 0, st, net->dZ, B, d[l + 1], net->gb[l]);
        SPDK_LAUNCH_CHECK("k_colsum");
        if (l == 0) break;
        // dH = dZ W (B x d[l]), then dZ_{l-1} = dH * gelu'(Z_{l-1})
        GemmP q{};
        q.A = net->dZb[cur];
        q.sam = d[l + 1];
        q.sak = 1;
        q.B = net->WbT[l];  // B(k = out, n = in) = W[out][in] = WbT[in][out]
        q.sbk = 1;
        q.sbn = d[l + 1];
        q.M = B;
        q.N = d[l];
        q.K = d[l + 1];
        q.epi = EPI_GELU_BWD;
        q.Zp = net->Z[l - 1];
        q.ldz = d[l];
        q.C = net->dZ;
        q.ldc = d[l];
        q.Cb = net->dZb[1 - cur];
        q.ldcb = d[l];
        if ((rc = gemm(st, q))) return rc;
        cur = 1 - cur;
    }
    return SPDK_OK;
}

int check_net(spdk_mlp* net, int B) {
    if (!net) { set_error("mlp: null network"); return SPDK_E_ARG; }
    if (B < 1 || B > net->cap) { set_error("mlp: batch outside 1..capacity"); return SPDK_E_ARG; }
    return SPDK_OK;
}

}  // namespace

extern "C" {

int spdk_mlp_create(const int dims[4], int batch_cap, void* stream, spdk_mlp** out) {
    if (!dims || !out || batch_cap < 1) { set_error("mlp: bad argument"); return SPDK_E_ARG; }
    spdk_mlp* net = new spdk_mlp();
    net->din = dims[0];
    net->dout = dims[3];
    for (int i = 0; i < 4; i++) {
        if (dims[i] < 1 || dims[i] > 65536) { delete net; set_error("mlp: bad dimension"); return SPDK_E_ARG; }
        net->dims[i] = (dims[i] + 63) / 64 * 64;
    }
    net->cap = batch_cap;
    net->stream = (hipStream_t)stream;
    const int* d = net->dims;
    int rc = SPDK_OK;
    auto A = [&](void** p, size_t b) { if (!rc) rc = alloc(net, p, b); };
    int dmax = 0;
    for (int l = 0; l < 3; l++) {
        size_t nw = (size_t)d[l + 1] * d[l];
        A((void**)&net->W[l], nw * 4);
        A((void**)&net->b[l], d[l + 1] * 4);
        A((void**)&net->Wb[l], nw * 2);
        A((void**)&net->WbT[l], nw * 2);
        A((void**)&net->mW[l], nw * 4);
        A((void**)&net->vW[l], nw * 4);
        A((void**)&net->mb[l], d[l + 1] * 4);
        A((void**)&net->vb[l], d[l + 1] * 4);
        A((void**)&net->gW[l], nw * 4);
        A((void**)&net->gb[l], d[l + 1] * 4);
        A((void**)&net->Z[l], (size_t)batch_cap * d[l + 1] * 4);
        dmax = dmax > d[l + 1] ? dmax : d[l + 1];
    }
    dmax = dmax > d[0] ? dmax : d[0];
    A((void**)&net->X, (size_t)batch_cap * d[0] * 2);
    A((void**)&net->H[0], (size_t)batch_cap * d[1] * 2);
    A((void**)&net->H[1], (size_t)batch_cap * d[2] * 2);
    A((void**)&net->dZ, (size_t)batch_cap * dmax * 4);
    A((void**)&net->dZb[0], (size_t)batch_cap * dmax * 2);
    A((void**)&net->dZb[1], (size_t)batch_cap * dmax * 2);
    A((void**)&net->loss, 16);
    if (rc) { for (void* p : net->owned) (void)hipFree(p); delete net; return rc; }
    *out = net;
    return SPDK_OK;
}

int spdk_mlp_destroy(spdk_mlp* net) {
    if (!net) return SPDK_OK;
    (void)hipStreamSynchronize(net->stream);
    for (void* p : net->owned) (void)hipFree(p);
    delete net;
    return SPDK_OK;
}

int spdk_gather_rows(const float* src, long ld, const int* idx, int n, int cols, float* dst, void* stream) {
    if (!src || !idx || !dst || n < 0 || cols < 0) { set_error("gather: bad argument"); return SPDK_E_ARG; }
    if (n == 0 || cols == 0) return SPDK_OK;
    hipLaunchKernelGGL(k_gather, dim3(nb((long)n * cols, 256)), dim3(256), 0, (hipStream_t)stream, src, ld, idx, n, cols,
                       dst);
    SPDK_LAUNCH_CHECK("k_gather");
    return SPDK_OK;
}

int spdk_rows_finite(const float* Y, long ld, int rows, int cols, float maxabs, unsigned char* flags, void* stream) {
    if (!Y || !flags || rows < 0 || cols < 0) { set_error("rows_finite: bad argument"); return SPDK_E_ARG; }
    if (rows == 0) return SPDK_OK;
    hipLaunchKernelGGL(k_rows_ok, dim3(rows), dim3(256), 0, (hipStream_t)stream, Y, ld, rows, cols, maxabs, flags);
    SPDK_LAUNCH_CHECK("k_rows_ok");
    return SPDK_OK;
}

int spdk_mlp_set_stream(spdk_mlp* net, void* stream) {
    if (!net) return SPDK_E_ARG;
    SPDK_CHECK(hipStreamSynchronize(net->stream));
    net->stream = (hipStream_t)stream;
    return SPDK_OK;
}

int spdk_mlp_dims(spdk_mlp* net, int padded[4]) {
    if (!net || !padded) return SPDK_E_ARG;
    for (int i = 0; i < 4; i++) padded[i] = net->dims[i];
    return SPDK_OK;
}

// Parameters as one FP32 host array, unpadded: W1 (h1 x in), b1, W2, b2, W3, b3.
size_t spdk_mlp_param_count(spdk_mlp* net) {
    if (!net) return 0;
    size_t n = 0;
    const int u[4] = {net->din, net->dims[1], net->dims[2], net->dout};
    for (int l = 0; l < 3; l++) n += (size_t)u[l + 1] * u[l] + u[l + 1];
    return n;
}

static int xfer_params(spdk_mlp* net, float* host, bool to_dev) {
    const int u[4] = {net->din, net->dims[1], net->dims[2], net->dout};
    const int* d = net->dims;
    size_t o = 0;
    for (int l = 0; l < 3; l++) {
        std::vector<float> W((size_t)d[l + 1] * d[l], 0.f), b(d[l + 1], 0.f);
        if (!to_dev) {
            SPDK_CHECK(hipMemcpyAsync(W.data(), net->W[l], W.size() * 4, hipMemcpyDeviceToHost, net->stream));
            SPDK_CHECK(hipMemcpyAsync(b.data(), net->b[l], b.size() * 4, hipMemcpyDeviceToHost, net->stream));
            SPDK_CHECK(hipStreamSynchronize(net->stream));
        }
        for (int r = 0; r < u[l + 1]; r++)
            for (int k = 0; k < u[l]; k++) {
                if (to_dev) W[(size_t)r * d[l] + k] = host[o];
                else host[o] = W[(size_t)r * d[l] + k];
                o++;
            }
        for (int r = 0; r < u[l + 1]; r++) {
            if (to_dev) b[r] = host[o];
            else host[o] = b[r];
            o++;
        }
        if (to_dev) {
            SPDK_CHECK(hipMemcpyAsync(net->W[l], W.data(), W.size() * 4, hipMemcpyHostToDevice, net->stream));
            SPDK_CHECK(hipMemcpyAsync(net->b[l], b.data(), b.size() * 4, hipMemcpyHostToDevice, net->stream));
            const long nw = (long)W.size();
            hipLaunchKernelGGL(k_to_bf16, dim3(nb(nw, 256)), dim3(256), 0, net->stream, (const float*)net->W[l],
                               net->Wb[l], net->WbT[l], nw, d[l], d[l + 1]);
            SPDK_LAUNCH_CHECK("k_to_bf16");
            SPDK_CHECK(hipStreamSynchronize(net->stream));
        }
    }
    return SPDK_OK;
}

int spdk_mlp_set_params(spdk_mlp* net, const float* host) {
    if (!net || !host) return SPDK_E_ARG;
    net->step = 0;
    for (int l = 0; l < 3; l++) {
        const int* d = net->dims;
        long nw = (long)d[l + 1] * d[l];
        hipLaunchKernelGGL(k_fill, dim3(nb(nw, 256)), dim3(256), 0, net->stream, net->mW[l], nw, 0.f);
        hipLaunchKernelGGL(k_fill, dim3(nb(nw, 256)), dim3(256), 0, net->stream, net->vW[l], nw, 0.f);
        hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, net->stream, net->mb[l], (long)d[l + 1], 0.f);
        hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, net->stream, net->vb[l], (long)d[l + 1], 0.f);
    }
    return xfer_params(net, const_cast<float*>(host), true);
}

int spdk_mlp_get_params(spdk_mlp* net, float* host) {
    if (!net || !host) return SPDK_E_ARG;
    return xfer_params(net, host, false);
}

// Forward pass on B device rows of FP32 features (leading dimension ldx);
// writes the B x dout outputs (leading dimension ldy) to Y (device).
int spdk_mlp_forward(spdk_mlp* net, const float* X, long ldx, int B, float* Y, long ldy) {
    int rc = check_net(net, B);
    if (rc) return rc;
    hipLaunchKernelGGL(k_pack_input, dim3(nb((long)B * net->dims[0], 256)), dim3(256), 0, net->stream, X, ldx, B,
                       net->din, net->dims[0], net->X);
    SPDK_LAUNCH_CHECK("k_pack_input");
    if ((rc = forward(net, B))) return rc;
    SPDK_CHECK(hipMemcpy2DAsync(Y, ldy * 4, net->Z[2], (size_t)net->dims[3] * 4, (size_t)net->dout * 4, B,
                                hipMemcpyDeviceToDevice, net->stream));
    return SPDK_OK;
}

// Gradients of the loss mean((f(X) - T)^2) for B device rows, copied to the
// host in the layout of spdk_mlp_set_params; *loss (host) gets the loss.
int spdk_mlp_gradients(spdk_mlp* net, const float* X, long ldx, const float* T, long ldt, int B, float* host_grad,
                       float* loss) {
    int rc = check_net(net, B);
    if (rc) return rc;
    hipLaunchKernelGGL(k_pack_input, dim3(nb((long)B * net->dims[0], 256)), dim3(256), 0, net->stream, X, ldx, B,
                       net->din, net->dims[0], net->X);
    SPDK_LAUNCH_CHECK("k_pack_input");
    if ((rc = forward(net, B))) return rc;
    if ((rc = backward(net, B, T, ldt))) return rc;
    if (loss) {
        SPDK_CHECK(hipMemcpyAsync(loss, net->loss, 4, hipMemcpyDeviceToHost, net->stream));
    }
    SPDK_CHECK(hipStreamSynchronize(net->stream));
    if (host_grad) {
        const int u[4] = {net->din, net->dims[1], net->dims[2], net->dout};
        const int* d = net->dims;
        size_t o = 0;
        for (int l = 0; l < 3; l++) {
            std::vector<float> g((size_t)d[l + 1] * d[l]), gb(d[l + 1]);
            SPDK_CHECK(hipMemcpy(g.data(), net->gW[l], g.size() * 4, hipMemcpyDeviceToHost));
            SPDK_CHECK(hipMemcpy(gb.data(), net->gb[l], gb.size() * 4, hipMemcpyDeviceToHost));
            for (int r = 0; r < u[l + 1]; r++)
                for (int k = 0; k < u[l]; k++) host_grad[o++] = g[(size_t)r * d[l] + k];
            for (int r = 0; r < u[l + 1]; r++) host_grad[o++] = gb[r];
        }
    }
    return SPDK_OK;
}

// One Adam step on B device rows; *loss (host, may be NULL) gets the loss
// before the update (reading it synchronises).
int spdk_mlp_train_step(spdk_mlp* net, const float* X, long ldx, const float* T, long ldt, int B, float lr,
                        float* loss) {
    int rc = check_net(net, B);
    if (rc) return rc;
    hipLaunchKernelGGL(k_pack_input, dim3(nb((long)B * net->dims[0], 256)), dim3(256), 0, net->stream, X, ldx, B,
                       net->din, net->dims[0], net->X);
    SPDK_LAUNCH_CHECK("k_pack_input");
    if ((rc = forward(net, B))) return rc;
    if ((rc = backward(net, B, T, ldt))) return rc;
    net->step++;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f / (1.0f - powf(b1, (float)net->step)), c2 = 1.0f / (1.0f - powf(b2, (float)net->step));
    const int* d = net->dims;
    for (int l = 0; l < 3; l++) {
        long nw = (long)d[l + 1] * d[l];
        hipLaunchKernelGGL(k_adam, dim3(nb(nw, 256)), dim3(256), 0, net->stream, net->W[l], net->mW[l], net->vW[l],
                           (const float*)net->gW[l], nw, d[l], lr, b1, b2, eps, c1, c2, net->Wb[l], net->WbT[l],
                           d[l + 1]);
        hipLaunchKernelGGL(k_adam, dim3(1), dim3(256), 0, net->stream, net->b[l], net->mb[l], net->vb[l],
                           (const float*)net->gb[l], (long)d[l + 1], d[l + 1], lr, b1, b2, eps, c1, c2,
                           (bf16*)nullptr, (bf16*)nullptr, 1);
        SPDK_LAUNCH_CHECK("k_adam");
    }
    if (loss) {
        SPDK_CHECK(hipMemcpyAsync(loss, net->loss, 4, hipMemcpyDeviceToHost, net->stream));
        SPDK_CHECK(hipStreamSynchronize(net->stream));
    }
    return SPDK_OK;
}

}  // extern "C"
