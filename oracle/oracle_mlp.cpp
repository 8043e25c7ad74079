// ORACLE — TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
// CPU restatement of the per-element MLP of the learned Picard guess, the
// checker of the GPU forward and backward kernels. No reference counterpart
// (the network is new). Network: in -> GELU(W1 x + b1) -> GELU(W2 h1 + b2)
// -> W3 h2 + b3, loss mean((y - t)^2) over the batch and outputs.
//
// Two modes:
//  exact (bf16 = 0): every value in FP64; pinned by central finite
//    differences of the loss in the tests;
//  bf16  (bf16 = 1): BF16 rounding (round to nearest even) at the points the
//    GPU kernels round: inputs, weights, hidden activations, and the layer
//    gradients that feed the products (bias gradients and the GELU'
//    product use the unrounded values), FP64 accumulation elsewhere.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

double bf16r(double x) {  // FP64 -> FP32 -> BF16 (round to nearest even) -> FP64
    float f = (float)x;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return x;  // inf / nan unchanged
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    std::memcpy(&f, &u, 4);
    return (double)f;
}
double gelu(double z) { return 0.5 * z * (1.0 + std::erf(z / std::sqrt(2.0))); }
double gelu_d(double z) {
    return 0.5 * (1.0 + std::erf(z / std::sqrt(2.0))) + z * std::exp(-0.5 * z * z) / std::sqrt(2.0 * M_PI);
}

struct Net {
    int d[4];
    std::vector<double> W[3], b[3];
    Net(const int* dims, const float* p, bool rb) {
        for (int i = 0; i < 4; i++) d[i] = dims[i];
        size_t o = 0;
        for (int l = 0; l < 3; l++) {
            W[l].resize((size_t)d[l + 1] * d[l]);
            b[l].resize(d[l + 1]);
            for (auto& x : W[l]) x = rb ? bf16r(p[o++]) : p[o++];
            for (auto& x : b[l]) x = p[o++];
        }
    }
};

// forward for one row; z[l] pre-activations, h[l] layer inputs (h[0] = x)
void fwd(const Net& n, const double* x, bool rb, std::vector<double> h[3], std::vector<double> z[3]) {
    h[0].assign(x, x + n.d[0]);
    if (rb)
        for (auto& v : h[0]) v = bf16r(v);
    for (int l = 0; l < 3; l++) {
        z[l].assign(n.d[l + 1], 0.0);
        for (int r = 0; r < n.d[l + 1]; r++) {
            double s = 0.0;
            for (int k = 0; k < n.d[l]; k++) s += n.W[l][(size_t)r * n.d[l] + k] * h[l][k];
            z[l][r] = s + n.b[l][r];
        }
        if (l < 2) {
            h[l + 1].resize(n.d[l + 1]);
            for (int r = 0; r < n.d[l + 1]; r++) h[l + 1][r] = rb ? bf16r(gelu(z[l][r])) : gelu(z[l][r]);
        }
    }
}

}  // namespace

extern "C" {

double orc_bf16_round(double x) { return bf16r(x); }

// Y (B x d3) = network(X (B x d0)); params in the layout of spdk_mlp_set_params.
int orc_mlp_forward(const int* dims, const float* params, const double* X, int B, int bf16, double* Y) {
    Net n(dims, params, bf16 != 0);
    std::vector<double> h[3], z[3];
    for (int m = 0; m < B; m++) {
        fwd(n, X + (size_t)m * dims[0], bf16 != 0, h, z);
        for (int r = 0; r < dims[3]; r++) Y[(size_t)m * dims[3] + r] = z[2][r];
    }
    return 0;
}

// Gradient of mean((y - t)^2) in the parameter layout; returns the loss.
double orc_mlp_grad(const int* dims, const float* params, const double* X, const double* T, int B, int bf16,
                    double* grad) {
    const bool rb = bf16 != 0;
    Net n(dims, params, rb);
    std::vector<double> gW[3], gb[3];
    for (int l = 0; l < 3; l++) {
        gW[l].assign(n.W[l].size(), 0.0);
        gb[l].assign(n.b[l].size(), 0.0);
    }
    const double inv = 1.0 / ((double)B * dims[3]);
    double loss = 0.0;
    std::vector<double> h[3], z[3];
    for (int m = 0; m < B; m++) {
        fwd(n, X + (size_t)m * dims[0], rb, h, z);
        std::vector<double> dz(dims[3]), dzr;
        for (int r = 0; r < dims[3]; r++) {
            double e = z[2][r] - T[(size_t)m * dims[3] + r];
            loss += e * e * inv;
            dz[r] = 2.0 * e * inv;
        }
        for (int l = 2; l >= 0; l--) {
            dzr = dz;
            if (rb)
                for (auto& v : dzr) v = bf16r(v);
            for (int r = 0; r < n.d[l + 1]; r++) {
                gb[l][r] += dz[r];
                for (int k = 0; k < n.d[l]; k++) gW[l][(size_t)r * n.d[l] + k] += dzr[r] * h[l][k];
            }
            if (l == 0) break;
            std::vector<double> dh(n.d[l], 0.0);
            for (int k = 0; k < n.d[l]; k++) {
                double s = 0.0;
                for (int r = 0; r < n.d[l + 1]; r++) s += dzr[r] * n.W[l][(size_t)r * n.d[l] + k];
                dh[k] = s * gelu_d(z[l - 1][k]);
            }
            dz = dh;
        }
    }
    size_t o = 0;
    for (int l = 0; l < 3; l++) {
        for (double v : gW[l]) grad[o++] = v;
        for (double v : gb[l]) grad[o++] = v;
    }
    return loss;
}

}  // extern "C"
