// ORACLE — TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
// Basis, quadrature, operator matrices and geometry.
//
// Pinned by tests/test_oracle_matrices.py: Gauss-Legendre nodes/weights
// against numpy.polynomial.legendre.leggauss, Lagrange interpolation
// exactness on polynomials of degree <= p, derivative exactness,
// cv_to_sp * sp_to_cv = I, sp_to_cv rows integrating monomials exactly,
// invader * ader = I, and the hydro golden (which fixes the matrices'
// combined effect to 1e-6).
#include "oracle_core.hpp"

namespace orc {

// src/polynomials.cpp:3-11
double lagrange(const double* x, double y, int i, int n) {
    double lag = 1.;
    for (int j = 0; j < n; j++)
        if (j != i) lag *= (y - x[j]) / (x[i] - x[j]);
    return lag;
}

// src/polynomials.cpp:21-36
double lagrange_prime(const double* x, double y, int i, int n) {
    double lagp = 0;
    for (int k = 0; k < n; k++) {
        if (k == i) continue;
        double lag = 1;
        for (int j = 0; j < n; j++)
            if (j != i && j != k) lag *= (y - x[j]) / (x[i] - x[j]);
        lagp += lag / (x[i] - x[k]);
    }
    return lagp;
}

// src/polynomials.cpp:13-19 (rows: destination points, cols: source basis)
void lagrange_matrix(Mat& m, const double* xa, const double* xb, int na, int nb) {
    m.init(nb, na);
    for (int j = 0; j < nb; j++)
        for (int i = 0; i < na; i++) m(j, i) = lagrange(xa, xb[j], i, na);
}

// src/polynomials.cpp:38-44
void lagrange_prime_matrix(Mat& m, const double* xa, const double* xb, int na, int nb) {
    m.init(nb, na);
    for (int j = 0; j < nb; j++)
        for (int i = 0; i < na; i++) m(j, i) = lagrange_prime(xa, xb[j], i, na);
}

// src/polynomials.cpp:46-79: Newton iteration on P_n from the Tricomi-type
// initial guess cos(pi_7digits*(i+3/4)/(n+1/2)), eps 3e-14.
void gauss_legendre(double xi, double xf, int n, double* x, double* w) {
    const double eps = 3E-14;
    int m = (n + 1) / 2;
    double xm = 0.5 * (xf + xi);
    double xl = 0.5 * (xf - xi);
    if (n > 0) {
        for (int i = 0; i < m; i++) {
            double z = std::cos(3.141592654 * (i + 0.75) / (n + 0.5));
            double z1 = 0.0, p1 = 0, p2 = 0, p3 = 0, pp = 0;
            while (std::fabs(z - z1) > eps) {
                p1 = 1.0;
                p2 = 0.0;
                for (int j = 0; j < n; j++) {
                    p3 = p2;
                    p2 = p1;
                    p1 = ((2.0 * j + 1) * z * p2 - j * p3) / (j + 1);
                }
                pp = n * (z * p1 - p2) / (z * z - 1.0);
                z1 = z;
                z = z1 - p1 / pp;
            }
            x[i] = xm - xl * z;
            x[n - i - 1] = xm + xl * z;
            w[i] = (2.0 * xl) / ((1.0 - z * z) * pp * pp);
            w[n - i - 1] = w[i];
        }
    } else {
        x[0] = 0.5;
        w[0] = 1;
    }
}

// src/polynomials.cpp:81-87
void flux_points(double* x_fp, const double* x, int n) {
    x_fp[0] = 0.0;
    for (int i = 1; i < n + 1; i++) x_fp[i] = x[i - 1];
    x_fp[n + 1] = 1.0;
}

// src/polynomials.cpp:89-93 (Chebyshev-Gauss points on [0,1])
void solution_points(double* x_sp, int n) {
    for (int i = 0; i <= n; i++) x_sp[i] = 0.5 * (1.0 - std::cos((2 * i + 1) / (2.0 * (n + 1)) * PI));
}

// src/polynomials.cpp:95-101
void ader_matrix(Mat& m, const double* xt, const double* wt, int n) {
    m.init(n, n);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++)
            m(j, i) = lagrange(xt, 1, i, n) * lagrange(xt, 1, j, n) - lagrange_prime(xt, xt[i], j, n) * wt[i];
}

// src/polynomials.cpp:103-124: GL(p) quadrature of the source Lagrange basis
// over each control volume [x_fp[k], x_fp[k+1]], divided by its width.
void integral_matrix(Mat& m, const double* x_fp, const double* x_sp, int n_cv, int n_sp) {
    m.init(n_cv, n_sp);
    int p = n_sp - 1;
    std::vector<double> x(p > 0 ? p : 1), w(p > 0 ? p : 1);
    if (p > 0) gauss_legendre(0.0, 1.0, p, x.data(), w.data());
    for (int k = 0; k < n_cv; k++) {
        if (p > 0) gauss_legendre(x_fp[k], x_fp[k + 1], p, x.data(), w.data());
        for (int j = 0; j < n_sp; j++) {
            double integral;
            if (p > 0) {
                integral = 0.0;
                for (int i = 0; i < p; i++) integral += lagrange(x_sp, x[i], j, p + 1) * w[i];
            } else
                integral = 1.0;
            m(k, j) = integral / (x_fp[k + 1] - x_fp[k]);
        }
    }
}

// src/polynomials.cpp:126-198: Doolittle LU without pivoting, then one
// forward/back substitution per unit column.
void inverse(const Mat& A, Mat& C, int n) {
    C.init(n, n);
    std::vector<double> b(n, 0.0), d(n), x(n), B((size_t)n * n), Lm((size_t)n * n), U((size_t)n * n);
    for (int j = 0; j < n; j++) {
        b[j] = 0.0;
        for (int i = 0; i < n; i++) {
            B[i + j * n] = A(j, i);
            U[i + j * n] = 0;
            Lm[i + j * n] = 0;
        }
    }
    for (int k = 0; k < n - 1; k++)
        for (int i = k + 1; i < n; i++) {
            double coeff = B[k + i * n] / B[k + k * n];
            Lm[k + i * n] = coeff;
            for (int j = k + 1; j < n; j++) B[j + i * n] -= coeff * B[j + k * n];
        }
    for (int i = 0; i < n; i++) Lm[i + i * n] = 1.0;
    for (int j = 0; j < n; j++)
        for (int i = 0; i <= j; i++) U[j + i * n] = B[j + i * n];
    for (int k = 0; k < n; k++) {
        b[k] = 1.0;
        d[0] = b[0];
        for (int i = 1; i < n; i++) {
            d[i] = b[i];
            for (int j = 0; j < i; j++) d[i] -= Lm[j + i * n] * d[j];
        }
        x[n - 1] = d[n - 1] / U[n * n - 1];
        for (int i = n - 2; i >= 0; i--) {
            x[i] = d[i];
            for (int j = n - 1; j >= i + 1; j--) x[i] = x[i] - U[j + i * n] * x[j];
            x[i] = x[i] / U[i + i * n];
        }
        for (int i = 0; i < n; i++) C(i, k) = x[i];
        b[k] = 0.0;
    }
}

// Operator set built once per run: src/main.cpp:210-218 and
// src/hydro_ader.hpp:97-141 (induction_ader.hpp:129-158 is identical).
void Ops::build(int p_, bool ader_int) {
    p = p_;
    xgl.assign(p > 0 ? p : 1, 0.0);
    wgl.assign(p > 0 ? p : 1, 0.0);
    gauss_legendre(0.0, 1.0, p, xgl.data(), wgl.data());
    x_sp.assign(p + 1, 0.0);
    x_fp.assign(p + 2, 0.0);
    flux_points(x_fp.data(), xgl.data(), p);
    solution_points(x_sp.data(), p);
    xx.assign(p + 1, 0.0);
    wx.assign(p + 1, 0.0);
    gauss_legendre(0.0, 1.0, p + 1, xx.data(), wx.data());
    if (ader_int) {
        T = p + 1;
        xt = xx;
        wt = wx;
    } else {
        T = 1;
        xt.assign(1, 0.0);
        wt.assign(1, 1.0);
    }
    lagrange_matrix(sp_to_fp, x_sp.data(), x_fp.data(), p + 1, p + 2);
    lagrange_matrix(fp_to_sp, x_fp.data(), x_sp.data(), p + 2, p + 1);
    lagrange_prime_matrix(dfp_to_sp, x_fp.data(), x_sp.data(), p + 2, p + 1);
    integral_matrix(sp_to_cv, x_fp.data(), x_sp.data(), p + 1, p + 1);
    integral_matrix(fp_to_cv, x_fp.data(), x_fp.data(), p + 1, p + 2);
    inverse(sp_to_cv, cv_to_sp, p + 1);
    if (ader_int) {
        ader_matrix(ader, xt.data(), wt.data(), p + 1);
        inverse(ader, invader, p + 1);
    }
}

// src/structs.hpp:40-75: element flux-point coordinates and the FV sub-grid.
void Dim::init(int dim_, int Nglob_, int N_, int p_, int start, double L_, const double* x_fp, bool active) {
    dim = dim_;
    Nglob = Nglob_;
    p = p_;
    L = L_;
    h = L / Nglob;
    N = N_;
    if (active) {
        nsp = p + 1;
        nfp = p + 2;
        Ntot = N + 2 * NGH;
        fvn = N * nsp + 2 * nGH;
        fvnf = fvn + 1;
        idL = nsp - nGH;
        idR = fvn + idL;
    }
    sdf.init(Ntot, nfp);
    sdc.init(Ntot, nsp);
    fvf.assign(fvnf, 0.0);
    fvc.assign(fvn, 0.0);
    for (int j = 0; j < Ntot; j++) {
        for (int i = 0; i < nfp; i++) {
            sdf(j, i) = (start + j - NGH + x_fp[i]) * h;
            if ((i + j * nsp) >= idL && (i + j * nsp) < idR) fvf[i - idL + j * nsp] = sdf(j, i);
        }
        for (int i = 0; i < nsp; i++) {
            // an inactive axis has a single flux point; its centre is never read
            double right = (i + 1 < nfp) ? sdf(j, i + 1) : sdf(j, i);
            sdc(j, i) = 0.5 * (right + sdf(j, i));
            if ((i + j * nsp) >= idL && (i + j * nsp) < idR) fvc[i - idL + j * nsp] = sdc(j, i);
        }
    }
}

}  // namespace orc
