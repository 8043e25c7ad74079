// ORACLE — TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
// ctypes entry points for the tests: operator matrices.
#include "oracle_core.hpp"

extern "C" {

// Fills the operator matrices of degree p (row-major):
// x_sp[p+1], x_fp[p+2], xt[T], wt[T], sp_to_fp[(p+2)(p+1)], dfp_to_sp[(p+1)(p+2)],
// sp_to_cv[(p+1)^2], cv_to_sp[(p+1)^2], fp_to_cv[(p+1)(p+2)], invader[T^2]
// with T = p+1 (ADER).
int orc_matrices(int p, double* x_sp, double* x_fp, double* xt, double* wt, double* sp_to_fp, double* dfp_to_sp,
                 double* sp_to_cv, double* cv_to_sp, double* fp_to_cv, double* ader, double* invader) {
    orc::Ops op;
    op.build(p, true);
    auto cp = [](const std::vector<double>& v, double* o) { for (size_t i = 0; i < v.size(); i++) o[i] = v[i]; };
    cp(op.x_sp, x_sp);
    cp(op.x_fp, x_fp);
    cp(op.xt, xt);
    cp(op.wt, wt);
    cp(op.sp_to_fp.a, sp_to_fp);
    cp(op.dfp_to_sp.a, dfp_to_sp);
    cp(op.sp_to_cv.a, sp_to_cv);
    cp(op.cv_to_sp.a, cv_to_sp);
    cp(op.fp_to_cv.a, fp_to_cv);
    cp(op.ader.a, ader);
    cp(op.invader.a, invader);
    return 0;
}

int orc_gauss_legendre(double a, double b, int n, double* x, double* w) {
    orc::gauss_legendre(a, b, n, x, w);
    return 0;
}

}  // extern "C"
