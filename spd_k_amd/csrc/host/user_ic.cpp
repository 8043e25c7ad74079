// User-defined initial condition, selected at runtime with problem=user
// (the hook of src/user_ic.hpp:28-39). Edit this function and rebuild.
//
// It returns PRIMITIVE variables for the point (x, y, z): var 0 density,
// 1..3 velocities, 4 pressure. Inactive directions have y = 0 / z = 0; ay/az
// say whether y/z are active. The driver averages the returned values over
// each sub-cell with Gauss-Legendre quadrature, so discontinuous profiles are
// fine. All <problem> parameters arrive through pp (amp, v1..v3, d0, d1, p0,
// p1, radius, sigma, dir).
//
// Default: a Gaussian pulse in v_y varying only in x, carried by a uniform
// flow v1 at uniform density and pressure (a pure shear layer).
#include <cmath>
#include "setup.hpp"

namespace spdk_host {

double user_ic(int var, double x, double y, double z, double gm, bool ay, bool az, const ProblemParams& pp) {
    (void)y; (void)z; (void)gm; (void)ay; (void)az;
    double dx = x - pp.radius;
    double pulse = pp.amp * std::exp(-dx * dx / (2 * pp.sigma * pp.sigma));
    if (var == 0) return pp.d0;
    if (var == 1) return pp.v1;
    if (var == 2) return pulse;
    if (var == 3) return 0.0;
    if (var == 4) return pp.p0;
    return 0;
}

}  // namespace spdk_host
