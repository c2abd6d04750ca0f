// problems.hpp -- initial conditions used by the oracle (test infrastructure).
//
// Orszag-Tang: mhd2d_glm_cpu.cpp:486-505 (identical in mhd2d_glm_gpu.cu:443-451).
// kh (reference Kelvin-Helmholtz): mhd2d_glm_gpu.cu:421-441.
// khb: the Kelvin-Helmholtz variant of the 1024^2 benchmark workload
//   (rho=1, p=2.5, vx=0.5*(tanh((y-0.25)/0.05)-tanh((y-0.75)/0.05)-1),
//    vy=0.01*sin(4 pi x)*(exp(-(y-0.25)^2/0.005)+exp(-(y-0.75)^2/0.005)),
//    Bx=0.1); it has no reference implementation, so its arithmetic is this
//   file's and the product host mirrors it.
// otv: Orszag-Tang with the ensemble workload's variations, params = {B
//   amplitude, phases added to the sine arguments of vx, vy, Bx, By}; with
//   {1, 0, 0, 0, 0} it is `ot` bit for bit (no reference counterpart either).
#ifndef ORACLE_PROBLEMS_HPP
#define ORACLE_PROBLEMS_HPP

#include <cmath>
#include <string>

namespace oracle {

// Fills P[9] (primitives + psi) at cell (i,j) of an nx*ny grid on [0,1]^2.
template <class R>
inline void initial_primitive(const std::string& problem, int i, int j, R dx,
                              R dy, R P[9], const double* params = nullptr) {
  const R x_min = (R)0.0, y_min = (R)0.0;
  const R x = x_min + ((R)i + (R)0.5) * dx;
  const R y = y_min + ((R)j + (R)0.5) * dy;
  const R pi = std::acos((R)-1.0);
  const R two_pi = (R)2.0 * pi;
  if (problem == "kh") {
    const R a = (R)0.025, sigma = (R)0.05, delta = (R)0.01;
    const R S = (R)0.5 * (std::tanh((y - (R)0.25) / a) -
                          std::tanh((y - (R)0.75) / a));
    const R g1 = (y - (R)0.25) / sigma;
    const R g2 = (y - (R)0.75) / sigma;
    P[0] = (R)1.0 + S;
    P[1] = -(R)0.5 + S;
    P[2] = delta * std::sin((R)2.0 * two_pi * x) *
           (std::exp(-(R)0.5 * g1 * g1) + std::exp(-(R)0.5 * g2 * g2));
    P[3] = (R)0.0;
    P[4] = (R)2.5;
    P[5] = (R)0.2;
    P[6] = (R)0.0;
    P[7] = (R)0.0;
    P[8] = (R)0.0;
  } else if (problem == "khb") {
    const R w = (R)0.05;
    const R g1 = y - (R)0.25, g2 = y - (R)0.75;
    P[0] = (R)1.0;
    P[1] = (R)0.5 * (std::tanh(g1 / w) - std::tanh(g2 / w) - (R)1.0);
    P[2] = (R)0.01 * std::sin((R)2.0 * two_pi * x) *
           (std::exp(-(g1 * g1) / (R)0.005) + std::exp(-(g2 * g2) / (R)0.005));
    P[3] = (R)0.0;
    P[4] = (R)2.5;
    P[5] = (R)0.1;
    P[6] = (R)0.0;
    P[7] = (R)0.0;
    P[8] = (R)0.0;
  } else if (problem == "otv") {
    const R b = (R)params[0];
    const R inv_sqrt_4pi = (R)1.0 / std::sqrt((R)4.0 * pi);
    P[0] = (R)25.0 / ((R)36.0 * pi);
    P[1] = -std::sin(two_pi * y + (R)params[1]);
    P[2] = std::sin(two_pi * x + (R)params[2]);
    P[3] = (R)0.0;
    P[4] = (R)5.0 / ((R)12.0 * pi);
    P[5] = -b * std::sin(two_pi * y + (R)params[3]) * inv_sqrt_4pi;
    P[6] = b * std::sin((R)2.0 * two_pi * x + (R)params[4]) * inv_sqrt_4pi;
    P[7] = (R)0.0;
    P[8] = (R)0.0;
  } else {
    const R inv_sqrt_4pi = (R)1.0 / std::sqrt((R)4.0 * pi);
    P[0] = (R)25.0 / ((R)36.0 * pi);
    P[1] = -std::sin(two_pi * y);
    P[2] = std::sin(two_pi * x);
    P[3] = (R)0.0;
    P[4] = (R)5.0 / ((R)12.0 * pi);
    P[5] = -std::sin(two_pi * y) * inv_sqrt_4pi;
    P[6] = std::sin((R)2.0 * two_pi * x) * inv_sqrt_4pi;
    P[7] = (R)0.0;
    P[8] = (R)0.0;
  }
}

}  // namespace oracle

#endif  // ORACLE_PROBLEMS_HPP
