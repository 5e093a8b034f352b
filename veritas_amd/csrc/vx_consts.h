// Physical constants with the reference's truncated digits (veritas.hpp:22-29).
// The method's results depend on these exact values (for example
// eps0_inv * eps0 - 1 = -8.0e-6), so they are kept verbatim.
#pragma once

#define vx_eps0_inv 1.1294e+11
#define vx_eps0 8.854187817e-12
#define vx_mu_inv 795774.715482
#define vx_cs 299792458.0
#define vx_c_inv 3.33564095e-9
#define vx_DPI 6.28318530718

namespace vx {
// electron mass and elementary charge as used by the reference's example
// application (veritas.cpp:15-16)
constexpr double m_electron = 9.10938291e-31;
constexpr double e_charge = 1.60217657e-19;

// Runge-Kutta tableau shared by the FCT advance (Rectangle.cpp:1396-1517) and
// the transverse field advance (EMSolver.cpp:208-337).  Row s holds the stage
// coefficients a_{s,k}; row 5 holds the weights b_k.
constexpr double rk_a[6][6] = {
    {0.5, 0, 0, 0, 0, 0},
    {0.221776, 0.110224, 0, 0, 0, 0},
    {-0.04884659515311857, -0.17772065232640102, 0.8465672474795197, 0, 0, 0},
    {-0.15541685842491548, -0.3567050098221991, 1.0587258798684427, 0.30339598837867193, 0, 0},
    {0.2014243506726763, 0.008742057842904185, 0.15993995707168115, 0.4038290605220775,
     0.22606457389066084, 0},
    {0.15791629516167136, 0.0, 0.18675894052400077, 0.6805652953093346, -0.27524053099500667,
     0.25},
};
}  // namespace vx
