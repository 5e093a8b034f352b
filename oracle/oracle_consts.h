// CPU ORACLE -- test infrastructure only.
//
// The reference's constants and Runge-Kutta tableau, restated here for the
// oracle on its own (not shared with the GPU library), so that a slip on
// either side shows up as a parity difference.
#pragma once

namespace orc {

// veritas.hpp:22-29 (SI values with the reference's truncated digits)
constexpr double kEps0Inv = 1.1294e+11;
constexpr double kEps0 = 8.854187817e-12;
constexpr double kMuInv = 795774.715482;
constexpr double kC = 299792458.0;
constexpr double kCInv = 3.33564095e-9;

// Rectangle.cpp:1396-1517 and EMSolver.cpp:208-337: the stage coefficients
// a_{s,k} (rows 0-4) and the weights b_k (row 5) of the six-stage scheme.
constexpr double kRK[6][6] = {
    {0.5, 0, 0, 0, 0, 0},
    {0.221776, 0.110224, 0, 0, 0, 0},
    {-0.04884659515311857, -0.17772065232640102, 0.8465672474795197, 0, 0, 0},
    {-0.15541685842491548, -0.3567050098221991, 1.0587258798684427, 0.30339598837867193, 0, 0},
    {0.2014243506726763, 0.008742057842904185, 0.15993995707168115, 0.4038290605220775,
     0.22606457389066084, 0},
    {0.15791629516167136, 0.0, 0.18675894052400077, 0.6805652953093346, -0.27524053099500667,
     0.25},
};

}  // namespace orc
