// Drop-in counterpart of the reference's include/globalParameters.hpp: the
// mutable _GLOBAL_ settings and recompute() (globalParameters.hpp:45-222).
// The PR-EOS solve and unit conversions run in the library (mts_globals);
// this namespace mirrors the results for code that reads them.
#pragma once

#include <cmath>
#include <stdexcept>
#include <string>

#include "mantiis_hip.h"
#include "utils.hpp"

namespace _GLOBAL_ {
inline double Cl = 1.0e-9;       // resolution [m]
inline double Fbody = 1e-8;      // body force, lattice units
inline double T_phy = 300.0;     // [K]
inline double P_phy = 2.0e6;     // [Pa]
inline double mfp = 7.730846760658538e-10;
inline double rhoin = 0.235554544075989;
inline double Cf = 7.72066669163199e16;
inline double T_lu = 0.11479923396990134;
inline double a_alpha_lu = 0.03308078656450196;
inline double Crho = 57.00646021398789;
inline double Cu = 1163.7651955341285;
inline double ZG = 0.9579541318502685;
inline double Gfs = -12.7;  // solid-fluid interaction strength
inline double Gff = -1.0;   // fluid-fluid interaction strength
inline const double A1 = 1.2, A2 = 0.9;

// globalParameters.hpp:180-221
inline void recompute() {
    mts_params p;
    mts_default_params(&p);
    p.dx = Cl;
    p.Fbody = Fbody;
    p.T = T_phy;
    p.P = P_phy;
    double out[10];
    if (mts_globals(&p, out) != MTS_OK) throw std::runtime_error(mts_last_error());
    rhoin = out[0];
    mfp = out[1];
    Cf = out[2];
    T_lu = out[3];
    a_alpha_lu = out[4];
    ZG = out[5];
    Crho = out[6];
    Cu = out[7];
}
}  // namespace _GLOBAL_
