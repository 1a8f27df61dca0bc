// Kulla-Conty energy table of the compensated rough conductor (host side).
#pragma once

#include <vector>

namespace mrg {

constexpr int kEnergyRes = 64;  // nodes per axis: x = cos(theta_o) in [0,1], y = roughness in [0,1]

// The conductor energy table as float4 per node, [y][x] = (A, B, row average, 0): A + B is the
// directional albedo of the white GGX conductor and the row average is 2 * int E(mu) mu dmu
// (src/render/microfacet_energy.cpp:56-67, src/render/ibl.cpp:240-297). Computed once per process
// in FP64 on the host, as the reference does, and rounded to float.
const std::vector<float>& conductor_energy_table();

}  // namespace mrg
