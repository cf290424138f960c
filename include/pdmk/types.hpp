/// Geometry types of the C++ API (same layout as cpp/include/pdmk/types.hpp:9-29):
/// Mat3 is column-major, cell[j][i] = component i of lattice vector j.
#pragma once

#include <array>

namespace pdmk {

using Vec3 = std::array<double, 3>;
using Mat3 = std::array<std::array<double, 3>, 3>;
using Vec2 = std::array<double, 2>;
using Mat2 = std::array<std::array<double, 2>, 2>;

template <int DIM> struct dim_types;
template <> struct dim_types<2> { using Vec = Vec2; using Mat = Mat2; };
template <> struct dim_types<3> { using Vec = Vec3; using Mat = Mat3; };
template <int DIM> using VecND = typename dim_types<DIM>::Vec;
template <int DIM> using MatND = typename dim_types<DIM>::Mat;

}  // namespace pdmk
