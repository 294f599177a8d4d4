// Sizes shared by the device code and the host-side scratch layout.
#pragma once

namespace glx {

constexpr int MAX_GRID_BLOCKS = 4096;  // workgroups of a persistent launch
constexpr int NREL = 16;               // grid-barrier release words
constexpr int REL_STRIDE = 16;         // u64 elements between release words (128 bytes)

}  // namespace glx
