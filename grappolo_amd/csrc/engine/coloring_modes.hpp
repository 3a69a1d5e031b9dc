// coloring_modes.hpp -- coloring = 2 (balanced) and coloring = 3
// (incomplete multi-hash) of runMultiPhaseColoring.
#pragma once

#include "common.hpp"

namespace glv {

// algoColoringMultiHashMaxMin (Coloring/coloringMultiHashMaxMin.cpp:60-226)
// with hashed random values: every vertex gets a color in [0, maxColor];
// vertices left uncoloured hold maxColor.  Returns maxColor.
int color_multihash(const DevGraph &g, int32_t *colors, int nHash, int nItrs, hipStream_t s);

// vBaseRedistribution (Coloring/vBase.cpp:6-168) as deterministic rounds:
// recolors a proper coloring of ncolors colors in place (type 0: first fit,
// 1: least used).  Returns the number of rounds.
int balance_colors(const DevGraph &g, int32_t *colors, int ncolors, int type, hipStream_t s);

}  // namespace glv
