// compare.hpp -- cluster-comparison metrics on the device
// (Utility/utilityClusterComparisonMetrics.cpp:50-330).
#pragma once

#include "common.hpp"

namespace glv {

struct CompareResult {
    i64 nC1 = 0, nC2 = 0;
    i64 SameSame = 0, SameDiff = 0, DiffSame = 0, Agree = 0;
    double precision = 0, recall = 0, fScore = 0, gini1 = 0, gini2 = 0;
};

// computeCommunityComparisons on device labels C1 (truth) and C2 of N
// vertices each.  Both are relabelled in place as the reference does (-1 in
// C1 becomes a new last community, -1 in C2 becomes N + 1).  The pair counts
// are the reference's at one thread.
void compare_communities(i64 *C1, i64 *C2, i64 N, CompareResult &r, hipStream_t s);

// computeGiniCoefficient of n device sizes, sorted ascending in place;
// numColors enters the formula as the reference's int parameter.
double gini_coefficient(i64 *d_sizes, i64 n, int numColors, hipStream_t s);

}  // namespace glv
