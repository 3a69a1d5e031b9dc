/* Drop-in for the cluster-comparison part of
 * DefineStructure/utilityClusteringFunctions.h (C++: takes std::vector). */
#ifndef GRAPPOLO_AMD_UTILITY_CLUSTERING_FUNCTIONS_H
#define GRAPPOLO_AMD_UTILITY_CLUSTERING_FUNCTIONS_H
#include <vector>

#include "defs.h"

/* utilityClusteringFunctions.h:71 -- utilityClusterComparisonMetrics.cpp:50-304:
   prints the comparison report of truth C1 against C2 (N1 == N2); both
   vectors are relabelled in place (-1 handling of the reference).  Pair
   counts are the reference's at one thread. */
void computeCommunityComparisons(std::vector<long> &C1, long N1, std::vector<long> &C2, long N2);

/* utilityClusteringFunctions.h:73 -- sorts colorSize in place; see basic_util.h */
double computeGiniCoefficient(long *colorSize, int numColors);
#endif
