/* Drop-in for DefineStructure/basic_util.h (the hot-path part). */
#ifndef GRAPPOLO_AMD_BASIC_UTIL_H
#define GRAPPOLO_AMD_BASIC_UTIL_H
#include "defs.h"

/* basic_util.h:5 -- first-appearance relabel in place; returns k. Negative ids kept. */
long renumberClustersContiguously(long *C, long size);

/* basic_util.h:6 -- contract the communities of Gin into Gout (rows sorted by tail).
   Gout->edgeListPtrs and Gout->edgeList are malloc()ed; the caller frees them. */
double buildNextLevelGraphOpt(graph *Gin, graph *Gout, long *C, long numUniqueClusters,
                              int nThreads);
/* basic_util.h:14 -- vertexFollowing.cpp:46-89: -1 for isolated vertices, the
   neighbour for degree-one vertices that follow it; returns the count. */
long vertexFollowing(graph *G, long *C);

/* basic_util.h:15 -- vertexFollowing.cpp:93-258: contract without zero
   self-loops, integer-truncated weights; outputs malloc()ed as above. */
double buildNewGraphVF(graph *Gin, graph *Gout, long *C, long numUniqueClusters);

/* basic_util.h:18 -- utilityClusterComparisonMetrics.cpp:309-330: sorts
   colorSize[0..numColors) ascending in place and returns the Gini
   coefficient of the sizes. */
double computeGiniCoefficient(long *colorSize, int numColors);
#endif
