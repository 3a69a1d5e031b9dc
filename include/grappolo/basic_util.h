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
#endif
