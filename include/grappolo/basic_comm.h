/* Drop-in for DefineStructure/basic_comm.h: the plain (uncolored) Louvain path. */
#ifndef GRAPPOLO_AMD_BASIC_COMM_H
#define GRAPPOLO_AMD_BASIC_COMM_H
#include "defs.h"

/* basic_comm.h:5 -- takes ownership of G (frees it); writes C_orig[NV]. */
void runMultiPhaseBasic(graph *G, long *C_orig, int basicOpt, long minGraphSize,
                        double threshold, double C_threshold, int numThreads, int threadsOpt);

/* basic_comm.h:21 -- Jacobi sweeps from singletons; returns prevMod, C = pastCommAss. */
double parallelLouvianMethod(graph *G, long *C, int nThreads, double Lower, double thresh,
                             double *totTime, int *numItr);

/* basic_comm.h:28 -- as above, starting from initCommAssOpt's greedy first move. */
double parallelLouvianMethodNoMap(graph *G, long *C, int nThreads, double Lower, double thresh,
                                  double *totTime, int *numItr);
#endif
