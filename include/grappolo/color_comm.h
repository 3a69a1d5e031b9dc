/* Drop-in for DefineStructure/color_comm.h: the colored multi-phase Louvain. */
#ifndef GRAPPOLO_AMD_COLOR_COMM_H
#define GRAPPOLO_AMD_COLOR_COMM_H
#include "basic_comm.h"
#include "coloring.h"

/* color_comm.h:7 -- takes ownership of G (frees it); writes C_orig[NV]. */
void runMultiPhaseColoring(graph *G, long *C_orig, int coloring, int numColors, int replaceMap,
                           long minGraphSize, double threshold, double C_threshold,
                           int numThreads, int threadsOpt);

/* color_comm.h:10 -- one colored phase; C is output (initialised inside). */
double algoLouvainWithDistOneColoring(graph *G, long *C, int nThreads, int *color, int numColor,
                                      double Lower, double thresh, double *totTime, int *numItr);

/* color_comm.h:13 -- same result as the map variant (same tie rule). */
double algoLouvainWithDistOneColoringNoMap(graph *G, long *C, int nThreads, int *color,
                                           int numColor, double Lower, double thresh,
                                           double *totTime, int *numItr);
#endif
