/* Drop-in for DefineStructure/coloring.h (distance-1 coloring). */
#ifndef GRAPPOLO_AMD_COLORING_H
#define GRAPPOLO_AMD_COLORING_H
#include "basic_util.h"

/* coloring.h:8 -- valid distance-1 coloring (self-loops ignored); returns the
   largest color used (the caller adds 1).  Colors come from Jones-Plassmann
   rounds with hash priorities, so they differ from the reference's speculative
   coloring; only validity is part of the contract. */
int algoDistanceOneVertexColoringOpt(graph *G, int *vtxColor, int nThreads, double *totTime);
#endif
