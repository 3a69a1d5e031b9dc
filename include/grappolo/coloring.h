/* Drop-in for DefineStructure/coloring.h (distance-1 coloring). */
#ifndef GRAPPOLO_AMD_COLORING_H
#define GRAPPOLO_AMD_COLORING_H
#include "basic_util.h"

/* coloring.h:8 -- valid distance-1 coloring (self-loops ignored); returns the
   largest color used (the caller adds 1).  Colors come from Jones-Plassmann
   rounds, larger degree first and then a hash of the id, so they differ from
   the reference's speculative coloring; only validity is part of the
   contract. */
int algoDistanceOneVertexColoringOpt(graph *G, int *vtxColor, int nThreads, double *totTime);

/* coloring.h:12 -- coloringMultiHashMaxMin.cpp:60-226: incomplete coloring by
   nItrs iterations of nHash max/min passes; vertices left uncoloured hold the
   returned maxColor.  Random values come from a hash of the vertex id (the
   reference draws rand()). */
int algoColoringMultiHashMaxMin(graph *G, int *vtxColor, int nThreads, double *totTime, int nHash, int nItrs);

/* coloring.h:15 -- vBase.cpp:6-168: rebalances a proper coloring of ncolors
   colors in place (type 0 first fit, 1 least used) in deterministic rounds;
   returns the number of rounds. */
int vBaseRedistribution(graph *G, int *vtxColor, int ncolors, int type);
#endif
