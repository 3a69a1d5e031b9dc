/*
 * grappolo_amd -- data model of the drop-in C++ API.
 *
 * These records keep the exact field order, types and sizes of Grappolo's
 * own data model (DefineStructure/defs.h:66-92) so that code written against
 * the reference compiles and links unchanged against libgrappolo_hip.so:
 *   edge   {long head; long tail; double weight;}           24 bytes (AoS)
 *   graph  {numVertices, sVertices, numEdges, edgeListPtrs, edgeList}
 *   Comm   {long size; double degree;}
 * `long` is 64-bit on every target this library builds for (LP64).
 *
 * A graph row i is edgeList[edgeListPtrs[i] .. edgeListPtrs[i+1]).  Every
 * undirected edge is stored in both rows; self-loops appear once.  For input
 * graphs numEdges counts undirected edges (half the entries); for graphs built
 * by buildNextLevelGraphOpt it is NE_out + (number of vertices)
 * (Utility/buildNextPhase.cpp:237), exactly as the reference writes it.
 */
#ifndef GRAPPOLO_AMD_DEFS_H
#define GRAPPOLO_AMD_DEFS_H

#include <stddef.h>

typedef struct comm {
    long size;     /* number of vertices in the community          */
    double degree; /* a_c: sum of the weighted degrees of members  */
} Comm;

typedef struct {
    long cid;       /* community id                                 */
    double Counter; /* weight from one vertex to that community     */
} mapElement;

typedef struct {
    long head;
    long tail;
    double weight;
} edge;

typedef struct {
    long numVertices;
    long sVertices;
    long numEdges;
    long *edgeListPtrs;
    edge *edgeList;
} graph;

#ifdef __cplusplus
static_assert(sizeof(long) == 8, "grappolo_amd requires LP64");
static_assert(sizeof(edge) == 24, "edge must be 24 bytes like the reference");
static_assert(offsetof(edge, tail) == 8 && offsetof(edge, weight) == 16, "edge layout");
static_assert(sizeof(graph) == 40, "graph layout");
static_assert(sizeof(Comm) == 16, "Comm layout");
#endif

#endif /* GRAPPOLO_AMD_DEFS_H */
