/* glv_status.h -- status codes returned by every glv_* entry point.
 * The reference reports failure with assert()/exit(1) (e.g.
 * Coloring/coloringUtils.cpp:57-60); the C ABI returns these codes instead and
 * keeps a per-thread message (glv_last_error in grappolo_hip.h). */
#ifndef GLV_STATUS_H
#define GLV_STATUS_H
enum {
    GLV_OK = 0,
    GLV_E_INVALID_GRAPH = 1,  /* malformed CSR: offsets, tails or sizes out of range */
    GLV_E_OOM = 2,            /* host or device allocation failed */
    GLV_E_DEVICE = 3,         /* a HIP call failed or no device is present */
    GLV_E_TOO_MANY_COLORS = 4,/* a coloring needed more than 4096 colors */
    GLV_E_UNSUPPORTED = 5,    /* option outside the implemented path (e.g. syncType) */
    GLV_E_INVALID_ARG = 6,    /* bad scalar argument or null pointer */
    GLV_E_IO = 7              /* file could not be read or written */
};
#endif
