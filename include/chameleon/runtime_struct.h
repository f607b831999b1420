/*
 * chameleon/runtime_struct.h -- sequence and request records of the
 * asynchronous API (reference include/chameleon/runtime_struct.h:78-103).
 * A sequence groups *_Tile_Async submissions that are waited on together;
 * its status is the first non-zero status any of them reported.
 */
#ifndef CHAMELEON_AMD_RUNTIME_STRUCT_H
#define CHAMELEON_AMD_RUNTIME_STRUCT_H

#ifdef __cplusplus
extern "C" {
#endif

#define RUNTIME_PRIORITY_MIN 0

typedef struct runtime_request_s {
    int   status;
    int   flush;
    int   priority;
    void *parent;
} RUNTIME_request_t;

#define RUNTIME_REQUEST_INITIALIZER { 0, 1, RUNTIME_PRIORITY_MIN, 0 }

typedef struct runtime_sequence_s {
    int                status;
    int                myrank;
    RUNTIME_request_t *request;
    void              *schedopt;  /* device-side state of the sequence */
    int                comm;
} RUNTIME_sequence_t;

#ifdef __cplusplus
}
#endif

#endif
