// Failure behaviour of the pattern API: a failing call prints the library's
// message and exits, as the reference's CUDA_SAFE_CALL does
// (kmeans/lib/macro.h:9-16), so applications behave unchanged.
#ifndef PAT_API_CHECK_H_
#define PAT_API_CHECK_H_

#include <cstdio>
#include <cstdlib>

#include "pattern.h"

#define PAT_API_CALL(call)                                                                                             \
    do {                                                                                                               \
        int pat_rc_ = (call);                                                                                          \
        if (pat_rc_ != PAT_OK) {                                                                                       \
            fprintf(stderr, "pattern error %d in %s at %s:%d: %s\n", pat_rc_, #call, __FILE__, __LINE__,             \
                    pat_last_error());                                                                                 \
            exit(EXIT_FAILURE);                                                                                        \
        }                                                                                                              \
    } while (0)

#endif
