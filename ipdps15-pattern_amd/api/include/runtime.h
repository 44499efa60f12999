// RuntimeInit / RuntimeFinalize (kmeans/include/runtime.h:4-5).  The
// reference initialises MPI; with one GPU holding each whole domain they
// select device 0 (or $PAT_DEVICE) and wait for outstanding work.
#ifndef INCLUDE_RUNTIME_H_
#define INCLUDE_RUNTIME_H_

#include <cstdlib>

#include "../pat_check.h"

inline void RuntimeInit(int argc, char *argv[])
{
    (void)argc;
    (void)argv;
    const char *dev = getenv("PAT_DEVICE");
    PAT_API_CALL(pat_init(dev ? atoi(dev) : 0));
}

inline void RuntimeFinalize() { PAT_API_CALL(pat_finalize()); }

#endif
