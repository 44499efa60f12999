// Shared index types of the pattern API (kmeans/lib/common.h:9-25).
#ifndef LIB_COMMON_H_
#define LIB_COMMON_H_

#include <stdint.h>

typedef int64_t PSIndex;
typedef int64_t IRIndex;

#endif
