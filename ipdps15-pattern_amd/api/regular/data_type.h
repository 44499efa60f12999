// Unit descriptor and reduction output of the generalized-reduction API
// (kmeans/regular/data_type.h:21-38).
#ifndef REGULAR_DATA_TYPE_H_
#define REGULAR_DATA_TYPE_H_

#include <cstddef>

// the smallest data unit processed by one map call
struct element_t {
    size_t offset;  // byte offset into the input
    size_t size;    // bytes
};
typedef struct element_t Offset;

// key/value pairs of the reduction object after RegularStart
struct output {
    char *output_keys;
    char *output_vals;
    unsigned int *key_index;
    unsigned int *val_index;
    unsigned int key_num;
    unsigned int key_size;
    unsigned int value_size;
};

#endif
