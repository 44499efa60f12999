// Edge and reduction value types of the irregular-reduction API
// (kmeans/irregular/data_type.h:5-19, kmeans/irregular/parameters.h:19-28).
#ifndef IRREGULAR_DATA_TYPE_H_
#define IRREGULAR_DATA_TYPE_H_

struct edge {
    int idx0;
    int idx1;
};
typedef struct edge EDGE;

typedef int KEY;

struct value {
    float x;
    float y;
    float z;
};
typedef struct value VALUE;

#endif
