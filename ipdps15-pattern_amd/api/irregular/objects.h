// The reduction objects an irregular functor inserts into: Sobject on the
// device (the reference's LDS hash of kmeans/irregular/ros.h:12-27), Cobject
// on the host (roc.h:19).  Both keep the insert(&key, &value, reduce_ptr)
// call: the first value for a key is copied, later ones folded with the
// reduce functor.
//
// Sobject is one node's object on the device path (api/gpu/irregular_table.hpp
// gives every node a partition of its own): it holds that node's value in
// registers and refuses -- returning false and remembering it -- any insert
// for another key.
#ifndef IRREGULAR_OBJECTS_H_
#define IRREGULAR_OBJECTS_H_

#include "data_type.h"
#include "pattern_api.h"

typedef DenseObject<KEY, VALUE> Cobject;

#ifdef __HIPCC__
#include <hip/hip_runtime.h>

class Sobject {
  public:
    __device__ explicit Sobject(KEY own) : own_(own), used_(false), refused_(false) {}

    template <typename REDUCE>
    __device__ bool insert(KEY *key, VALUE *value, REDUCE reduce_ptr)
    {
        if (*key != own_) {
            refused_ = true;
            return false;
        }
        if (!used_) {
            val_ = *value;
            used_ = true;
        } else {
            reduce_ptr(&val_, value);
        }
        return true;
    }
    __device__ bool used() const { return used_; }
    __device__ bool refused() const { return refused_; }
    __device__ VALUE value() const { return val_; }

  private:
    KEY own_;
    VALUE val_;
    bool used_, refused_;
};
#endif

#endif
