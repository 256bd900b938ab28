// sps/thrust_vectors.hpp -- the thrust::device_vector forms of the
// reference's delegate calls (Clusterization/select_delegates.cuh:4-26),
// over the libsps C ABI.  Include it from a source compiled with hipcc:
// thrust here is rocThrust, whose device_vector owns HIP device memory.
#pragma once

#include <thrust/device_vector.h>

#include <string>

#include "search.hpp"

namespace select_delegates {

// select_delegates.cu:16-51: d_vctGroupBeginIndices[g] opens group g of
// d_vctIndices (the last group runs to its end); d_vctNumDelegates[g]
// receives the group's candidate index with the largest msnr, the first
// position on ties (thrust::max_element's rule).
inline void select_delegates(const thrust::device_vector<Cand>& d_vctCandHeap,
                             const thrust::device_vector<int>& d_vctIndices,
                             const thrust::device_vector<int>& d_vctGroupBeginIndices,
                             thrust::device_vector<int>& d_vctNumDelegates) {
    d_vctNumDelegates.resize(d_vctGroupBeginIndices.size());
    select_delegates(thrust::raw_pointer_cast(d_vctCandHeap.data()), thrust::raw_pointer_cast(d_vctIndices.data()),
                     (int)d_vctIndices.size(), thrust::raw_pointer_cast(d_vctGroupBeginIndices.data()),
                     (int)d_vctGroupBeginIndices.size(), thrust::raw_pointer_cast(d_vctNumDelegates.data()));
}

// select_delegates.cu:127-177: delegate i and the size of group i.
inline void write_log(const Cand* d_parrCand, const int QUantCandidates,
                      const thrust::device_vector<int>& d_vctGroupBeginIndices,
                      const thrust::device_vector<int>& d_vctIndices,
                      const thrust::device_vector<int>& d_vctNumDelegates, const std::string& filename,
                      const float delt_bin, const float t_chunkBegin) {
    if (d_vctNumDelegates.size() != d_vctGroupBeginIndices.size())
        throw std::invalid_argument("one delegate per group");
    write_log(d_parrCand, QUantCandidates, thrust::raw_pointer_cast(d_vctGroupBeginIndices.data()),
              (int)d_vctGroupBeginIndices.size(), (int)d_vctIndices.size(),
              thrust::raw_pointer_cast(d_vctNumDelegates.data()), filename, delt_bin, t_chunkBegin);
}

}  // namespace select_delegates
