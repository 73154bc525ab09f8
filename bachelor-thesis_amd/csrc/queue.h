// queue.h -- the batched executor's work-queue mapping, shared by the
// persistent kernels (device) and the library's host copy used to simulate
// the queues.
//
// A queue is the concatenation of the tile ranges of njobs jobs:
// prefix[0] = 0 <= prefix[1] <= ... <= prefix[njobs] = total items (jobs
// may have no tiles).  Workgroups claim items from one atomic counter; item
// it < total belongs to the last job j with prefix[j] <= it (hence never to
// an empty job), as tile it - prefix[j].  A claim at or past the total tells
// the workgroup that the queue is drained, so every workgroup leaves the
// loop after exactly one failing claim and the counter ends at total + grid.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace mgx {

__host__ __device__ inline bool queue_map(const int64_t *prefix, int njobs, int64_t item, int *job, int64_t *tile)
{
    if (njobs <= 0 || item < 0 || item >= prefix[njobs]) return false;
    int lo = 0, hi = njobs - 1;   // last job with prefix <= item
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= item) lo = mid;
        else hi = mid - 1;
    }
    *job = lo;
    *tile = item - prefix[lo];
    return true;
}

}  // namespace mgx
