// queue.h -- the batched executor's work-queue mapping, shared by the
// persistent kernels (device) and the library's host copy used to simulate
// the queues.
//
// A queue is the concatenation of the tile ranges of njobs jobs:
// prefix[0] = 0 <= prefix[1] <= ... <= prefix[njobs] = total items (jobs
// may have no tiles); item it < total belongs to the last job j with
// prefix[j] <= it (hence never to an empty job), as tile it - prefix[j].
//
// Claiming: the items are split into QRANGES contiguous ranges (range x =
// [x per, min((x + 1) per, total)), per = ceil(total / QRANGES)), one atomic
// counter each.  A workgroup starts at the range of its XCD (workgroup ids
// go round-robin over the XCDs), so the workgroups an XCD runs together take
// neighbouring tiles, which share operands in that XCD's L2; when a claim in
// its current range fails, the range is drained and it moves on to the next
// one, and after QRANGES moves it is done.  Every item is claimed exactly
// once (each counter hands out each of its values once), every workgroup
// makes exactly one failing claim per range, so counter x ends at the size
// of range x + grid, and the counters sum to total + QRANGES grid.
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

constexpr int QRANGES = 8;   // one item range per XCD
constexpr int QPITCH = 16;   // the ranges' device counters sit 128 bytes apart

// One claim attempt of a workgroup at range `cur` having moved `moved`
// times: the claimed item; -1 when range cur is drained (cur and moved
// advance); -2 when every range has been found drained.  add(x) increments
// range x's counter and returns its previous value.
template <class Add>
__host__ __device__ inline int64_t queue_claim_once(Add add, int64_t total, int &cur, int &moved)
{
    if (moved >= QRANGES) return -2;
    const int64_t per = (total + QRANGES - 1) / QRANGES;
    const int64_t c = (int64_t)add(cur), it = (int64_t)cur * per + c;
    if (c < per && it < total) return it;
    cur = cur + 1 == QRANGES ? 0 : cur + 1;
    moved++;
    return -1;
}

}  // namespace mgx
