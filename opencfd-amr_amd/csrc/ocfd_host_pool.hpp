// Host worker threads of the library: the regrid's host work (clustering the
// tags, the interface lists of a new flux register) splits into independent
// pieces, and an AMR coarse step waits for it with the GPU idle.
#pragma once

#include <functional>

namespace ocfd {

// Number of threads parallel_for runs on (the caller included): the
// environment's OCFD_HOST_THREADS when set (1 = serial), else the machine's
// hardware threads, at most 8.
int host_threads();

// fn(t) for every t in [0, n), spread over the host threads; returns when all
// have run.  Calls from a worker (nested) run serially on that worker.
void parallel_for(int n, const std::function<void(int)>& fn);

}  // namespace ocfd
