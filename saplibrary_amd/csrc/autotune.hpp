// Self-written autotuner: per (kernel family, half-bandwidth, element size)
// it times the candidate values of a launch parameter on the caller's own
// data the first time that combination is set up, and keeps the fastest for
// the rest of the process.  Only parameters that do not change the
// arithmetic are tuned (tile widths of pure data movement, staging depths),
// so the results are the same whichever value wins.
#pragma once

#include "sap_internal.hpp"

#include <functional>
#include <vector>

namespace sap_amd {

enum TuneFamily {
    TUNE_B2C_TILE = 0,     // columns per workgroup of the row band -> column band conversion
    TUNE_SWEEP_NBUF = 1,   // LDS block buffers of the staged triangular sweeps
};

struct TuneEntry {
    int family = 0, k = 0, elem = 0;
    int value = 0;            // the chosen parameter
    int default_value = 0;    // what the untuned launch uses
    double best_ms = 0, default_ms = 0;
    int candidates = 0;
};

class Autotuner {
public:
    static Autotuner& instance();
    // the stored choice for (family, k, elem), or false
    bool find(int family, int k, int elem, int* value);
    // Times run(candidate) for every candidate (reps runs each, the fastest
    // counted) with HIP events on st, stores and returns the fastest.
    int tune(int family, int k, int elem, int default_value, const std::vector<int>& candidates,
             const std::function<void(int)>& run, hipStream_t st, int reps = 2);
    std::vector<TuneEntry> entries();
    void clear();
};

}  // namespace sap_amd
