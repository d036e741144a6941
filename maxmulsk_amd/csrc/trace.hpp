// Phase instrumentation of the host runtime: per-phase event counters (always
// on; an atomic increment each) and, when MAXMULSK_ROCTX=1, roctx ranges around
// the host phases so a rocprofv3 --marker-trace shows them beside the kernels.
// Counterpart of the reference's instrumentation scopes (allocation, pack_a,
// pack_b, compute, writeback; sme/support/lab_hooks.hpp:10-22) and of the
// call counts NB::counts derives from the loop structure
// (sme/v4/sme-1x4-kcout-ncblock-apack4za.cpp:448-466).
#pragma once

#include <cstdint>

namespace mmsk {

enum PhaseCounter : int {
    kCntAllocation = 0,  // device buffers allocated or grown by the library
    kCntStageIn,         // host operands copied to the device
    kCntPack,            // pack launches (operands rewritten into fragments)
    kCntCompute,         // GEMM kernel launches (tiled, packed, grouped)
    kCntReduce,          // split-K reduce launches
    kCntWriteback,       // results copied back to host memory
    kCntZeroFill,        // C zero-filled for K == 0
    kNumPhaseCounters
};

void count(PhaseCounter c, uint64_t n = 1);
uint64_t counter(int i);
const char* counter_name(int i);
void reset_counters();

// roctx range for the lifetime of the object (no-op unless enabled).
class Scope {
public:
    explicit Scope(const char* name);
    ~Scope();
    Scope(const Scope&) = delete;
    Scope& operator=(const Scope&) = delete;

private:
    bool on_;
};

}  // namespace mmsk
