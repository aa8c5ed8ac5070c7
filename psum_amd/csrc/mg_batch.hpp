// Batched multigrid solves: the cores of several independent fixed-cycle
// solves recorded as ops and run as batched launches (one launch per phase
// and kernel variant over all solvers).  Implemented in mg.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace psum_hip {

class HipMultigrid;

// One recorded op: the kernel body and template variant (kind), its block
// count, its argument struct, and a key that orders it (keys grow along each
// solver's sequence; ops of different solvers with equal keys are
// independent and share a launch).
struct MgRecOp {
    uint32_t key;
    uint16_t kind;
    unsigned nblocks;
    std::vector<unsigned char> args;
};

struct MgBatchPlan;

// solve() in parts: begin (replace, pin; returns the cycles' right-hand
// side), core (the cycles), end (gauge; no sync)
const double* mg_solve_begin(HipMultigrid* h, double* b);
void mg_solve_core(HipMultigrid* h, const double* b, const double* bs, double* x);
void mg_solve_end(HipMultigrid* h, const double* b, double* x);
// whether the core can be recorded (2D fused hierarchy, fixed cycle count,
// no one-workgroup tail), the right-hand side the core reads, and its ops
bool mg_recordable(HipMultigrid* h);
const double* mg_core_rhs(HipMultigrid* h, double* b);
void mg_record_core(HipMultigrid* h, const double* b, const double* bs, double* x, std::vector<MgRecOp>& ops);

MgBatchPlan* mg_batch_plan_create();
void mg_batch_plan_destroy(MgBatchPlan* p);
// merges the solvers' ops into batched launches and uploads their argument
// tables (host-synchronous; call outside any stream capture)
void mg_batch_plan_build(MgBatchPlan* p, const std::vector<const std::vector<MgRecOp>*>& per);
void mg_batch_plan_launch(const MgBatchPlan* p, hipStream_t st);
int64_t mg_batch_plan_launches(const MgBatchPlan* p);
int64_t mg_batch_plan_ops(const MgBatchPlan* p);

}  // namespace psum_hip
