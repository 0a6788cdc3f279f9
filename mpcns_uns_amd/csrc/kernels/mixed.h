#pragma once
// Mixed-precision implicit induction solver (FP32 multigrid-preconditioned inner FGMRES inside FP64
// iterative refinement).
namespace mhdg
{
struct MixedSolver;
} // namespace mhdg
