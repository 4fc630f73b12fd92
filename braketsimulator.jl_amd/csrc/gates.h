// Host-side description of one gate application on the 2^vq vector and the
// launchers of the per-gate kernels (gates.hip).  The fused planner
// (fused.hip) and the adjoint sweep reuse the same description.
#pragma once

#include <vector>

#include "common.h"

namespace bsim {

struct GateSpec {
    int k = 0;                 // number of matrix bits
    int tb[16] = {0};          // vector bit position of matrix bit b
    uint64_t cmask = 0;        // control bits (vector positions)
    uint64_t cval = 0;         // required values of the control bits
    bool diag = false;         // m holds 2^k diagonal entries
    std::vector<double2> m;    // column-major 2^k x 2^k (already conjugated if requested)
};

// Validate targets/controls (reference qubit numbers) against vq and fill
// the vector bit positions.  `sorted` reorders targets to ascending vector
// bit (the k-qubit fallback convention).
int make_spec(int vq, int k, const int* targets, int nc, const int* controls, const int* cvals, const double* m,
              bool diag, bool conj, bool sorted, GateSpec& out);

// One launch applying `g` to the whole vector (the reference's per-gate
// loop shape).  `scratch` holds device copies of large matrices.
int launch_gate(double2* d, int vq, const GateSpec& g, Workspace& scratch, hipStream_t s);

// DoubleExcitation kernel (src/custom_gates.jl:390-416).
int launch_double_excitation(double2* d, int vq, const int* e4, double c, double sn, hipStream_t s);

}  // namespace bsim
