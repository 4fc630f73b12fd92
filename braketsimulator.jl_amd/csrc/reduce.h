// Device-side reduction entry points shared by the result, sampling and
// adjoint code (reduce.hip).
#pragma once

#include "common.h"

namespace bsim {

struct PauliMask {
    uint64_t x = 0, y = 0, z = 0;
    int ny = 0;
};

// Parse one Pauli word of nq letters; qubit q acts on bit nq-1-q.
int parse_pauli(int nq, const char* word, PauliMask& pm);
// probabilities into a device buffer of 2^nq doubles
int probabilities_device(bsim_state* h, double* dp);
// Per-bit probability sums and the total into dout[0..37).  *layout = -1:
// dout[b] is the sum over bit b; layout L >= 0: bits b < L at dout[16 + b],
// bits b >= L at dout[b - L]; dout[36] is the total.
int zsums_device(bsim_state* h, double* dout, int* layout);
// <P> (re, im) into dout2 (device)
int pauli_expect_device(bsim_state* h, const PauliMask& pm, double* dout2);
// conj(a).b (re, im) into dout2 (device)
int dot_device(const double2* a, const double2* b, int64_t len, Workspace& red, hipStream_t s, double* dout2);

}  // namespace bsim
