// Device-side reduction entry points shared by the result, sampling and
// adjoint code (reduce.hip).
#pragma once

#include "common.h"

namespace bsim {

struct PauliMask {
    uint64_t x = 0, y = 0, z = 0;
    int ny = 0;
};

// One term of a Pauli sum H = sum_w c_w P_w on the device: P_w |i> =
// phase_w(i) |i ^ f>, f = x | y bits, with (-1) for each y bit of i that is
// 0 and each z bit that is 1, times i^ny (lambda_i = sum_w c_w phase_w(i)
// psi_{i ^ f_w}).
struct PWord {
    uint64_t f, y, z;
    int32_t ny, pad_;
    double c;
};

__device__ __forceinline__ double2 pphase(uint64_t i, const PWord& w) {
    int neg = (__popcll(~i & w.y) + __popcll(i & w.z)) & 1;
    double s = neg ? -w.c : w.c;
    switch (w.ny & 3) {
        case 0: return make_double2(s, 0);
        case 1: return make_double2(0, s);
        case 2: return make_double2(-s, 0);
        default: return make_double2(0, -s);
    }
}

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
