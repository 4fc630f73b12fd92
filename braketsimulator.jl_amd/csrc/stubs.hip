// Entry points whose engines are not built yet report BSIM_ENOTSUP.
#include "common.h"

using namespace bsim;

extern "C" {

int bsim_batch_run(const bsim_task*, int, double*) { return fail(BSIM_ENOTSUP, "not built"); }

}
