// C-ABI export of the reference's micro_cholesky (src/micro_chol.hpp:14-44),
// compiled straight from the reference source tree by oracle/Makefile into
// oracle/_ref/.  Test infrastructure only: used to pin orc_potrf_f32_micro.
#include "micro_chol.hpp"   // resolved with -I <reference>/src

extern "C" int ref_micro_cholesky(float* A, int n, int ld) {
    return micro_cholesky(A, n, ld);
}
