// Declarations of the CPU oracle (oracle/pfx_oracle.c) used as the checker
// by the C++ tests.  Test infrastructure only.
#pragma once
#include <stdint.h>
extern "C" {
int orc_c2c_batched(double *data, int64_t length, int64_t batch, int64_t stride, int64_t dist,
                    int sign, int nthreads);
int orc_r2c_inplace(double *padded, int64_t n, int64_t batch, int64_t stride, int64_t dist,
                    int nthreads);
int orc_c2r_inplace(double *padded, int64_t n, int64_t batch, int64_t stride, int64_t dist,
                    int nthreads);
int orc_parafaft_c2c(int D, const int *shape, double *data, int dir, int nthreads);
int orc_parafaft_r2c(int D, const int *real_shape, double *padded, int dir, int nthreads);
int orc_generate_shape(double *out, int D, int N, int shape_id);
void orc_uniform(double *out, uint64_t n, uint64_t seed);
}
