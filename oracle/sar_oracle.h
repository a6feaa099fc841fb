/*
 * sar_oracle.h -- FP64 CPU restatement of the reference SAR operators.
 * TEST INFRASTRUCTURE ONLY (the checker, never the product path).
 */
#ifndef SAR_ORACLE_H
#define SAR_ORACLE_H

#include <stddef.h>

#include "../include/sarsolver_hip.h" /* SarCalculationInfo layout */

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_NEAREST_REF 0
#define ORACLE_NEAREST 1
#define ORACLE_LINEAR 2

int oracle_fft(double *data, size_t n, int sign);
double oracle_bistatic_range(const double *tx, const double *rx, const double *x);
size_t oracle_working_len(size_t n, double upsample_ratio);
long oracle_bin(double dr, double by_dx, size_t nw, int interp, int adjoint, double *frac);
int oracle_forward(SarCalculationInfo *info, int interp, int num_threads);
int oracle_adjoint(SarCalculationInfo *info, int interp, int num_threads);
int oracle_differential_ranges(const SarCalculationInfo *info, double *dr_out);

#ifdef __cplusplus
}
#endif

#endif
