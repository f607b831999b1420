/*
 * chameleon/flops.h -- floating-point operation counts used to turn a time
 * into a rate.  Same formulas as the reference include/chameleon/flops.h
 * (:101-108, :300-311): a Cholesky factorisation of order n counts
 * n^3/3 + n^2/2 + n/6 flops and the two triangular solves with nrhs
 * right-hand sides count 2 n^2 nrhs.
 */
#ifndef CHAMELEON_AMD_FLOPS_H
#define CHAMELEON_AMD_FLOPS_H

static inline double flops_dpotrf(double n)
{
    return n * ((1.0 / 3.0) * n * n + 0.5 * n + 1.0 / 6.0);
}

static inline double flops_dpotrs(double n, double nrhs)
{
    return 2.0 * n * n * nrhs;
}

static inline double flops_dposv(double n, double nrhs)
{
    return flops_dpotrf(n) + flops_dpotrs(n, nrhs);
}

static inline double flops_dgemm(double m, double n, double k) { return 2.0 * m * n * k; }
static inline double flops_dsyrk(double k, double n) { return k * n * (n + 1.0); }
static inline double flops_dtrsm_left(double m, double n) { return n * m * m; }

#endif
