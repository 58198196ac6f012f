/*
 * sg_oracle.c -- CPU restatement of the reference split-field PML solver
 * (3DCPUwithPML_3DArray_GridIncludePMLandOrder).  TEST INFRASTRUCTURE ONLY:
 * the checker and the CPU baseline, never the product path.
 *
 * Built by oracle/Makefile into oracle/liboracle.so with -O2
 * -ffp-contract=off (plus OpenMP over the outer i loop, which does not change
 * any result: every point of a half-step is independent).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sg_oracle.h"

/* The reference's PI literal (CPUPML/ElasticMacro.h:16). */
#define ORACLE_PI 3.1415926f

void oracle_ricker(float f0, float dt, int nt, int delay_mode, float *s_out)
{
    /* sourceFunction (CPUPML/ElasticWaveEquation3D.cpp:114-119):
     *   t0  = ceil(1/(f0*dt))                         float
     *   tmp = pow(PI*f0*((it - t0)*dt - 1/f0), 2)     pow(float,int) -> double, stored float
     *   s   = (1 - 2*tmp) * exp(-tmp)                 float exp */
    float t0 = delay_mode == ORACLE_RICKER_CEIL_DELAY ? ceilf(1.0f / (f0 * dt)) : 0.0f;
    for (int it = 0; it < nt; ++it) {
        float arg = ORACLE_PI * f0 * (((float)it - t0) * dt - 1 / f0);
        float tmp = (float)pow((double)arg, 2.0);
        s_out[it] = (1 - 2 * tmp) * expf(-tmp);
    }
}

static inline float log10_f32(float x) { return log10f(x); }
static inline float pow_f32(float x, float y) { return powf(x, y); }
static inline double log10_f64(double x) { return log10(x); }
static inline double pow_f64(double x, double y) { return pow(x, y); }

#define REAL float
#define FN(name) name##_f32
#include "sg_oracle_impl.h"
#undef REAL
#undef FN

#define REAL double
#define FN(name) name##_f64
#include "sg_oracle_impl.h"
#undef REAL
#undef FN

void oracle_damping_profile_f32(int n, int pml, float delta, int mode, float v, float r, float tau,
                                float *out)
{
    damping_profile_f32(n, pml, delta, mode, v, r, tau, out);
}

void oracle_damping_profile_f64(int n, int pml, double delta, int mode, double v, double r,
                                double tau, double *out)
{
    damping_profile_f64(n, pml, delta, mode, v, r, tau, out);
}

int oracle_run_f32(const oracle_params *p, const float *b, const float *lam, const float *mu,
                   const float *wavelet, const int *stations, int nst, unsigned comp_mask,
                   float *st_out, float *fields_out)
{
    return run_f32(p, b, lam, mu, wavelet, stations, nst, comp_mask, st_out, fields_out);
}

int oracle_run_f64(const oracle_params *p, const double *b, const double *lam, const double *mu,
                   const double *wavelet, const int *stations, int nst, unsigned comp_mask,
                   double *st_out, double *fields_out)
{
    return run_f64(p, b, lam, mu, wavelet, stations, nst, comp_mask, st_out, fields_out);
}
