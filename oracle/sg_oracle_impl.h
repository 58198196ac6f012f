/*
 * sg_oracle_impl.h -- precision-generic body of the oracle, included twice by
 * sg_oracle.c with REAL = float and REAL = double.  TEST INFRASTRUCTURE ONLY.
 *
 * Every arithmetic expression keeps the reference's operand order so that the
 * float build reproduces the reference CPU path's rounding (the build uses
 * -ffp-contract=off, matching an x86-64 -O2 build of the reference, which has
 * no FMA).  Paths are relative to the reference root;
 * CPUPML = 3DCPUwithPML_3DArray_GridIncludePMLandOrder.
 */

/* DIFF_10 (CPUPML/ElasticMacro.h:63-70): ((((c1*a + c2*b) + c3*c) + c4*d) + c5*e) / delta,
 * a = f[F] - f[B1] ...; the reference writes "-8.97217e-2f * (...)" as a
 * subtraction of a positive product, which is bit-identical to adding the
 * product with the negated coefficient. */
#define ORC_DIFF(c, f, F0, B0, stride, delta)                                                      \
    ((((((c)[0] * ((f)[(F0)] - (f)[(B0)]) + (c)[1] * ((f)[(F0) + (stride)] - (f)[(B0) - (stride)])) \
        + (c)[2] * ((f)[(F0) + 2 * (stride)] - (f)[(B0) - 2 * (stride)]))                          \
       + (c)[3] * ((f)[(F0) + 3 * (stride)] - (f)[(B0) - 3 * (stride)]))                           \
      + (c)[4] * ((f)[(F0) + 4 * (stride)] - (f)[(B0) - 4 * (stride)])) / (delta))

/* DIFF_X(phi, i, ...) reads phi[i..i+4] - phi[i-1..i-5] ("backward" at node i);
 * DIFF_X(phi, i+1, ...) reads phi[i+1..i+5] - phi[i..i-4] ("forward"). */
#define ORC_DBWD(c, f, p, s, h) ORC_DIFF(c, f, (p), (p) - (s), s, h)
#define ORC_DFWD(c, f, p, s, h) ORC_DIFF(c, f, (p) + (s), (p), s, h)

static void FN(damping_profile)(int n, int pml, REAL delta, int mode, REAL v, REAL r, REAL tau,
                                REAL *out)
{
    /* CPUPML/ElasticWaveEquation3D.cpp:271-296: with N the interior size,
     * i < pml+5 gives dist = pml+5-i; i >= N+pml+5 gives dist = N+pml+5-i-1
     * (negative, squared by powf).  N+pml+5 == n-pml-5 for total extent n. */
    for (int i = 0; i < n; ++i) {
        out[i] = 0;
        int dist;
        if (i < 5 || i >= n - 5) continue; /* halo: never updated */
        if (i < pml + 5)
            dist = pml + 5 - i;
        else if (i >= n - pml - 5)
            dist = (n - pml - 5) - i - 1;
        else
            continue;
        REAL ratio = (REAL)dist * (REAL)1 / ((REAL)pml * (REAL)1);
        REAL d0;
        if (mode == ORACLE_DAMP_REFERENCE) {
            /* DAMPING0 = log10(1/R) * tau * V / (pml * delta), left to right */
            d0 = FN(log10)((REAL)1 / r) * tau * v / ((REAL)pml * delta);
        } else {
            /* d0 = -3 V ln(R) / (2 L), L = pml*delta, evaluated in double */
            d0 = (REAL)(-3.0 * (double)v * log((double)r) / (2.0 * (double)pml * (double)delta));
        }
        out[i] = d0 * FN(pow)(ratio, (REAL)2);
    }
}

/* [i][j][k] index with k contiguous (the reference's float*** rows). */
#define ORC_IX(i, j, k) (((size_t)(i) * (size_t)NY + (size_t)(j)) * (size_t)NZ + (size_t)(k))
/* x-fastest index of the solver layout */
#define ORC_XF(i, j, k) ((size_t)(i) + (size_t)NX * ((size_t)(j) + (size_t)NY * (size_t)(k)))

static int FN(run)(const oracle_params *p, const REAL *b_in, const REAL *lam_in, const REAL *mu_in,
                   const REAL *wavelet, const int *stations, int nst, unsigned comp_mask,
                   REAL *st_out, REAL *fields_out)
{
    const int NX = p->nx, NY = p->ny, NZ = p->nz, P = p->pml;
    if (NX < 11 || NY < 11 || NZ < 11 || P < 0) return -1;
    if (2 * P > NX - 10 || 2 * P > NY - 10 || 2 * P > NZ - 10) return -1;
    for (int a = 0; a < 3; ++a) {
        int n = a == 0 ? NX : (a == 1 ? NY : NZ);
        if (p->src[a] < 5 || p->src[a] >= n - 5) return -2;
    }
    for (int s = 0; s < nst; ++s) {
        if (stations[3 * s] < 0 || stations[3 * s] >= NX || stations[3 * s + 1] < 0 ||
            stations[3 * s + 1] >= NY || stations[3 * s + 2] < 0 || stations[3 * s + 2] >= NZ)
            return -2;
    }
    const size_t N = (size_t)NX * NY * NZ;
    const REAL dt = (REAL)p->dt, hx = (REAL)p->dx, hy = (REAL)p->dy, hz = (REAL)p->dz;
    REAL c[5];
    for (int n = 0; n < 5; ++n) c[n] = (REAL)p->coef[n];

    /* 9 full fields + 27 splits + 3 model arrays, zero-initialised like
     * ALLOCATE_3D_VelocityStress (CPUPML/ElasticMacro.h:151-193). */
    REAL *mem = (REAL *)calloc(N * 39, sizeof(REAL));
    if (!mem) return -3;
    REAL *F[9], *S[9][3];
    for (int f = 0; f < 9; ++f) F[f] = mem + N * f;
    for (int f = 0; f < 9; ++f)
        for (int a = 0; a < 3; ++a) S[f][a] = mem + N * (9 + 3 * f + a);
    REAL *B = mem + N * 36, *LAM = mem + N * 37, *MU = mem + N * 38;
    for (int k = 0; k < NZ; ++k)
        for (int j = 0; j < NY; ++j)
            for (int i = 0; i < NX; ++i) {
                B[ORC_IX(i, j, k)] = b_in[ORC_XF(i, j, k)];
                LAM[ORC_IX(i, j, k)] = lam_in[ORC_XF(i, j, k)];
                MU[ORC_IX(i, j, k)] = mu_in[ORC_XF(i, j, k)];
            }

    REAL *dpx = (REAL *)calloc((size_t)NX + NY + NZ, sizeof(REAL));
    if (!dpx) { free(mem); return -3; }
    REAL *dpy = dpx + NX, *dpz = dpy + NY;
    /* calculateDamping's 3-D arrays are separable: damping_x[i][j][k] is the
     * profile of i wherever it is set and 0 elsewhere, and the profile is 0
     * for interior i, so a 1-D profile per axis reproduces them exactly. */
    FN(damping_profile)(NX, P, hx, p->damp_mode, (REAL)p->damp_v, (REAL)p->damp_r, (REAL)p->damp_tau, dpx);
    FN(damping_profile)(NY, P, hy, p->damp_mode, (REAL)p->damp_v, (REAL)p->damp_r, (REAL)p->damp_tau, dpy);
    FN(damping_profile)(NZ, P, hz, p->damp_mode, (REAL)p->damp_v, (REAL)p->damp_r, (REAL)p->damp_tau, dpz);

    REAL *vx = F[OR_VX], *vy = F[OR_VY], *vz = F[OR_VZ];
    REAL *sxx = F[OR_SXX], *syy = F[OR_SYY], *szz = F[OR_SZZ];
    REAL *sxy = F[OR_SXY], *sxz = F[OR_SXZ], *syz = F[OR_SYZ];
    const size_t si = (size_t)NY * NZ, sj = (size_t)NZ, sk = 1;
    const size_t src = ORC_IX(p->src[0], p->src[1], p->src[2]);
    int ncomp = 0;
    for (int f = 0; f < 9; ++f) ncomp += (comp_mask >> f) & 1u;

#ifdef _OPENMP
    int nthr = p->nthreads > 0 ? p->nthreads : 1;
#endif

    /* the source point lies in the undamped interior box (no PML axis) */
    int src_interior = 1;
    for (int a = 0; a < 3; ++a) {
        const int n = a == 0 ? NX : (a == 1 ? NY : NZ);
        if (p->src[a] < 5 + P || p->src[a] >= n - 5 - P) src_interior = 0;
    }
    for (int it = 0; it < p->nt; ++it) {
        /* loadSeismicSource (ElasticWaveEquation3D.cpp:121-130) */
        if (p->inject_mode == ORACLE_INJECT_DOUBLE) {
            double amp = (double)p->dt * (double)p->dt * (double)wavelet[it];
            sxx[src] = (REAL)((double)sxx[src] + amp);
            syy[src] = (REAL)((double)syy[src] + amp);
            szz[src] = (REAL)((double)szz[src] + amp);
        } else if (p->inject_mode == ORACLE_INJECT_ACCUMULATE) {
            /* The CUDA path's interior sigma is unsplit (..._concise/
             * ElasticWaveEquation.cu:256-258), so what locateSource adds
             * (.cu:216-218) stays in sigma for good; its PML regions rebuild
             * sigma from the splits (ElasticMacro.h:466-543), which drops the
             * injection after one velocity update, as here.  In this
             * split-everywhere form the interior case adds the amount to one
             * split as well, where the rebuild keeps it. */
            REAL amp = dt * dt * wavelet[it];
            sxx[src] += amp;
            syy[src] += amp;
            szz[src] += amp;
            if (src_interior) {
                S[OR_SXX][0][src] += amp;
                S[OR_SYY][0][src] += amp;
                S[OR_SZZ][0][src] += amp;
            }
        } else {
            REAL amp = dt * dt * wavelet[it];
            sxx[src] += amp;
            syy[src] += amp;
            szz[src] += amp;
        }

        /* velocityUpdate (ElasticWaveEquation3D.cpp:134-183) */
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthr) schedule(static)
#endif
        for (int i = 5; i < NX - 5; i++) {
            for (int j = 5; j < NY - 5; j++) {
                for (int k = 5; k < NZ - 5; k++) {
                    const size_t o = ORC_IX(i, j, k);
                    const REAL hdt = (REAL)0.5 * dt;
                    const REAL mX = (REAL)1 - hdt * dpx[i], mY = (REAL)1 - hdt * dpy[j], mZ = (REAL)1 - hdt * dpz[k];
                    const REAL pX = (REAL)1 + hdt * dpx[i], pY = (REAL)1 + hdt * dpy[j], pZ = (REAL)1 + hdt * dpz[k];
                    REAL db;
                    /* BUOYANCYBARX = (b[i] + b[i+1]) / 2 (ElasticMacro.h:101-104) */
                    db = dt * ((B[o] + B[o + si]) / 2);
                    S[OR_VX][0][o] = (mX * S[OR_VX][0][o] + db * ORC_DFWD(c, sxx, o, si, hx)) / pX;
                    S[OR_VX][1][o] = (mY * S[OR_VX][1][o] + db * ORC_DBWD(c, sxy, o, sj, hy)) / pY;
                    S[OR_VX][2][o] = (mZ * S[OR_VX][2][o] + db * ORC_DBWD(c, sxz, o, sk, hz)) / pZ;
                    vx[o] = S[OR_VX][0][o] + S[OR_VX][1][o] + S[OR_VX][2][o];

                    db = dt * ((B[o] + B[o + sj]) / 2);
                    S[OR_VY][0][o] = (mX * S[OR_VY][0][o] + db * ORC_DBWD(c, sxy, o, si, hx)) / pX;
                    S[OR_VY][1][o] = (mY * S[OR_VY][1][o] + db * ORC_DFWD(c, syy, o, sj, hy)) / pY;
                    S[OR_VY][2][o] = (mZ * S[OR_VY][2][o] + db * ORC_DBWD(c, syz, o, sk, hz)) / pZ;
                    vy[o] = S[OR_VY][0][o] + S[OR_VY][1][o] + S[OR_VY][2][o];

                    db = dt * ((B[o] + B[o + sk]) / 2);
                    S[OR_VZ][0][o] = (mX * S[OR_VZ][0][o] + db * ORC_DBWD(c, sxz, o, si, hx)) / pX;
                    S[OR_VZ][1][o] = (mY * S[OR_VZ][1][o] + db * ORC_DBWD(c, syz, o, sj, hy)) / pY;
                    S[OR_VZ][2][o] = (mZ * S[OR_VZ][2][o] + db * ORC_DFWD(c, szz, o, sk, hz)) / pZ;
                    vz[o] = S[OR_VZ][0][o] + S[OR_VZ][1][o] + S[OR_VZ][2][o];
                }
            }
        }

        /* stressUpdate (ElasticWaveEquation3D.cpp:185-249) */
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthr) schedule(static)
#endif
        for (int i = 5; i < NX - 5; i++) {
            for (int j = 5; j < NY - 5; j++) {
                for (int k = 5; k < NZ - 5; k++) {
                    const size_t o = ORC_IX(i, j, k);
                    const REAL hdt = (REAL)0.5 * dt;
                    const REAL mX = (REAL)1 - hdt * dpx[i], mY = (REAL)1 - hdt * dpy[j], mZ = (REAL)1 - hdt * dpz[k];
                    const REAL pX = (REAL)1 + hdt * dpx[i], pY = (REAL)1 + hdt * dpy[j], pZ = (REAL)1 + hdt * dpz[k];
                    const REAL dX = ORC_DBWD(c, vx, o, si, hx);
                    const REAL dY = ORC_DBWD(c, vy, o, sj, hy);
                    const REAL dZ = ORC_DBWD(c, vz, o, sk, hz);
                    const REAL lam = LAM[o], mu = MU[o];

                    S[OR_SXX][0][o] = (mX * S[OR_SXX][0][o] + dt * (lam + (REAL)2 * mu) * dX) / pX;
                    S[OR_SXX][1][o] = (mY * S[OR_SXX][1][o] + dt * lam * dY) / pY;
                    S[OR_SXX][2][o] = (mZ * S[OR_SXX][2][o] + dt * lam * dZ) / pZ;
                    sxx[o] = S[OR_SXX][0][o] + S[OR_SXX][1][o] + S[OR_SXX][2][o];

                    S[OR_SYY][0][o] = (mX * S[OR_SYY][0][o] + dt * lam * dX) / pX;
                    S[OR_SYY][1][o] = (mY * S[OR_SYY][1][o] + dt * (lam + (REAL)2 * mu) * dY) / pY;
                    S[OR_SYY][2][o] = (mZ * S[OR_SYY][2][o] + dt * lam * dZ) / pZ;
                    syy[o] = S[OR_SYY][0][o] + S[OR_SYY][1][o] + S[OR_SYY][2][o];

                    S[OR_SZZ][0][o] = (mX * S[OR_SZZ][0][o] + dt * lam * dX) / pX;
                    S[OR_SZZ][1][o] = (mY * S[OR_SZZ][1][o] + dt * lam * dY) / pY;
                    S[OR_SZZ][2][o] = (mZ * S[OR_SZZ][2][o] + dt * (lam + (REAL)2 * mu) * dZ) / pZ;
                    szz[o] = S[OR_SZZ][0][o] + S[OR_SZZ][1][o] + S[OR_SZZ][2][o];

                    /* MUBAR = 4 / (1/mu1 + 1/mu2 + 1/mu3 + 1/mu4) (ElasticMacro.h:106-114) */
                    const REAL muxy = (REAL)4 / ((REAL)1 / MU[o] + (REAL)1 / MU[o + si] + (REAL)1 / MU[o + sj] + (REAL)1 / MU[o + si + sj]);
                    const REAL muxz = (REAL)4 / ((REAL)1 / MU[o] + (REAL)1 / MU[o + si] + (REAL)1 / MU[o + sk] + (REAL)1 / MU[o + si + sk]);
                    const REAL muyz = (REAL)4 / ((REAL)1 / MU[o] + (REAL)1 / MU[o + sj] + (REAL)1 / MU[o + sk] + (REAL)1 / MU[o + sj + sk]);

                    S[OR_SXY][0][o] = (mX * S[OR_SXY][0][o] + dt * muxy * ORC_DFWD(c, vy, o, si, hx)) / pX;
                    S[OR_SXY][1][o] = (mY * S[OR_SXY][1][o] + dt * muxy * ORC_DFWD(c, vx, o, sj, hy)) / pY;
                    S[OR_SXY][2][o] = mZ * S[OR_SXY][2][o] / pZ;
                    sxy[o] = S[OR_SXY][0][o] + S[OR_SXY][1][o] + S[OR_SXY][2][o];

                    S[OR_SXZ][0][o] = (mX * S[OR_SXZ][0][o] + dt * muxz * ORC_DFWD(c, vz, o, si, hx)) / pX;
                    S[OR_SXZ][1][o] = mY * S[OR_SXZ][1][o] / pY;
                    S[OR_SXZ][2][o] = (mZ * S[OR_SXZ][2][o] + dt * muxz * ORC_DFWD(c, vx, o, sk, hz)) / pZ;
                    sxz[o] = S[OR_SXZ][0][o] + S[OR_SXZ][1][o] + S[OR_SXZ][2][o];

                    S[OR_SYZ][0][o] = mX * S[OR_SYZ][0][o] / pX;
                    S[OR_SYZ][1][o] = (mY * S[OR_SYZ][1][o] + dt * muyz * ORC_DFWD(c, vz, o, sj, hy)) / pY;
                    S[OR_SYZ][2][o] = (mZ * S[OR_SYZ][2][o] + dt * muyz * ORC_DFWD(c, vy, o, sk, hz)) / pZ;
                    syz[o] = S[OR_SYZ][0][o] + S[OR_SYZ][1][o] + S[OR_SYZ][2][o];
                }
            }
        }

        /* record after the stress update (ElasticWaveEquation3D.cpp:67-69) */
        if (st_out) {
            REAL *row = st_out + (size_t)it * nst * ncomp;
            for (int s = 0; s < nst; ++s) {
                const size_t o = ORC_IX(stations[3 * s], stations[3 * s + 1], stations[3 * s + 2]);
                int q = 0;
                for (int f = 0; f < 9; ++f)
                    if ((comp_mask >> f) & 1u) row[(size_t)s * ncomp + q++] = F[f][o];
            }
        }
    }

    if (fields_out) {
        for (int f = 0; f < 9; ++f)
            for (int k = 0; k < NZ; ++k)
                for (int j = 0; j < NY; ++j)
                    for (int i = 0; i < NX; ++i)
                        fields_out[N * f + ORC_XF(i, j, k)] = F[f][ORC_IX(i, j, k)];
    }
    free(dpx);
    free(mem);
    return 0;
}

#undef ORC_IX
#undef ORC_XF
#undef ORC_DIFF
#undef ORC_DBWD
#undef ORC_DFWD
