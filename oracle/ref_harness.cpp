/*
 * ref_harness.cpp -- drives the UNMODIFIED reference CPU solver
 * 3DCPUwithPML_3DArray_GridIncludePMLandOrder/ElasticWaveEquation3D.cpp,
 * compiled from where it lies under the reference root (see Makefile, target
 * "ref").  TEST INFRASTRUCTURE ONLY: it produces golden vectors that pin the
 * oracle restatement in sg_oracle_impl.h; nothing ships or runs it on a GPU.
 *
 * Macro-level adjustments made here, before the reference TU is included (no
 * reference file is edited or copied):
 *  - DEF_3D_VelocityStress is re-declared variadic: the reference destructor
 *    (ElasticWaveEquation3D.cpp:43-46) calls the 2-argument macro
 *    (ElasticMacro.h:140) with 5 arguments, which GCC rejects.  The expansion
 *    is unchanged (it declares nine pointers).
 *  - private members are made public so the harness can read the arrays.
 *  - With -DSG_WORKLOAD_CONSTANTS the three constants that differ from the
 *    stated workload are overridden: DIFF_10's coefficients (exact rationals),
 *    DAMPING0 (d0 = -3 V ln(R) / (2 P delta)) with a run-time V, and the Ricker
 *    delay (ceil() -> 0, so the peak sits at 1/f0).
 *
 * I/O (little-endian):
 *   in : int32 nxi nyi nzi pml nt sx sy sz nst; float32 dx dy dz dt f0 dampv;
 *        int32 stations[3*nst]; float32 b[N] lam[N] mu[N]   (x-fastest, total grid)
 *   out: float32 wavelet[nt]; traces[nt][nst][9]; fields[9][N] (x-fastest);
 *        damping_x[NX] (line j=k=5), damping_y[NY], damping_z[NZ]
 * The third argument is a scratch directory: calculateDamping() writes its
 * dampingX/Y/Z.txt dumps into the working directory.
 *
 * Timing mode (fourth argument "time"): damping stays zero (the update
 * arithmetic does not depend on its values, so the work per step is the
 * same), no file is written by the reference, and only the wall time of the
 * step loop is reported, as "seconds <t>" on stderr; out receives nothing.
 */
#include <chrono>
#include <unistd.h>
#include <string>
#include <vector>

#include "ElasticMacro.h"

#undef DEF_3D_VelocityStress
#define DEF_3D_VelocityStress(type, name, ...)   \
    DEF_3D_POINTER(type, stressXX##name)         \
    DEF_3D_POINTER(type, stressYY##name)         \
    DEF_3D_POINTER(type, stressZZ##name)         \
    DEF_3D_POINTER(type, stressXY##name)         \
    DEF_3D_POINTER(type, stressXZ##name)         \
    DEF_3D_POINTER(type, stressYZ##name)         \
    DEF_3D_POINTER(type, velocityX##name)        \
    DEF_3D_POINTER(type, velocityY##name)        \
    DEF_3D_POINTER(type, velocityZ##name)

#ifdef SG_WORKLOAD_CONSTANTS
static float sg_damp_v = 3000.0f;
static const float SGC0 = (float)(19845.0 / 16384.0);
static const float SGC1 = (float)(-735.0 / 8192.0);
static const float SGC2 = (float)(567.0 / 40960.0);
static const float SGC3 = (float)(-405.0 / 229376.0);
static const float SGC4 = (float)(35.0 / 294912.0);
#undef DIFF_10
#define DIFF_10(phiF, phiB1, phiF1, phiB2, phiF2, phiB3, phiF3, phiB4, phiF4, phiB5, delta)        \
    ((SGC0 * (phiF - phiB1) + SGC1 * (phiF1 - phiB2) + SGC2 * (phiF2 - phiB3) + SGC3 * (phiF3 - phiB4) \
      + SGC4 * (phiF4 - phiB5)) / delta)
#undef DAMPING0
#define DAMPING0(V, pml, delta) \
    ((float)(-3.0 * (double)(V) * log((double)R) / (2.0 * (double)(pml) * (double)(delta))))
#undef Vs
#define Vs sg_damp_v
#define ceil(x) 0.0f
#endif

#define private public
#include "ElasticWaveEquation3D.cpp"
#undef private

static bool rd(FILE *f, void *p, size_t n) { return fread(p, 1, n, f) == n; }
static bool wr(FILE *f, const void *p, size_t n) { return fwrite(p, 1, n, f) == n; }

int main(int argc, char **argv)
{
    if (argc != 4 && argc != 5) {
        fprintf(stderr, "usage: %s in.bin out.bin scratch_dir\n", argv[0]);
        return 2;
    }
    FILE *fi = fopen(argv[1], "rb");
    if (!fi) return 3;
    int hdr[9];
    float fh[6];
    if (!rd(fi, hdr, sizeof hdr) || !rd(fi, fh, sizeof fh)) return 4;
    const int nxi = hdr[0], nyi = hdr[1], nzi = hdr[2], pml = hdr[3], nt = hdr[4];
    const int nst = hdr[8];
    const int NX = nxi + 2 * pml + ORDER, NY = nyi + 2 * pml + ORDER, NZ = nzi + 2 * pml + ORDER;
    const size_t N = (size_t)NX * NY * NZ;
    std::vector<int> st(3 * (size_t)nst);
    std::vector<float> model(3 * N);
    if (nst && !rd(fi, st.data(), st.size() * sizeof(int))) return 5;
    if (!rd(fi, model.data(), model.size() * sizeof(float))) return 6;
    fclose(fi);
#ifdef SG_WORKLOAD_CONSTANTS
    sg_damp_v = fh[5];
#endif

    DEF_3D_POINTER(float, buoyancy);
    DEF_3D_POINTER(float, lambda);
    DEF_3D_POINTER(float, mu);
    ALLOCATE_3D(float, buoyancy, NX, NY, NZ);
    ALLOCATE_3D(float, lambda, NX, NY, NZ);
    ALLOCATE_3D(float, mu, NX, NY, NZ);
    for (int k = 0; k < NZ; ++k)
        for (int j = 0; j < NY; ++j)
            for (int i = 0; i < NX; ++i) {
                size_t o = (size_t)i + (size_t)NX * ((size_t)j + (size_t)NY * k);
                buoyancy[i][j][k] = model[o];
                lambda[i][j][k] = model[N + o];
                mu[i][j][k] = model[2 * N + o];
            }

    GRID grid = {nxi, nyi, nzi};
    SOURCELOCATION src = {hdr[5], hdr[6], hdr[7]};
    ElasticWaveEquation3D *ewq =
        new ElasticWaveEquation3D(grid, nt, pml, buoyancy, lambda, mu, src, fh[0], fh[1], fh[2], fh[3], fh[4]);

    if (chdir(argv[3]) != 0) return 7;
    std::cout.setstate(std::ios_base::failbit); /* the reference prints every step */

    const bool timing = argc > 4 && std::string(argv[4]) == "time";
    if (timing) {
        const auto t0 = std::chrono::steady_clock::now();
        for (int it = 0; it < nt; ++it) {
            ewq->loadSeismicSource(it);
            ewq->velocityUpdate();
            ewq->stressUpdate();
        }
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::fprintf(stderr, "seconds %.6f\n", sec);
        return 0;
    }
    std::vector<float> wavelet(nt), traces((size_t)nt * nst * 9);
    /* The loop of propagate() (ElasticWaveEquation3D.cpp:53-85) without its
     * file output: damping once, then source -> velocity -> stress -> record. */
    ewq->calculateDamping();
    for (int it = 0; it < nt; ++it) {
        wavelet[it] = ewq->sourceFunction(it);
        ewq->loadSeismicSource(it);
        ewq->velocityUpdate();
        ewq->stressUpdate();
        for (int s = 0; s < nst; ++s) {
            int i = st[3 * s], j = st[3 * s + 1], k = st[3 * s + 2];
            float *row = &traces[((size_t)it * nst + s) * 9];
            row[0] = ewq->velocityX[i][j][k];
            row[1] = ewq->velocityY[i][j][k];
            row[2] = ewq->velocityZ[i][j][k];
            row[3] = ewq->stressXX[i][j][k];
            row[4] = ewq->stressYY[i][j][k];
            row[5] = ewq->stressZZ[i][j][k];
            row[6] = ewq->stressXY[i][j][k];
            row[7] = ewq->stressXZ[i][j][k];
            row[8] = ewq->stressYZ[i][j][k];
        }
    }

    FILE *fo = fopen(argv[2], "wb");
    if (!fo) return 8;
    if (!wr(fo, wavelet.data(), wavelet.size() * sizeof(float))) return 9;
    if (!wr(fo, traces.data(), traces.size() * sizeof(float))) return 9;
    float ***fields[9] = {ewq->velocityX, ewq->velocityY, ewq->velocityZ, ewq->stressXX, ewq->stressYY,
                          ewq->stressZZ,  ewq->stressXY,  ewq->stressXZ,  ewq->stressYZ};
    std::vector<float> buf(N);
    for (int f = 0; f < 9; ++f) {
        for (int k = 0; k < NZ; ++k)
            for (int j = 0; j < NY; ++j)
                for (int i = 0; i < NX; ++i)
                    buf[(size_t)i + (size_t)NX * ((size_t)j + (size_t)NY * k)] = fields[f][i][j][k];
        if (!wr(fo, buf.data(), N * sizeof(float))) return 9;
    }
    std::vector<float> line;
    for (int i = 0; i < NX; ++i) line.push_back(ewq->damping_x[i][5][5]);
    for (int j = 0; j < NY; ++j) line.push_back(ewq->damping_y[5][j][5]);
    for (int k = 0; k < NZ; ++k) line.push_back(ewq->damping_z[5][5][k]);
    if (!wr(fo, line.data(), line.size() * sizeof(float))) return 9;
    fclose(fo);
    return 0;
}
