// sg_stencil.hpp -- point-update pieces shared by the simple and tiled sweeps.
#pragma once

#include <hip/hip_runtime.h>

#include "sg_precision.hpp"

namespace sg {

// Split-field PML update of one component at one point.
//   fold : the stored component before the update
//   inc  : per contributing axis, the undamped increment dt*coef*D_a
//   dm   : whether that axis damps here
//   sp   : split arrays of the contributing axes (indexed by `slot`)
// Damped axes:   f_a <- A_a f_a + B_a inc_a   (== ((1-dt d/2) f_a + inc_a)/(1+dt d/2),
//                3DCPUwithPML.../ElasticWaveEquation3D.cpp:153-158)
// Result: the sum of the new splits when every axis damps (as the reference
// rebuilds the component, .cpp:159), else the undamped remainder
// (fold - old damped splits), plus undamped increments, plus new splits.
template <class C, int NA>
__device__ __forceinline__ C pml_update(C fold, const C (&inc)[NA], const bool (&dm)[NA], C *const (&sp)[NA],
                                        long long slot, const C (&A)[NA], const C (&B)[NA])
{
    C so[NA], sn[NA];
    bool all = true;
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        all = all && dm[a];
        if (dm[a]) {
            so[a] = sp[a][slot];
            sn[a] = A[a] * so[a] + B[a] * inc[a];
            sp[a][slot] = sn[a];
        } else {
            so[a] = (C)0;
            sn[a] = (C)0;
        }
    }
    if (all) {
        C r = sn[0];
#pragma unroll
        for (int a = 1; a < NA; ++a) r += sn[a];
        return r;
    }
    C r = fold;
#pragma unroll
    for (int a = 0; a < NA; ++a) r = dm[a] ? r - so[a] : r + inc[a];
#pragma unroll
    for (int a = 0; a < NA; ++a)
        if (dm[a]) r += sn[a];
    return r;
}

// The same update with the split arrays accessed in the global address space
// (for callers whose pointers come from a descriptor in device memory).
template <class C, int NA>
__device__ __forceinline__ C pml_update_g(C fold, const C (&inc)[NA], const bool (&dm)[NA], C *const (&sp)[NA],
                                          long long slot, const C (&A)[NA], const C (&B)[NA])
{
    typedef __attribute__((address_space(1))) C gC;
    C so[NA], sn[NA];
    bool all = true;
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        all = all && dm[a];
        if (dm[a]) {
            gC *p = (gC *)sp[a];
            so[a] = p[slot];
            sn[a] = fmad(A[a], so[a], B[a] * inc[a]);
            p[slot] = sn[a];
        } else {
            so[a] = (C)0;
            sn[a] = (C)0;
        }
    }
    if (all) {
        C r = sn[0];
#pragma unroll
        for (int a = 1; a < NA; ++a) r += sn[a];
        return r;
    }
    C r = fold;
#pragma unroll
    for (int a = 0; a < NA; ++a) r = dm[a] ? r - so[a] : r + inc[a];
#pragma unroll
    for (int a = 0; a < NA; ++a)
        if (dm[a]) r += sn[a];
    return r;
}

}  // namespace sg
