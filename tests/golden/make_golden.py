"""Writes the golden vectors G0-G2 from the UNMODIFIED reference CPU solver.

Run in the build container (it needs the reference tree to build the harness):

    make -C oracle ref && python tests/golden/make_golden.py

The reference CPU class (3DCPUwithPML_3DArray_GridIncludePMLandOrder) is
compiled through oracle/ref_harness.cpp into oracle/_ref/; this script feeds it
inputs and stores its outputs as small .npz fixtures next to this file.

  G0  native reference constants (DIFF_10 literals, tau=3.4/V=2000 damping,
      ceil() source delay), homogeneous medium as in the reference main.cpp,
      non-cubic grid, final 9 fields + station traces + damping lines.
  G1  stated-workload constants (exact coefficients, d0=-3V ln(R)/(2L), delay
      1/f0), seeded heterogeneous medium (pins the buoyancy/harmonic-mu
      staggering), unequal dx/dy/dz, final fields + traces.
  G2  the 100^3 CPU-runnable workload (P=10, 500 steps): 20 stations x
      (vx,vy,vz,sxx) traces and the final fields on a stride-5 subsample.
Nothing else comes from the reference except one data file kept next to
this script: stationData_concise_gpu.txt, the reference's own
3DGPUwithPML_1DArraySingleGPU_concise/stationData.txt (a fixture).
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402


def homog(shape, vp, vs, rho):
    rho = np.full(shape, rho, np.float64)
    mu = rho * vs * vs
    lam = rho * vp * vp - 2 * mu
    return (1 / rho).astype(np.float32), lam.astype(np.float32), mu.astype(np.float32)


def save(name, meta, **arrays):
    path = os.path.join(HERE, name)
    np.savez_compressed(path, meta=np.array(json.dumps(meta)), **arrays)
    print("wrote", path, os.path.getsize(path), "bytes")


def g0():
    interior, pml, nt = (12, 10, 8), 3, 160
    NX, NY, NZ = [n + 2 * pml + 10 for n in interior]
    # reference main.cpp:34-39: rho 2500, V = Vs = 2000, lambda = rho V^2, mu = rho V^2 / 3
    rho = np.full((NZ, NY, NX), 2500.0, np.float32)
    V = np.float32(2000.0)
    b = (np.float32(1.0) / rho).astype(np.float32)
    lam = (rho * np.float32(V * V)).astype(np.float32)
    mu = (rho * np.float32(V * V) / np.float32(3.0)).astype(np.float32)
    src = (NX // 2 + 1, NY // 2, NZ // 2 - 1)
    st = np.array([[src[0] + 2, src[1], src[2]], [6, 7, 8], [src[0], src[1] - 3, src[2] + 1],
                   [NX - 6, NY - 6, NZ - 6], [src[0], src[1], src[2]]], np.int32)
    dx = dy = dz = 5.0
    dt, f0 = 5e-4, 30.0
    R = O.ref_run("native", interior, pml, nt, src, st, dx, dy, dz, dt, f0, 2000.0, b, lam, mu)
    meta = dict(kind="native", interior=interior, total=(NX, NY, NZ), pml=pml, nt=nt, src=src,
                dx=dx, dy=dy, dz=dz, dt=dt, f0=f0, damp_v=2000.0, rho=2500.0, V=2000.0,
                note="model: b=1/rho, lam=rho*V^2, mu=rho*V^2/3 in float32")
    save("g0_native_homog.npz", meta, stations=st, wavelet=R["wavelet"], traces=R["traces"],
         fields=R["fields"], damp_x=R["damping"][0], damp_y=R["damping"][1], damp_z=R["damping"][2])


def g1():
    interior, pml, nt = (12, 10, 8), 3, 120
    NX, NY, NZ = [n + 2 * pml + 10 for n in interior]
    rng = np.random.default_rng(20240601)
    vp = 3000.0 * (1.0 + 0.05 * np.clip(rng.standard_normal((NZ, NY, NX)), -4, 4))
    vs = vp / np.sqrt(3.0)
    rho = 310.0 * vp ** 0.25
    mu = (rho * vs * vs).astype(np.float32)
    lam = (rho * vp * vp - 2 * rho * vs * vs).astype(np.float32)
    b = (1.0 / rho).astype(np.float32)
    src = (NX // 2 - 2, NY // 2 + 1, NZ // 2)
    st = np.array([[src[0] + 1, src[1] + 1, src[2] + 1], [7, 6, 9], [NX - 7, NY // 2, NZ // 2],
                   [src[0], src[1], NZ - 7], [src[0], src[1], src[2]]], np.int32)
    dx, dy, dz = 10.0, 12.0, 8.0
    dt, f0, dampv = 1e-3, 15.0, 3000.0
    R = O.ref_run("workload", interior, pml, nt, src, st, dx, dy, dz, dt, f0, dampv, b, lam, mu)
    meta = dict(kind="workload", interior=interior, total=(NX, NY, NZ), pml=pml, nt=nt, src=src,
                dx=dx, dy=dy, dz=dz, dt=dt, f0=f0, damp_v=dampv)
    save("g1_workload_hetero.npz", meta, stations=st, b=b, lam=lam, mu=mu, wavelet=R["wavelet"],
         traces=R["traces"], fields=R["fields"], damp_x=R["damping"][0], damp_y=R["damping"][1],
         damp_z=R["damping"][2])


def g2():
    # 100^3 total incl. a 10-cell PML and the 5-cell halo, dx=10 m, dt=1 ms, 500 steps,
    # Vp 3000 / Vs 1732 / rho 2000, Ricker 15 Hz delayed 1/f0 at the centre,
    # 20 stations on a line at z=20, y=50, x=20+3r.
    interior, pml, nt = (70, 70, 70), 10, 500
    NX, NY, NZ = [n + 2 * pml + 10 for n in interior]
    b, lam, mu = homog((NZ, NY, NX), 3000.0, 1732.0, 2000.0)
    src = (50, 50, 50)
    st = np.array([[20 + 3 * r, 50, 20] for r in range(20)], np.int32)
    t = time.time()
    R = O.ref_run("workload", interior, pml, nt, src, st, 10.0, 10.0, 10.0, 1e-3, 15.0, 3000.0, b, lam, mu)
    print("G2 reference run", time.time() - t, "s")
    sub = R["fields"][:, 2::5, 2::5, 2::5]
    meta = dict(kind="workload", interior=interior, total=(NX, NY, NZ), pml=pml, nt=nt, src=src,
                dx=10.0, dy=10.0, dz=10.0, dt=1e-3, f0=15.0, damp_v=3000.0, vp=3000.0, vs=1732.0, rho=2000.0,
                subsample="fields[:, 2::5, 2::5, 2::5]", components="vx,vy,vz,sxx",
                field_sumsq=[float(np.sum(R["fields"][f].astype(np.float64) ** 2)) for f in range(9)])
    save("g2_plain100_traces.npz", meta, stations=st, wavelet=R["wavelet"], traces=R["traces"][:, :, :4],
         fields_sub=sub)


if __name__ == "__main__":
    if not (O.ref_available("native") and O.ref_available("workload")):
        sys.exit("reference harness not built: run `make -C oracle ref` first")
    g0()
    g1()
    g2()
