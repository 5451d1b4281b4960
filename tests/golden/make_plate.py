"""Writes tests/golden/plate_32x16_p{1,3}.npz: the oracle's converged steady
state of the small laminar flat plate -- 32 x 16 quads stretched to y+ = 1
at Re 1e5, M 0.2, the reference's boundary-layer VR settings, Roe + viscous
fluxes, implicit Euler with LU-SGS in the natural cell order (the
reference's sequential sweeps), local time steps at CFL 20, 1 Jacobi sweep
per iteration -- iterated until the L1 norm of the density increment is
1e-11 of the first iteration's.  The GPU test drives the multicolour LU-SGS
to the same drop and compares the steady states.

Run from the repository root: python tests/golden/make_plate.py (about a
minute per order on 8 threads)."""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

NX, NY, RE, CFL, DROP, MAX_ITER, CHUNK = 32, 16, 1e5, 20.0, 1e-11, 100000, 500


def plate_oracle(p):
    import oracle
    from dnds_amd import cases
    case = cases.flat_plate(NX, NY, reynolds=RE, rec_order=p, cfl=CFL, nthreads=4)
    st, g = case.vfv.setting, case.gas
    orc = oracle.Oracle(case.mesh.arrays(), case.vfv.arrays(), riemann_solver=1, far_field=case.far,
                        jacobi_relax=st.jacobi_relax, tang_weight=st.tang_weight,
                        tang_weight_direction=st.tang_weight_direction, mu_gas=g["muGas"], pr_gas=g["prGas"],
                        r_gas=g["Rgas"], t_ref=g["TRef"], c_sutherland=g["CSutherland"])
    return case, orc


def converge(case, orc):
    u, ur = case.u0.copy(), np.zeros((case.n_cell, 36))
    first, its, hist = None, 0, []
    while its < MAX_ITER:
        u, ur, h = orc.implicit_euler(u, ur, CHUNK, 1, CFL, dt=1e100, threshold=0.0)
        if first is None:
            first = h[0, 0]
        its += len(h)
        hist.append(h[-1, 0])
        if h[-1, 0] <= DROP * first:
            break
    return u, ur, its, first, np.array(hist)


def main():
    import oracle
    oracle.set_threads(8)
    for p in (1, 3):
        t0 = time.time()
        case, orc = plate_oracle(p)
        u, ur, its, first, hist = converge(case, orc)
        drop = hist[-1] / first
        print(f"p={p}: {its} iterations, drop {drop:.2e}, {time.time() - t0:.0f}s", flush=True)
        if drop > DROP:
            print(f"p={p}: not converged; no fixture written")
            continue
        np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), f"plate_{NX}x{NY}_p{p}.npz"),
                            u=u, urec=ur, iterations=its, first_increment=first, history=hist,
                            settings=np.array([NX, NY, RE, CFL, DROP]))


if __name__ == "__main__":
    main()
