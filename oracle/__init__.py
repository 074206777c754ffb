"""CPU FP64 oracle for the Takagi-Taupin hot path.  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package, and only as the reference results are compared
against (or the timed CPU baseline):
the product (dyfox_amd) never calls it, and has no CPU fallback.

    volterra.py      H = 2 pi g.u of straight dislocations (NumPy), after
                     the reference's dislocations.py:68-90, 113-168, 171-211
    laue.py          the Laue depth march in the reference's direct form,
                     vectorised over (s_dev, rows, x), after
                     laue/kernels.py:48-352 and laue/solver.py:34-118
    bragg.py         the Bragg shooting solve: _march_down, perfect_symbol,
                     the Fourier preconditioner and SciPy GMRES, after
                     bragg/solver.py:28-60, 135-379
    closed_forms.py  Pendelloesung (discrete M^(Nz-1) and continuous),
                     semi-infinite and finite-thickness Riccati
                     reflectivities in closed form, Darwin width

How each function is pinned (golden vectors, closed forms, invariants)
is recorded beside its tests in tests/test_oracle_*.py.
"""
