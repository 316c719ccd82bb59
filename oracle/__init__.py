"""CPU oracle for the FIT leapfrog and wake/impedance path.

TEST INFRASTRUCTURE ONLY. This package is a numpy/scipy restatement of the
reference Wakis v0.7.0 algorithm (the scipy CSR path of ``SolverFIT3D`` plus
``wakesolve`` and the ``WakeSolver`` post-processing). It exists to check the
HIP product path. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it. The product package
``wakis_amd`` never imports it, and nothing here is ever the thing measured
as GPU throughput.

Pinning status of each function (what fixes its numbers besides itself):

* ``grid_metrics``            closed form on uniform grids (tests/test_oracle_core.py)
* ``bc_column_row_masks``     brute force against the literal assignment order,
                              plus test_002 (PEC box resonance)
* ``assemble_operators``      dense-matrix brute force of the curl identities
                              (C is a discrete curl: div(curl)=0, curl(grad)=0)
* ``OracleSolver.one_step``   test_002 closed form (TE101 frequency within 5%),
                              lossless energy invariant, test_001 golden
* ``Beam`` update             closed form (Gaussian line charge integrates to q)
* ``wake_long_WP_3d``         impulse/constant closed forms, test_001 golden WP
* ``wake_trans_WP``           closed form (linear-in-x WP gives constant WPx)
* ``wake_long_Z``/``trans_Z`` numpy.fft as the textbook DFT, test_001 golden Z
* ``init_pml``                closed-form geomspace profile, test_005 invariant
* ``emsolve`` ABC copy        the reference's snapshot protocol (get_abc /
                              update_abc with the n-2 rotation) run literally on
                              numpy views of tiny grids; copied snapshots (a true
                              two-step delay) are shown to give other fields
"""
