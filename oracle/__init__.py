"""CPU oracle for the Wigner split-operator hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker.  The product package ``pywignercuda_amd`` never imports it and has no
CPU fallback: it raises when its HIP library is missing.

Modules
-------
wigner2d      complex128, FFT-ordered restatement of GPU_Wigner2D_FFT.py (Lie
              step) and GPU_Wigner2D_GPitaevskii.py (Strang step), with the
              Caldeira-Leggett order-3 dissipator, the Gross-Pitaevskii shifted
              densities, the absorbing mask and the per-step observables.
cpu_notebook  real-FFT Strang propagator of CPU_GPU_Wigner2D.ipynb cells 6-11,
              the reference's own CPU path (pinned by its recorded outputs).
dirac44       complex128 restatement of GPU_WignerDirac2D_4x4.py.
"""
