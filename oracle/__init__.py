"""CPU oracle for the split-step propagation path -- TEST INFRASTRUCTURE ONLY.

This package restates, in numpy/scipy, the reference CPU path of
Quantum-Optics-LKB/NLSE v2.3.0 (pyFFTW + numba): the ``out_field`` ->
``split_step`` loop of NLSE, NLSE_1d, NLSE_3d, GPE, CNLSE, CNLSE_1d and
DDGPE.  It exists to check the HIP engine in ``nlse_amd``.

Rules for its use:
  * only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
    leg of ``bench.py`` may import it, and only as the checker / the timed
    CPU baseline -- never as a code path of the product;
  * ``nlse_amd`` never imports it (a test asserts this).

Two numeric modes:
  * ``"c64"``  -- reference-CPU semantics: complex64 storage, single
    precision FFTs, FP64 elementwise intermediates exactly where numba
    promotes (``NLSE/kernels_cpu.py``), numpy-1.26 value-based casting in the
    host expressions (the reference pins numpy 1.26.4, requirements.txt:3);
  * ``"c128"`` -- the FP64 oracle: the same scheme carried out in complex128
    (propagator not cast to complex64, V' kept in float64).

Pinning (what checks each function; see tests/test_oracle_*.py):
  * kernels.*: closed forms evaluated independently in mpmath-free FP64 /
    special cases (Isat=inf -> s=1, alpha=0 preserves |A|, V=None == V=0,
    Rabi pi/2 swap, two half steps == one step);
  * propagators: the reference tests' closed forms (test_nlse.py:32-50,
    test_nlse_1d.py:28-37, test_nlse3d.py:34-59, test_gpe.py:25-50,
    test_cnlse_1d.py:29-49, test_ddgpe.py:203-235);
  * FFT: numpy.fft (independent pocketfft build) and DFT-by-matrix on tiny
    sizes;
  * whole loop: plane-wave closed-form recursion, exact linear diffraction
    in the DFT basis (n2=0, alpha=0), exponential loss decay, norm
    conservation (test_nlse.py:287-311), Lie/Strang convergence order, the
    step-count rule brute force;
  * nonlocal kernel: scipy.signal.convolve(mode="same") of the K0 profile.
"""

from . import kernels  # noqa: F401
from .solvers import (  # noqa: F401
    CNLSE,
    CNLSE_1d,
    DDGPE,
    GPE,
    NLSE,
    NLSE_1d,
    NLSE_3d,
    count_steps,
)
