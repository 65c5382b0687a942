"""FP64 NumPy restatement of the BATS reference kernels -- TEST INFRASTRUCTURE ONLY.

This package is the checker, never the product. Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it; the ``bats_amd`` package never does, and its GPU path fails loudly
when the HIP library is missing instead of falling back to anything here.

Numerical model ("FP64 arithmetic, f32 storage"): every array the reference
keeps as a float32 device array (exps, solver outputs a/x/t/z, errors after the
reset-chain scan, f1/f2, the per-sample weight gradient, losses, Adam moments)
is rounded to float32 at the same point here; every scalar local variable of a
reference kernel (cumul_a, cumul_b, the discriminant, the root, the spike time,
the running error sums) is a float64. Reduction order inside a kernel follows
the reference where the order is defined (the solver's event loop, the reset
chain) and is NumPy's own where the reference leaves it to a library reduction.

Pins (each checked in ``tests/test_oracle_*.py``):
  * ``lif.pre_exps``            closed form ``np.exp`` (compute_pre_exps_tests.py:19-23 style)
  * ``lif.sort_spikes``         permutation / ascending / neuron = slot // K invariants
  * ``lif.lif_forward``         independent brute-force integrator of the membrane
                                (``integrator.py``), closed forms for one input spike,
                                the constant-drive multi-spike sequence, the seed-42
                                reference test inputs regenerated with RandomState
  * ``lif.recurrent_errors``    the reference test's O(N^2) formula on its hand arrays
  * ``lif.factors``             f1 = dt/da, f2 = -dt/db by FP64 central differences
  * ``lif.weights_gradient`` /
    ``lif.pre_spike_errors``    masked brute force and FP64 central differences of
                                sum(e * t) with respect to weights and input times
  * ``losses``                  closed forms + the TTFS hand arrays
  * ``optim``                   closed form + convergence of 0.5*|w|^2
  * ``events.events_to_per_neuron``  the hand KAT of convert_to_spike_per_neuron_tests.py
"""
