"""CPU oracle for the state-vector hot path (TEST INFRASTRUCTURE ONLY).

This package is a numpy restatement of the reference's numpy backend and the
QuantumFlow layers above it that the GPU engine replaces.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker.  The product package ``quantumflow_amd``
never imports anything from here.

Every function cites the reference file:line it restates.  Functions whose
parity is not pinned by a reference known-answer test, a closed form or a
brute-force check are marked "parity unpinned" in their docstring; at present
there are none (see tests/test_oracle.py for the pins).
"""
