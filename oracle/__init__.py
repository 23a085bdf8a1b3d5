"""CPU oracle for the loom prove phase -- TEST INFRASTRUCTURE ONLY.

This package is a plain numpy restatement of the reference Go prover's algorithm
(Consensys loom, KoalaBear field, Poseidon2 width 24, multi-degree FRI).  It is
the checker the HIP path is compared against: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
it, and only to check or to time it beside the GPU.  Nothing in ``loom_amd``
imports it; the product path fails loudly when its HIP library is missing.

Every value is kept as a canonical integer in [0, p) (numpy uint64 arrays);
Montgomery form (the reference's in-memory ``koalabear.Element``) appears only at
the boundary (``field.to_mont`` / ``field.from_mont``) and in the proof
encoding.

Parity status per function is listed in each module's header.  Functions whose
result depends on gnark-crypto constants that no in-tree fixture fixes (the
2^24-th root of unity and the multiplicative generator, see ``field.py``) say
"parity unpinned" there.
"""
