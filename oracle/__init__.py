"""CPU oracle of the global-placement iteration -- TEST INFRASTRUCTURE ONLY.

A from-scratch numpy restatement of the reference's CPU (`--gpu -1`) path and of the device kernels' arithmetic,
used to check the HIP path. Only `tests/`, `__graft_entry__.smoke()` and the `cpu_baseline` leg of `bench.py` may
import it, and only as the checker or the CPU baseline: the product (`autoplace_amd`) never calls it.

Modules and what pins each one (tests/test_oracle_*.py):
    wa.py        WA wirelength value / gradient / HPWL     closed form of a 2-pin net, invariants, FP64 brute
                                                             force, central finite differences, reference goldens
    density.py   density map, overflow, field gather       total charge, brute-force overlap, aligned-node
                                                             patterns, linearity, reference goldens
    poisson.py   spectral field (mirrored FFT and DCT)     DCT closed form vs mirrored FFT, cosine mode with an
                                                             analytic field, uniform map -> zero field
    control.py   BB step, scheduler, stop rules (FP64),    reference ParamScheduler goldens, hand-derived cases;
                 Nesterov step, BB norms                   elementwise algebra against hand-computed values
    cluster.py   star-matching clustering (torch CPU)      reference cluster-id goldens, integer invariants
    place.py     the whole level / two-level flow (CPU)    reference full-run goldens on the 10k-cell design
"""
