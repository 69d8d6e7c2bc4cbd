"""How the run's parameters follow the design (src/resolution.py:21-56 of the reference).

    cells, free area     -> force grid: about one cell per free bin             (force_num_bin)
    force / overflow grid -> refine                                             (refine)
    overflow / refine    -> target error of the WA model; gamma follows it      (smoothing)
    gamma                -> growth of the density weight                        (growth)
    gamma of clusters    -> hand-over to the cells                              (handover)

The device control kernel computes the same in double precision, in the same order.
"""
import math


def force_num_bin(num_nodes, free_fraction, num_bin):
    """2^round(log2 sqrt(nodes / free fraction)), never coarser than num_bin (resolution.py:21-25)."""
    return max(num_bin, 2 ** int(round(math.log2(math.sqrt(num_nodes / max(free_fraction, 1e-9))))))


def refine(density):
    """Force bins per overflow bin per side (resolution.py:28-30)."""
    return density.nbx / density.judge.nbx if density.judge is not None else 1.0


def smoothing(gamma, hpwl, wa, overflow, refine, unit, mu):
    """gamma x sqrt(target / error) within [1/m, m], m = max(mu, 1/mu); kept in [0.01, 100] x unit
    (resolution.py:33-41)."""
    error = max(1e-6, 1 - wa / max(hpwl, 1e-12))
    m = max(mu, 1.0 / mu)
    g = gamma * min(m, max(1.0 / m, math.sqrt(max(overflow, 1e-3) / refine / error)))
    return min(100 * unit, max(0.01 * unit, g))


def growth(motion, pitch, gamma, hpwl_fell):
    """1.05 x 0.98 while the HPWL falls, else 1.05 x max(1.05^(-motion pitch / gamma), 0.95) (resolution.py:44-50)."""
    if hpwl_fell:
        return 1.05 * 0.98
    return 1.05 * max(1.05 ** (-motion * pitch / gamma), 0.95)


def handover(gamma, start_gamma):
    return gamma <= start_gamma
