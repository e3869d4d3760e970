"""CPU oracle for the coupled D3Q19 lattice-Boltzmann step (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and the cpu_baseline leg of bench.py
import this package.  The product (``xadvcool_amd``) never does.
"""
