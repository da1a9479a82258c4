"""CPU oracle for the DF-RCCSD(T) path: test infrastructure only.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg import
this package, as the checker; the product (``fermi.jl_amd``) never does.  There is no
``oracle/_ref``: the reference (Fermi.jl) is Julia and no Julia toolchain is available, so
nothing of it can be compiled here.
"""
