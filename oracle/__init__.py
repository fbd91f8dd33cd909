"""CPU oracle for the Natrix fluid step (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
import this package.  The product path (``natrix_amd``) never does.
"""
