"""CPU oracle for chompack_amd -- TEST INFRASTRUCTURE ONLY.

``oracle.numeric`` restates the reference's numeric sweeps in NumPy;
``oracle.ref`` drives the reference's own C kernels built by oracle/Makefile
into oracle/_ref/.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
baseline leg may import this package; the product never does.
"""
