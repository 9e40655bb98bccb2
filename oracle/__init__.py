"""CPU oracle for the Imogen fluid step (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker. The product library
(gpuimogen_amd/libimogen_hip.so) never loads it.
"""
