"""CPU checkers for the HIP Stokes library (test infrastructure only).

``oracle.oracle``    ctypes binding of the C restatement (liboracle.so).
``oracle.reference`` ctypes binding of the reference's own Fortran compiled
                     from source by ``make -C oracle ref`` (oracle/_ref/).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package; the product library never does.
"""
