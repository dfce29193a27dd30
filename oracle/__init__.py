"""CPU oracle for the pressure-projection solve path (test infrastructure only).

Only the test suite, the smoke check and the CPU-baseline leg of the benchmark
may import this package.  The product library never calls it.
"""
