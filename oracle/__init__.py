"""ORACLE -- test infrastructure only (tests/, the smoke check and the CPU
baseline leg of the benchmark).  Never imported by the product."""
