"""Test infrastructure: the CPU restatement of the reference algorithm (the
oracle).  Imported only by tests, the smoke check and the CPU-baseline leg of
the benchmark; never by the product package."""
