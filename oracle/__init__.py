"""CPU oracle of the hot path -- test infrastructure only.

Imported only by tests/, the smoke check and the CPU-baseline leg of
bench.py, as the checker; the product path never imports it.
"""
