"""CPU oracle: restatement of the reference's CPU algorithms.

TEST INFRASTRUCTURE ONLY (the checker for tests/, the smoke check and the
CPU-baseline leg of bench.py); the product package never imports it.
"""
