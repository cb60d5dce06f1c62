"""The oracle: a CPU restatement of the reference solver and its schedule.

Test infrastructure only -- loaded by tests/, the smoke check and the
benchmark's CPU-baseline leg as the checker, never by the library.
"""
