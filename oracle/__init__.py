"""CPU oracle for the getrates path -- test infrastructure only.

Imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg, as the
checker; the product path (chemgen_amd) never imports it.
"""
