"""CPU oracle for the VIE operator path (test infrastructure only).

Importable only from tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg.  The product (marie-tools_amd/) never imports it.
"""
