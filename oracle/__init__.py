"""Correctness oracle for the 2D TMz FDTD fitness path (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
package, and only to check results or as the timed CPU baseline.
"""
