"""CPU oracle for the vkzg hot path -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
package, and only as the checker.  The shipped library never imports it.
"""
