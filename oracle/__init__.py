"""CPU restatement of bevy_millions_ball's physics step (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
package, always as the checker (or the timed CPU baseline), never as the
product.  See mb_oracle.c for the reference citations and, per function, what
pins it.
"""
