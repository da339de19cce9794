"""CPU oracle for the generalized-coordinates step.

TEST INFRASTRUCTURE ONLY.  This package is a restatement of the reference's
algorithm (brax/generalized/*.py, brax/{kinematics,scan,base,math,actuator,
contact}.py, plus the third-party pieces jaxopt.ProjectedGradient and the MJX
plane collisions) in torch on the CPU, in float64 or float32.  Only tests,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU baseline leg may import it,
and only as the checker; the product path (``brax_amd``) never does.

Parity status of each part is recorded in ``oracle/generalized.py``'s header.
"""
