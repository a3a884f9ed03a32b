"""Host evaluation of a PlanBuilder's task arrays (test helper: checks the
plan builder independently of the device kernels)."""
from oracle import plan_exec


def apply_numpy(builder, vin, out_len):
    c1, o1, c2, o2 = builder.arrays()
    return plan_exec.execute(builder.arena_array(), c1, o1, c2, o2, builder.mid_len, vin, out_len)
