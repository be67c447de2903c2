"""Host logic of the benchmark workloads (CPU): the uneven set's sides and
the pass sizes of the stacked engines."""
from _pkg import tomo  # noqa: F401  (loads the package as tomo_amd)
import tomo_amd.workloads as W


def test_c4_sides_are_the_seeded_draw():
    sides = W.c4_sides()
    assert len(sides) == 64 and set(sides) <= set(W.C4_SIDES)
    counts = {n: sides.count(n) for n in W.C4_SIDES}
    assert counts == {128: 11, 192: 12, 256: 16, 384: 4, 512: 21}
    assert W.c4_sides() == sides  # seeded


def test_stack_pass_shots():
    # 512^2, 4098 levels, checkpoint 128: 197 levels of 1 MiB per shot; 2688 shots
    per = W.stack_pass_shots(512, 4098, 128, 2688, 190e9)
    assert per == 896 and 2688 % per == 0
    levels = 2 * (-(-4098 // 128)) + 128 + 5
    assert per * levels * 512 * 512 * 4 <= 190e9
    # a stack that fits runs in one pass
    assert W.stack_pass_shots(256, 2050, 128, 1024, 190e9) == 0
    # passes stay equal and within the budget when 128-shot rounds do not fit
    per = W.stack_pass_shots(512, 4098, 128, 2688, 100e9)
    assert per * levels * 512 * 512 * 4 <= 100e9 and per * (-(-2688 // per)) >= 2688
    # a pass never holds fewer than one shot
    assert W.stack_pass_shots(512, 4098, 128, 10, 1.0) == 1
