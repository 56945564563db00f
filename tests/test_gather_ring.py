"""CPU check of the LDS ring schedule of the pair-layout gather.

Restates the slot arithmetic of k_gather_lds_pair (claire_amd/csrc/gather_lds.hip,
PairGeo and the step lambda) for the D <= 2 geometry and walks a chunk step by
step: every stencil plane a step reads must sit in the slot the kernel computes
for it and must have been landed before the last barrier, and no slot may be
landed and read between the same two barriers (waves of one block run the
steps of a period without synchronizing).  A one-slot-smaller ring must fail
the second condition, so the check can fail.
"""
import pytest

D = 2
HLO, HHI = D + 1, D + 2
WIN = HLO + HHI + 1


def walk(period, w, chunk):
    slots = {k: (k - HLO, -1) for k in range(WIN + period - 1)}  # prologue, planes relative to x1s
    epoch = 0
    reads, writes = {}, {}
    for step in range(chunk):
        cs = step % w  # slot of plane step - HLO
        for o1 in range(-HLO, HHI - 3 + 1):  # first stencil plane x1 + o1
            sw = min(max(cs + HLO + o1, 0), 2 * w - 4)
            sw -= w if sw >= w else 0
            for a in range(4):
                s = sw + a
                s -= w if s >= w else 0  # the kernel's base[a] wrap
                plane, landed = slots.get(s, (None, None))
                assert plane == step + o1 + a, f"step {step}: slot {s} holds {plane}, not {step + o1 + a}"
                assert landed < epoch, f"step {step}: plane {plane} landed in this period"
                reads.setdefault(epoch, set()).add(s)
        spare = cs + WIN + period - 1
        spare -= w if spare >= w else 0
        slots[spare] = (step + HHI + period, epoch)
        writes.setdefault(epoch, set()).add(spare)
        if step % period == period - 1:
            epoch += 1
    for e, ws in writes.items():
        both = ws & reads.get(e, set())
        assert not both, f"period {e}: slots {sorted(both)} landed and read between the same barriers"


@pytest.mark.parametrize("period", [1, 2])
def test_ring_schedule(period):
    w = WIN + 2 * period - 1
    for chunk in range(1, 50):
        walk(period, w, chunk)


def test_smaller_ring_is_caught():
    with pytest.raises(AssertionError, match="landed and read"):
        walk(2, WIN + 2, 20)
