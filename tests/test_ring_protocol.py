"""CPU model of the four-slot LDS ring of the 128-query scan (scan_lds_kernel, knn_scan_bf16.hip).

Every wave of the workgroup stores its share of each 16-dimension chunk into the ring and reads
its own operands of a chunk from it.  The model runs one program per wave with the kernel's order
of ring operations and barriers:

    prologue: store chunks 0, 1, 2 into slots 0, 1, 2; barrier; read chunk 0 from slot 0;
              barrier                                   (the barrier after the prologue's read)
    chunk g:  read chunk g+1 from slot (g+1) % 4; store chunk g+3 into slot (g+3) % 4;
              barrier after every odd g

A read is correct when every wave's share of its slot holds the chunk it expects.  The search
walks every interleaving of the waves' programs that the barriers allow (waves run in any order
between barriers; LDS executes each wave's operations in program order) and reports the first
read that can see a wrong chunk.  The shipped order passes; the same order without the barrier
after the prologue's read (the form that shipped before) fails at chunk 1, where a fast wave
stores chunk 4 into slot 0 before a slow wave has read chunk 0.
"""
import random

import pytest

RING = 4


def program(chunks, prologue_barrier=True, barrier_every=2):
    ops = [("W", j, j) for j in range(3)] + [("B",), ("R", 0, 0)]
    if prologue_barrier:
        ops.append(("B",))
    for g in range(chunks):
        ops.append(("R", (g + 1) % RING, g + 1))
        ops.append(("W", (g + 3) % RING, g + 3))
        if g % barrier_every == barrier_every - 1:
            ops.append(("B",))
    return ops


def slot_state(progs, pcs):
    """Chunk held by each (slot, writer wave) after the executed prefixes."""
    state = {}
    for w, (prog, pc) in enumerate(zip(progs, pcs)):
        for op in prog[:pc]:
            if op[0] == "W":
                state[(op[1], w)] = op[2]
    return state


def find_violation(progs):
    """Depth-first search over all barrier-respecting interleavings; returns the first bad read
    (wave, slot, expected chunk, held chunks) or None."""
    n = len(progs)
    seen = set()
    stack = [tuple(0 for _ in progs)]
    while stack:
        pcs = stack.pop()
        if pcs in seen:
            continue
        seen.add(pcs)
        at_barrier = [pcs[w] < len(progs[w]) and progs[w][pcs[w]][0] == "B" for w in range(n)]
        if all(at_barrier):
            stack.append(tuple(p + 1 for p in pcs))
            continue
        for w in range(n):
            if pcs[w] >= len(progs[w]) or at_barrier[w]:
                continue
            op = progs[w][pcs[w]]
            if op[0] == "R":
                st = slot_state(progs, pcs)
                held = [st.get((op[1], v)) for v in range(n)]
                if any(h != op[2] for h in held):
                    return (w, op[1], op[2], held)
            nxt = list(pcs)
            nxt[w] += 1
            stack.append(tuple(nxt))
    return None


@pytest.mark.parametrize("waves", [2, 3])
@pytest.mark.parametrize("chunks", [4, 8, 12])
def test_shipped_ring_order_is_safe_under_every_interleaving(waves, chunks):
    progs = [program(chunks) for _ in range(waves)]
    assert find_violation(progs) is None


@pytest.mark.parametrize("waves", [2, 3])
def test_missing_prologue_barrier_is_caught(waves):
    # planted mistake: the order without the barrier after the prologue's read
    progs = [program(8, prologue_barrier=False) for _ in range(waves)]
    bad = find_violation(progs)
    assert bad is not None
    wave, slot, expected, held = bad
    assert slot == 0 and expected == 0 and 4 in held


def test_barrier_every_third_chunk_is_caught():
    # planted mistake: one barrier per three chunks lets a store overtake a pending read
    progs = [program(12, barrier_every=3) for _ in range(2)]
    assert find_violation(progs) is not None


def random_schedule_ok(progs, rng):
    """One random interleaving of eight waves; True if every read saw its chunk."""
    n = len(progs)
    pcs = [0] * n
    state = {}
    while any(pc < len(p) for pc, p in zip(pcs, progs)):
        at_barrier = [pcs[w] < len(progs[w]) and progs[w][pcs[w]][0] == "B" for w in range(n)]
        if all(at_barrier):
            pcs = [p + 1 for p in pcs]
            continue
        w = rng.choice([v for v in range(n) if pcs[v] < len(progs[v]) and not at_barrier[v]])
        op = progs[w][pcs[w]]
        if op[0] == "W":
            state[(op[1], w)] = op[2]
        elif op[0] == "R" and any(state.get((op[1], v)) != op[2] for v in range(n)):
            return False
        pcs[w] += 1
    return True


def test_eight_waves_random_schedules():
    rng = random.Random(0x6B4E)
    shipped = [program(24) for _ in range(8)]
    assert all(random_schedule_ok(shipped, rng) for _ in range(300))
    # the same search finds the planted mistake with eight waves
    planted = [program(24, prologue_barrier=False) for _ in range(8)]
    assert not all(random_schedule_ok(planted, rng) for _ in range(300))
