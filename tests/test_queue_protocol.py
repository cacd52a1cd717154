"""Model check of the keypoint kernels' work-queue protocol (WorkQueue in
sift-wgpu_amd/csrc/kernels_detect.hip), on the CPU.

The waves of one work-group are explored over every interleaving of their
steps, with the hardware barrier as it behaves on CDNA: it releases once every
wave that has not ended has arrived, whichever barrier instruction each one is
at. Thread 0 (in wave 0) takes items from a device counter; the other waves
learn the item through an LDS slot. Items may be left early (a keypoint with no
peak, or one for another instantiation).

- The protocol the kernels use (two slots; the next item is taken into the slot
  no one reads, at the start of an item; every item, left early or not, ends at
  a barrier; the slot is read after it) gives every wave the same item sequence
  and ends every wave, in every interleaving.
- A one-slot form in which an item left early skips the barrier that ends the
  item lets thread 0 overwrite the slot before a slower wave has read it: waves
  then work on different items and pair barriers of different items, which
  this check finds.
"""
import itertools


def run_model(protocol, n_waves, skip, n_items):
    """Explores every interleaving; returns the set of outcomes, an outcome being
    ("ok", item sequence) when all waves saw one sequence, else ("diverged", ...)
    or ("stuck", ...)."""
    # wave state: (pc, idx, it, seen-tuple); shared: (counter, slot0, slot1)
    # barrier: the set of waves waiting at a barrier
    start = (tuple((0, -1, 0, ()) for _ in range(n_waves)), (0, -1, -1), frozenset())
    outcomes = set()
    seen = set()
    stack = [start]
    while stack:
        st = stack.pop()
        if st in seen:
            continue
        seen.add(st)
        waves, shared, waiting = st
        live = [w for w in range(n_waves) if waves[w][0] != "end"]
        if not live:
            seqs = {waves[w][3] for w in range(n_waves)}
            outcomes.add(("ok", seqs.pop()) if len(seqs) == 1 else ("diverged", tuple(sorted(seqs))))
            continue
        moves = []
        # a barrier releases when every live wave waits at one
        if live and all(w in waiting for w in live):
            moves.append(("release", None))
        for w in live:
            if w not in waiting:
                moves.append(("step", w))
        if not moves:
            outcomes.add(("stuck", tuple(waves)))
            continue
        for kind, w in moves:
            if kind == "release":
                nw = tuple((wv[0] + 1 if i in waiting else wv[0],) + wv[1:] if wv[0] != "end" else wv
                           for i, wv in enumerate(waves))
                stack.append((nw, shared, frozenset()))
            else:
                stack.extend(protocol(w, waves, shared, waiting, skip, n_items))
    return outcomes


def _set(waves, w, **kw):
    pc, idx, it, seqs = waves[w]
    new = dict(pc=pc, idx=idx, it=it, seqs=seqs)
    new.update(kw)
    out = list(waves)
    out[w] = (new["pc"], new["idx"], new["it"], new["seqs"])
    return tuple(out)


def double_slot(w, waves, shared, waiting, skip, n_items):
    """WorkQueue: first() / prefetch() / next() around an item that may be left
    early (skip) but always reaches next()'s barrier."""
    pc, idx, it, seqs = waves[w]
    ctr, s0, s1 = shared
    slots = [s0, s1]
    if pc == 0:  # first(): thread 0 takes item 0 into slot 0
        if w == 0:
            return [(_set(waves, w, pc=1), (ctr + 1, ctr, s1), waiting)]
        return [(_set(waves, w, pc=1), shared, waiting)]
    if pc == 1:  # barrier of first()
        return [(waves, shared, waiting | {w})]
    if pc == 2:  # read slot 0
        return [(_set(waves, w, pc=3, idx=slots[0], it=0), shared, waiting)]
    if pc == 3:  # loop test
        if idx >= n_items:
            return [(_set(waves, w, pc="end"), shared, waiting)]
        return [(_set(waves, w, pc=4, seqs=seqs + (idx,)), shared, waiting)]
    if pc == 4:  # prefetch(): thread 0 takes the next item into the other slot
        if w == 0:
            sl = list(slots)
            sl[(it + 1) & 1] = ctr
            return [(_set(waves, w, pc=5), (ctr + 1, sl[0], sl[1]), waiting)]
        return [(_set(waves, w, pc=5), shared, waiting)]
    if pc == 5:  # the item: left early, or a barrier inside it
        if skip[idx]:
            return [(_set(waves, w, pc=7), shared, waiting)]
        return [(_set(waves, w, pc=6), shared, waiting | {w})]
    if pc == 6:  # (a barrier release moves pc 6 -> 7: handled by the release)
        raise AssertionError
    if pc == 7:  # next(): the barrier that ends the item
        return [(waves, shared, waiting | {w})]
    if pc == 8:  # next(): read the other slot
        it2 = it + 1
        return [(_set(waves, w, pc=3, idx=slots[it2 & 1], it=it2), shared, waiting)]
    raise AssertionError(pc)


def one_slot_skipping(w, waves, shared, waiting, skip, n_items):
    """A one-slot form: thread 0 writes the slot at the top of the loop, a barrier,
    everyone reads it; an item left early goes straight back to the top (no
    barrier at the end of the item); other items end at a barrier."""
    pc, idx, it, seqs = waves[w]
    ctr, s0, s1 = shared
    if pc == 0:  # top: thread 0 takes the next item into the slot
        if w == 0:
            return [(_set(waves, w, pc=1), (ctr + 1, ctr, s1), waiting)]
        return [(_set(waves, w, pc=1), shared, waiting)]
    if pc == 1:
        return [(waves, shared, waiting | {w})]
    if pc == 2:  # read the slot
        return [(_set(waves, w, pc=3, idx=s0), shared, waiting)]
    if pc == 3:
        if idx >= n_items:
            return [(_set(waves, w, pc="end"), shared, waiting)]
        seqs = seqs + (idx,)
        if skip[idx]:
            return [(_set(waves, w, pc=0, seqs=seqs), shared, waiting)]  # `continue`
        return [(_set(waves, w, pc=4, seqs=seqs), shared, waiting)]
    if pc == 4:  # the barrier that ends a worked item
        return [(waves, shared, waiting | {w})]
    if pc == 5:
        return [(_set(waves, w, pc=0), shared, waiting)]
    raise AssertionError(pc)


def test_double_slot_protocol_is_safe():
    for n_waves in (2, 3):
        for n_items in (1, 2, 3):
            for skip in itertools.product((False, True), repeat=n_items):
                out = run_model(double_slot, n_waves, skip, n_items)
                assert out == {("ok", tuple(range(n_items)))}, (n_waves, skip, out)


def test_one_slot_form_with_early_exit_diverges():
    out = run_model(one_slot_skipping, 2, (True, False, False), 3)
    assert any(o[0] != "ok" for o in out)
    # without early exits the same form is safe
    out = run_model(one_slot_skipping, 2, (False, False, False), 3)
    assert out == {("ok", (0, 1, 2))}
