"""A model, on the CPU, of the FOF root pass (k_fof_root_min in
skid_amd/csrc/skg_fof.hip): every thread i finds the root of point i in the
union-find parents and stores it into par[i]; the label pass then reads
cmin[par[i]], which is set for roots only.  Every interleaving of the memory
operations of two threads on a chain of seven nodes is run.  With a
path-halving find (which writes par[x] = par[par[x]] on its way), some
interleavings leave a non-root in par; with the read-only walk the pass uses,
none do."""
import itertools


def _thread(par, i, halving):
    """Thread i's loads and stores, one per step (yield after each)."""
    x = i
    while True:
        p = par[x]
        yield
        if p == x:
            break
        if halving:
            gp = par[p]
            yield
            if gp != p:
                par[x] = gp
                yield
            x = gp
        else:
            x = p
    par[i] = x
    yield


def _bad_schedules(halving):
    bad = total = 0
    for length in range(14):
        for sched in itertools.product((4, 6), repeat=length):
            par = [0, 0, 1, 2, 3, 4, 5]                    # the chain 6 -> 5 -> ... -> 0
            th = {t: _thread(par, t, halving) for t in (4, 6)}
            done, ok = set(), True
            for t in sched:
                if t in done:
                    ok = False
                    break
                try:
                    next(th[t])
                except StopIteration:
                    done.add(t)
            if not ok:
                continue
            for t in (4, 6):
                for _ in th[t]:
                    pass
            total += 1
            bad += par[4] != 0 or par[6] != 0
    return bad, total


def test_read_only_root_walk_leaves_roots():
    bad, total = _bad_schedules(halving=False)
    assert total > 10000 and bad == 0


def test_halving_root_walk_can_leave_a_non_root():
    bad, total = _bad_schedules(halving=True)
    assert total > 10000 and bad > 0
