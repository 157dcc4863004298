"""Per-task index plan of the batched recursion.  TEST INFRASTRUCTURE (oracle/__init__.py).

Restates reference blocksvibe/plan_ref.py:9-132 with plain dicts and loops (the semantics the vectorised
product Plan must reproduce): slot offsets, per-level (src, two_j') groups with their rows D' and prime-
pole columns, tasks (slot, group, row, M keys, Q key), and the physical slot assignment with LIFO reuse.
Pinned by tests/test_golden.py against the reference's own PlanRef on six operators with and without reuse
(tests/golden/plan.json: keys, offsets, groups, Cauchy differences, tasks, physical slots), by
tests/test_oracle_structure.py against the published slot counts (reference docs/BLOCKS_PLAN.md:209-210) and by
the invariants every slot is written before it is read and no physical slot is reused while live.
"""
from . import structure_ref as S


class PlanRef:
    def __init__(self, two_j12, two_j43, two_j_max, order, pp, reuse=False):
        self.tj12, self.tj43, self.tjmax, self.order, self.pp = two_j12, two_j43, two_j_max, order, pp
        SL = lambda p, tj: len(S.so3(pp[p][0], two_j12, tj))
        SR = lambda p, tj: len(S.so3(pp[p][1], two_j43, tj))
        self.offset, n = {}, 0
        for m in range(order):
            for jidx, tj in enumerate(S.two_j_values(m, order, two_j12, two_j_max)):
                self.offset[(m, jidx)] = n
                n += len(S.poles(tj, two_j12, two_j43, order - m - S.j_excess(tj, two_j_max)))
        self.nslots = n
        self.Mkeys, self.Qkeys = {}, {}
        self.levels = []
        for m in range(order):
            groups, tasks = {}, []
            for jidx, tj in enumerate(S.two_j_values(m, order, two_j12, two_j_max)):
                for pidx, P in enumerate(S.poles(tj, two_j12, two_j43, order - m - S.j_excess(tj, two_j_max))):
                    tjp = S.pole_two_j(P, tj)
                    dprime = S.delta_2(P, tj) + 2 * S.shift(P)
                    for p in range(2):
                        src = (p + S.parity(P)) % 2
                        if not (SL(p, tj) and SR(p, tj) and SL(src, tjp) and SR(src, tjp)):
                            continue
                        g = groups.get((src, tjp))
                        if g is None:
                            jpex = S.j_excess(tjp, two_j_max)
                            cols = [(src * n + self.offset[(m - S.shift(i), tjp // 2)]
                                     + S.position(tjp, two_j12, two_j43, order - m + S.shift(i) - jpex, i),
                                     S.delta_2(i, tjp)) for i in S.poles(tjp, two_j12, two_j43, m)]
                            g = groups[(src, tjp)] = {"rows": {}, "cols": cols, "gi": len(groups)}
                        r = g["rows"].setdefault(dprime, len(g["rows"]))
                        sign = -1 if (tjp - tj) % 4 else 1
                        Lk = self.Mkeys.setdefault((P, 0, two_j12, pp[src][0], tj), len(self.Mkeys))
                        Rk = self.Mkeys.setdefault((P, 1, two_j43, pp[src][1], tj), len(self.Mkeys))
                        Qk = self.Qkeys.setdefault((P, tj, sign), len(self.Qkeys))
                        tasks.append({"slot": p * n + self.offset[(m, jidx)] + pidx, "group": g["gi"], "row": r,
                                      "L": Lk, "R": Rk, "Q": Qk})
            self.levels.append(None if not groups else {"m": m, "groups": groups, "tasks": tasks})
        self._physical(reuse)

    def _physical(self, reuse):
        born, last = {}, {}
        for lev in self.levels:
            if lev is None:
                continue
            m = lev["m"]
            for t in lev["tasks"]:
                born[t["slot"]] = m
            for g in lev["groups"].values():
                for c, _ in g["cols"]:
                    last[c] = max(last.get(c, -1), m)
        if not reuse:
            self.phys, self.nphys = {t: t for t in born}, 2 * self.nslots
            return
        self.phys, free, release, nxt = {}, [], {}, 0
        for t in sorted(born, key=lambda t: (born[t], t)):
            m = born[t]
            for mm in [k for k in release if k <= m]:
                free += release.pop(mm)
            if free:
                ph = free.pop()
            else:
                ph, nxt = nxt, nxt + 1
            self.phys[t] = ph
            d = last.get(t, -1)
            release.setdefault(d if d > m else m + 1, []).append(ph)
        self.nphys = nxt
        self.born, self.last = born, last

    def phys_slot(self, logical):
        return self.phys.get(logical, self.nphys)
