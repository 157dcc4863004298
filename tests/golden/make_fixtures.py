"""Writes the golden vectors of tests/golden/ from the reference's own pure-Python modules (blocksvibe/structure.py,
blocksvibe/recursion_ref.py, blocksvibe/plan_ref.py of the reference checkout).  Run in a checkout that has the
reference beside it; the vectors are committed, the reference is not needed to use them:

    python tests/golden/make_fixtures.py [/path/to/reference]

  structure.json  pole bookkeeping (pole_range_k, poles, position, shift, pole_two_j, delta_2, parity, q_value,
                  two_j_values, so3, j_excess, cauchy_coeff) over a grid of spins and orders
  plan.json       PlanRef (slot offsets, M and Q keys in first-appearance order, per-level groups, Cauchy
                  differences and tasks, physical slots with and without reuse) for small operators
  recursion.json  recursion_ref.recursion in exact Fraction arithmetic on seeded small operators: inputs
                  (M matrices, h_inf) and every residue
Rationals are written as "p/q" strings.
"""
import json
import os
import random
import sys
from fractions import Fraction

HERE = os.path.dirname(os.path.abspath(__file__))


def q(x):
    x = Fraction(x)
    return f"{x.numerator}/{x.denominator}"


def structure_vectors(S):
    out = {"range": [], "poles": [], "two_j_values": [], "so3": [], "cauchy": []}
    for tj in range(0, 13):
        for tj12 in range(tj % 2, 9, 2):
            for tj43 in range(tj % 2, 9, 2):
                for order in range(-3, 9):                  # negative orders: the recursion's shifted orders
                    out["range"].append([tj, tj12, tj43, order, [S.pole_range_k(t, tj, tj12, tj43, order)
                                                                 for t in range(4)]])
                    pl = S.poles(tj, tj12, tj43, order)
                    out["poles"].append([tj, tj12, tj43, order, [
                        [t, k, S.position(tj, tj12, tj43, order, (t, k)), S.shift((t, k)), S.pole_two_j((t, k), tj),
                         S.delta_2((t, k), tj), S.parity((t, k)), q(S.q_value((t, k), tj)),
                         S.j_excess(S.pole_two_j((t, k), tj), 6)] for t, k in pl],
                        [S.position(tj, tj12, tj43, order, (t, 1)) for t in range(4)]])
    for order in range(1, 9):
        for m in range(order):
            for tj12 in (0, 1, 2, 3):
                for tjmax in (tj12 % 2, 4 + tj12 % 2, 8 + tj12 % 2):
                    out["two_j_values"].append([m, order, tj12, tjmax, S.two_j_values(m, order, tj12, tjmax)])
    for pp in (0, 1):
        for a in range(0, 9):
            for b in range(a % 2, 13, 2):
                out["so3"].append([pp, a, b, list(S.so3(pp, a, b))])
    for dp in range(-12, 13):
        for dd in range(-12, 13):
            out["cauchy"].append([dp, dd, q(S.cauchy_coeff(dp, dd))])
    return out


PLAN_CASES = [(3, 1, 3, 5, ((0, 1), (1, 0))), (4, 2, 4, 6, ((0, 0), (1, 1))), (6, 2, 4, 7, ((0, 1), (1, 0))),
              (8, 8, 2, 6, ((0, 0), (1, 1))), (0, 0, 6, 8, ((0, 0), (1, 1))), (5, 5, 3, 6, ((1, 0), (0, 1)))]


def plan_vectors(PR):
    out = []
    for tj12, tj43, tjmax, order, pp in PLAN_CASES:
        for reuse in (False, True):
            p = PR.PlanRef(tj12, tj43, tjmax, order, pp, reuse=reuse)
            levels = []
            for lev in p.levels:
                if lev is None:
                    levels.append(None)
                    continue
                levels.append({k: (lev[k].tolist() if hasattr(lev[k], "tolist") else lev[k])
                               for k in ("m", "nb", "Mr", "Kc", "gsrc", "gj", "col", "cdiff", "t_slot", "t_hrow",
                                         "t_L", "t_R", "t_q")})
            out.append({"case": [tj12, tj43, tjmax, order, [list(x) for x in pp], reuse],
                        "Smax": p.Smax, "Jh": p.Jh, "nslots": p.nslots, "nphys": p.nphys, "zero_slot": p.zero_slot,
                        "dmax": p.dmax,
                        "Mkeys": [[list(k[0]), k[1], k[2], k[3], k[4]] for k, _ in sorted(p.Mkeys.items(),
                                                                                          key=lambda kv: kv[1])],
                        "Qkeys": [[list(k[0]), k[1], k[2]] for k, _ in sorted(p.Qkeys.items(), key=lambda kv: kv[1])],
                        "offset": [[m, j, v] for (m, j), v in sorted(p.offset.items())],
                        "phys": sorted([t, v] for t, v in p.phys.items()),
                        "levels": levels})
    return out


RECURSION_CASES = [(3, 1, 3, 5, ((0, 1), (1, 0)), 1), (4, 2, 4, 6, ((0, 0), (1, 1)), 2),
                   (4, 2, 4, 6, ((0, 1), (1, 0)), 3), (6, 2, 2, 6, ((0, 1), (1, 0)), 4),
                   (8, 8, 2, 5, ((0, 0), (1, 1)), 5), (2, 2, 6, 7, ((0, 0), (1, 1)), 6),
                   (0, 0, 4, 8, ((0, 0), (1, 1)), 7), (5, 3, 3, 5, ((1, 0), (0, 1)), 8)]


def recursion_vectors(S, RR, PR):
    out = []
    for tj12, tj43, tjmax, order, pp, seed in RECURSION_CASES:
        rng = random.Random(seed)
        val = lambda: Fraction(rng.randint(-64, 64), rng.choice((16, 32, 64)))
        plan = PR.PlanRef(tj12, tj43, tjmax, order, pp)
        M = {}
        for (P, d, tjNN, ppp, tj) in plan.Mkeys:
            rows = len(S.so3((ppp + S.parity(P)) % 2, tjNN, tj))
            cols = len(S.so3(ppp, tjNN, S.pole_two_j(P, tj)))
            M[(P, d, tjNN, ppp, tj)] = [[val() for _ in range(cols)] for _ in range(rows)]
        h = {}
        for s in range(2):
            for m in range(order):
                for j in range(plan.Jh):
                    tj = tj12 % 2 + 2 * j
                    for a in range(len(S.so3(pp[s][0], tj12, tj))):
                        for b in range(len(S.so3(pp[s][1], tj43, tj))):
                            h[(s, m, j, a, b)] = val()
        res = RR.recursion(h, M, tj12, tj43, tjmax, order, pp, Fraction)
        out.append({"case": [tj12, tj43, tjmax, order, [list(x) for x in pp]],
                    "M": [[list(k[0]), k[1], k[2], k[3], k[4], [[q(x) for x in row] for row in v]] for k, v in M.items()],
                    "h_inf": [list(k) + [q(v)] for k, v in h.items()],
                    "res": [list(k) + [q(v)] for k, v in sorted(res.items())]})
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    sys.path.insert(0, ref)
    from blocksvibe import plan_ref as PR
    from blocksvibe import recursion_ref as RR
    from blocksvibe import structure as S
    for name, source, data in (("structure.json", "structure.py", structure_vectors(S)),
                               ("plan.json", "plan_ref.py", plan_vectors(PR)),
                               ("recursion.json", "recursion_ref.py", recursion_vectors(S, RR, PR))):
        path = os.path.join(HERE, name)
        with open(path, "w") as f:
            json.dump({"source": f"reference blocksvibe/{source}", "data": data}, f, separators=(",", ":"))
        print(f"{path}: {os.path.getsize(path) / 1e6:.2f} MB", flush=True)


if __name__ == "__main__":
    main()
