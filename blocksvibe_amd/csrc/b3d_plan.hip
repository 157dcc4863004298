// libb3dhip: the index plan of the batched radial recursion, built on the host (reference
// blocksvibe/recursion_gpu.py:136-365, plan_ref.py:13-121; pole bookkeeping of structure.py:16-74).
//
// For one operator (two_j12, two_j43, two_j_max, order, pseudo-parities): the logical residue slots, per radial
// level the Cauchy groups (source parity, pole spin) with their rows D' and prime-pole columns, the pole-step
// tasks with their M, Q and (L, Q) table keys (numbered by first appearance over all levels), and the physical
// slots (recycled through a LIFO stack once their last reader has run).  Integers only.  The library's
// recursion kernels read these tables; blocksvibe_amd/plan.py is the host-side object over them.
//
// Spins are doubled.  A pole is (type t in 0..3 = blocks_3d's I..IV, k >= 1); every count below is clipped at 0.
#include <stdint.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/b3dhip.h"

namespace b3d {
void set_error(const std::string& msg);
}

namespace {

typedef int64_t i64;
typedef std::vector<i64> V;

i64 fdiv(i64 a, i64 b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }   // floor division, b > 0
i64 fmod4(i64 a) { return ((a % 4) + 4) % 4; }

// pole counts per type of spin tj at radial order `order` (reference structure.py:16-24)
void n_poles(i64 tj, i64 tj12, i64 tj43, i64 order, i64 (&c)[4]) {
  const i64 half = tj / 2, small = std::min(half, std::min(tj12 / 2, tj43 / 2));
  const bool even = tj % 2 == 0;
  c[0] = order;
  c[1] = std::min(half, order);
  c[2] = even ? fdiv(order, 2) : std::min(small, fdiv(order, 2));
  c[3] = even ? std::min(small, fdiv(order + 1, 2)) : fdiv(order + 1, 2);
  for (i64& x : c) x = std::max<i64>(x, 0);
}
i64 n_poles_sum(i64 tj, i64 tj12, i64 tj43, i64 order) {
  i64 c[4];
  n_poles(tj, tj12, tj43, order, c);
  return c[0] + c[1] + c[2] + c[3];
}
// (type, k) of the idx-th pole in type-major order
void pole_of_index(i64 tj, i64 tj12, i64 tj43, i64 order, i64 idx, i64& t, i64& k) {
  i64 c[4];
  n_poles(tj, tj12, tj43, order, c);
  i64 before = 0;
  t = 0;
  while (t < 3 && idx >= before + c[t]) before += c[t++];
  k = idx - before + 1;
}
i64 index_of_pole(i64 tj, i64 tj12, i64 tj43, i64 order, i64 t, i64 k) {
  i64 c[4];
  n_poles(tj, tj12, tj43, order, c);
  i64 before = 0;
  for (i64 u = 0; u < t; ++u) before += c[u];
  return before + k - 1;
}
i64 shift(i64 t, i64 k) { return t < 2 ? k : (t == 2 ? 2 * k : 2 * k - 1); }         // structure.py:37-39
i64 pole_spin(i64 t, i64 k, i64 tj) { return t == 0 ? tj + 2 * k : (t == 1 ? tj - 2 * k : tj); }   // :42-44
i64 pole_delta2(i64 t, i64 k, i64 tj) {                                             // :47-50
  const i64 v[4] = {2 - tj - 2 * k, 4 + tj - 2 * k, 3 - 2 * k, 4 - 2 * k};
  return v[t];
}
i64 spin_excess(i64 tj, i64 tjmax) { return tj > tjmax ? (tj - tjmax) / 2 : 0; }
// SO(3) structures of the given pseudo-parity (structure.py:71-74)
i64 so3(i64 pp, i64 tj1, i64 tj2) {
  i64 n = 0;
  for (i64 c = std::abs(tj1 - tj2); c <= tj1 + tj2; c += 2)
    if (fmod4((tj1 + tj2 - c) - 2 * pp) == 0) ++n;
  return n;
}

// rank of every code among the distinct codes in order of first appearance, and those codes (a direct table
// over the codes' range when it is small, else a hash map)
void first_rank(const V& codes, V& rank, V& uniq) {
  rank.resize(codes.size());
  uniq.clear();
  if (codes.empty()) return;
  const auto mm = std::minmax_element(codes.begin(), codes.end());
  const i64 lo = *mm.first, span = *mm.second - lo + 1;
  if (span <= std::max<i64>(4 * (i64)codes.size(), 4096)) {
    V at(span, -1);
    for (size_t i = 0; i < codes.size(); ++i) {
      i64& r = at[codes[i] - lo];
      if (r < 0) {
        r = (i64)uniq.size();
        uniq.push_back(codes[i]);
      }
      rank[i] = r;
    }
    return;
  }
  std::unordered_map<i64, i64> at;
  at.reserve(codes.size() * 2 + 1);
  for (size_t i = 0; i < codes.size(); ++i) {
    auto it = at.find(codes[i]);
    if (it == at.end()) {
      it = at.emplace(codes[i], (i64)uniq.size()).first;
      uniq.push_back(codes[i]);
    }
    rank[i] = it->second;
  }
}

struct Level {
  bool present = false;
  i64 m = 0, nb = 0, Kc = 1, nH = 0, T = 0, dmax = 0;
  V gsrc, gtjp, ncol, col, cold, rg, rd, t_slot, t_hrow, t_L, t_R, t_q, t_qL;
  V kt, kk, ktj, kpl, kpr, ksign;   // task keys (t, k, tj, ppL, ppR, sign)
};

struct Plan {
  i64 tj12, tj43, tjmax, order, b, pp[2][2];
  i64 SLm = 0, SRm = 0, S2 = 0, Smax = 0, Jh = 0, nslots = 0, nphys = 0, dmax = 0;
  V SL[2], SR[2];
  std::vector<V> ltjs, lnpol, lfirst;
  V lstart, npol_flat, first_flat, phys;
  std::vector<Level> levels;
  V mkeys, qkeys, lqpairs;   // flattened (t, k, d, tjNN, ppp, tj), (t, k, tj, sign), (L, Q)

  i64 ji(i64 tj) const { return (tj - b) / 2; }
  i64 sl(i64 p, i64 tj) const { return SL[p][ji(tj)]; }
  i64 sr(i64 p, i64 tj) const { return SR[p][ji(tj)]; }

  bool build(bool reuse) {
    b = tj12 % 2;
    const i64 ntj = (tjmax + 2 * order + 2 - b) / 2 + 1;   // two_j = b, b + 2, .., <= tjmax + 2 order + 2
    for (int p = 0; p < 2; ++p) {
      SL[p].resize(ntj);
      SR[p].resize(ntj);
      for (i64 i = 0; i < ntj; ++i) {
        SL[p][i] = so3(pp[p][0], tj12, b + 2 * i);
        SR[p][i] = so3(pp[p][1], tj43, b + 2 * i);
      }
    }
    const i64 nsp = (tjmax + 2 * order - b) / 2 + 1;
    for (int p = 0; p < 2; ++p)
      for (i64 i = 0; i < nsp && i < ntj; ++i) {
        SLm = std::max(SLm, SL[p][i]);
        SRm = std::max(SRm, SR[p][i]);
      }
    S2 = SLm * SRm;
    Smax = std::max(SLm, SRm);
    Jh = (tjmax + 2 * order - b) / 2 + 1;
    // logical slot layout: level m, spin jidx, pole
    i64 n = 0;
    for (i64 m = 0; m < order; ++m) {
      V tjs, npol, first;
      for (i64 tj = tj12 % 2; tj <= tjmax + 2 * (order - m - 1); tj += 2) tjs.push_back(tj);
      for (i64 tj : tjs) {
        first.push_back(n);
        const i64 c = n_poles_sum(tj, tj12, tj43, order - m - spin_excess(tj, tjmax));
        npol.push_back(c);
        n += c;
      }
      lstart.push_back((i64)npol_flat.size());
      npol_flat.insert(npol_flat.end(), npol.begin(), npol.end());
      first_flat.insert(first_flat.end(), first.begin(), first.end());
      ltjs.push_back(tjs);
      lnpol.push_back(npol);
      lfirst.push_back(first);
    }
    nslots = n;
    levels.resize(order);
    for (i64 m = 0; m < order; ++m)
      if (!level(m, levels[m])) return false;
    keys();
    physical(reuse);
    return true;
  }

  bool level(i64 m, Level& lev) {
    const V& tjs = ltjs[m];
    const V& npol = lnpol[m];
    const i64 ns = nslots;
    // tasks: every (spin, pole) of the level, parity p innermost, kept when both its block and its source
    // block hold structures
    V jidx, pidx, tj, t, k, p, tjp, dprime, src;
    for (size_t j = 0; j < tjs.size(); ++j)
      for (i64 q = 0; q < npol[j]; ++q) {
        i64 tt, kk;
        pole_of_index(tjs[j], tj12, tj43, order - m - spin_excess(tjs[j], tjmax), q, tt, kk);
        const i64 tp = pole_spin(tt, kk, tjs[j]);
        for (i64 par = 0; par < 2; ++par) {
          const i64 s = (par + (tt == 3)) % 2;
          if (ji(tp) >= (i64)SL[0].size()) {
            b3d::set_error("b3d_plan_build: a pole's spin lies beyond the structure tables");
            return false;
          }
          if (!(sl(par, tjs[j]) > 0 && sr(par, tjs[j]) > 0 && sl(s, tp) > 0 && sr(s, tp) > 0)) continue;
          jidx.push_back((i64)j);
          pidx.push_back(q);
          tj.push_back(tjs[j]);
          t.push_back(tt);
          k.push_back(kk);
          p.push_back(par);
          tjp.push_back(tp);
          dprime.push_back(pole_delta2(tt, kk, tjs[j]) + 2 * shift(tt, kk));
          src.push_back(s);
        }
      }
    if (p.empty()) return true;   // nothing at this level
    lev.present = true;
    lev.m = m;
    const size_t T = p.size();
    // groups (src, two_j') and their rows D', each in order of first appearance
    // (codes compact over their ranges: group = src TJ + two_j', row = group DS + D' - min D')
    const i64 TJ = *std::max_element(tjp.begin(), tjp.end()) + 1;
    V codes(T), gi, gcodes;
    for (size_t i = 0; i < T; ++i) codes[i] = src[i] * TJ + tjp[i];
    first_rank(codes, gi, gcodes);
    const i64 nb = (i64)gcodes.size();
    lev.nb = nb;
    lev.gsrc.resize(nb);
    lev.gtjp.resize(nb);
    for (i64 g = 0; g < nb; ++g) {
      lev.gsrc[g] = gcodes[g] / TJ;
      lev.gtjp[g] = gcodes[g] % TJ;
    }
    const auto dmm = std::minmax_element(dprime.begin(), dprime.end());
    const i64 dmin = *dmm.first, DS = *dmm.second - dmin + 1;
    V ri, rcodes;
    for (size_t i = 0; i < T; ++i) codes[i] = gi[i] * DS + (dprime[i] - dmin);
    first_rank(codes, ri, rcodes);
    const i64 nrow = (i64)rcodes.size();
    // rows grouped, first appearance within each group (a counting sort by group)
    V gof(nrow), start(nb + 1, 0), by_g(nrow);
    for (i64 r = 0; r < nrow; ++r) ++start[(gof[r] = rcodes[r] / DS) + 1];
    for (i64 g = 0; g < nb; ++g) start[g + 1] += start[g];
    for (i64 r = 0; r < nrow; ++r) by_g[start[gof[r]]++] = r;
    V hrow_of(nrow);
    lev.nH = nrow;
    lev.rg.resize(nrow);
    lev.rd.resize(nrow);
    for (i64 h = 0; h < nrow; ++h) {
      hrow_of[by_g[h]] = h;
      lev.rg[h] = rcodes[by_g[h]] / DS;
      lev.rd[h] = rcodes[by_g[h]] % DS + dmin;
    }
    // columns: the prime poles of two_j' at order m, residues of level m - shift
    lev.ncol.resize(nb);
    i64 Kc = 1;
    for (i64 g = 0; g < nb; ++g) {
      lev.ncol[g] = n_poles_sum(lev.gtjp[g], tj12, tj43, m);
      Kc = std::max(Kc, lev.ncol[g]);
    }
    lev.Kc = Kc;
    lev.col.assign(nb * Kc, -1);
    lev.cold.assign(nb * Kc, 0);
    for (i64 g = 0; g < nb; ++g) {
      const i64 ctj = lev.gtjp[g];
      for (i64 c = 0; c < lev.ncol[g]; ++c) {
        i64 ti, ki;
        pole_of_index(ctj, tj12, tj43, m, c, ti, ki);
        const i64 si = shift(ti, ki), lvl = m - si;
        const i64 cord = order - m + si - spin_excess(ctj, tjmax);
        const i64 cpos = index_of_pole(ctj, tj12, tj43, cord, ti, ki);
        const i64 cji = ji(ctj);
        if (lvl < 0 || cji >= (i64)ltjs[lvl].size()) {
          b3d::set_error("b3d_plan_build: prime-pole spin outside its level");
          return false;
        }
        const i64 fi = lstart[lvl] + cji;
        if (cpos >= npol_flat[fi]) {
          b3d::set_error("b3d_plan_build: prime-pole position outside its residue block");
          return false;
        }
        lev.col[g * Kc + c] = lev.gsrc[g] * ns + first_flat[fi] + cpos;
        lev.cold[g * Kc + c] = pole_delta2(ti, ki, ctj);
      }
    }
    // largest |D' - D_i| of the level (the Cauchy kernel divides by it with one-limb arithmetic): per group
    // from the extremes of its rows and of its columns
    const i64 BIG = i64(1) << 62;
    V rlo(nb, BIG), rhi(nb, -BIG);
    for (i64 h = 0; h < nrow; ++h) {
      rlo[lev.rg[h]] = std::min(rlo[lev.rg[h]], lev.rd[h]);
      rhi[lev.rg[h]] = std::max(rhi[lev.rg[h]], lev.rd[h]);
    }
    i64 dm = 0;
    for (i64 g = 0; g < nb; ++g) {
      if (lev.ncol[g] == 0 || rlo[g] > rhi[g]) continue;
      const auto cm = std::minmax_element(lev.cold.begin() + g * Kc, lev.cold.begin() + g * Kc + lev.ncol[g]);
      dm = std::max({dm, std::abs(rhi[g] - *cm.first), std::abs(rlo[g] - *cm.second)});
    }
    lev.dmax = dm;
    // tasks
    lev.T = (i64)T;
    lev.t_slot.resize(T);
    lev.t_hrow.resize(T);
    for (size_t i = 0; i < T; ++i) {
      lev.t_slot[i] = p[i] * ns + lfirst[m][jidx[i]] + pidx[i];
      lev.t_hrow[i] = hrow_of[ri[i]];
      lev.kt.push_back(t[i]);
      lev.kk.push_back(k[i]);
      lev.ktj.push_back(tj[i]);
      lev.kpl.push_back(pp[src[i]][0]);
      lev.kpr.push_back(pp[src[i]][1]);
      lev.ksign.push_back((tjp[i] - tj[i]) % 4 != 0 ? -1 : 1);
    }
    return true;
  }

  // M keys (L before R per task), Q keys and (L, Q) pairs, by first appearance over all levels
  void keys() {
    const i64 KM = order + 2, TJM = tjmax + 2 * order + 4, PPM = 2;
    V mc, qc;
    for (const Level& lev : levels) {
      if (!lev.present) continue;
      for (i64 i = 0; i < lev.T; ++i) {
        const i64 tk = lev.kt[i] * KM + lev.kk[i];
        mc.push_back(((tk * 2 + 0) * PPM + lev.kpl[i]) * TJM + lev.ktj[i]);
        mc.push_back(((tk * 2 + 1) * PPM + lev.kpr[i]) * TJM + lev.ktj[i]);
        qc.push_back((tk * TJM + lev.ktj[i]) * 2 + (lev.ksign[i] > 0));
      }
    }
    V midx, mcodes, qidx, qcodes, lq, lqcodes;
    first_rank(mc, midx, mcodes);
    first_rank(qc, qidx, qcodes);
    for (i64 c : mcodes) {
      const i64 tj = c % TJM, u1 = c / TJM, ppp = u1 % PPM, u2 = u1 / PPM, d = u2 % 2, u3 = u2 / 2;
      const i64 kk = u3 % KM, tt = u3 / KM;
      const i64 v[6] = {tt, kk, d, d == 0 ? tj12 : tj43, ppp, tj};
      mkeys.insert(mkeys.end(), v, v + 6);
    }
    for (i64 c : qcodes) {
      const i64 s = c % 2, u = c / 2, tj = u % TJM, x = u / TJM;
      const i64 v[4] = {x / KM, x % KM, tj, s ? 1 : -1};
      qkeys.insert(qkeys.end(), v, v + 4);
    }
    const i64 nQ = (i64)qcodes.size();
    V lc(qidx.size());
    for (size_t i = 0; i < qidx.size(); ++i) lc[i] = midx[2 * i] * nQ + qidx[i];
    first_rank(lc, lq, lqcodes);
    for (i64 c : lqcodes) {
      lqpairs.push_back(c / nQ);
      lqpairs.push_back(c % nQ);
    }
    size_t o = 0;
    for (Level& lev : levels) {
      if (!lev.present) continue;
      for (i64 i = 0; i < lev.T; ++i, ++o) {
        lev.t_L.push_back(midx[2 * o]);
        lev.t_R.push_back(midx[2 * o + 1]);
        lev.t_q.push_back(qidx[o]);
        lev.t_qL.push_back(lq[o]);
      }
      dmax = std::max(dmax, lev.dmax);
    }
  }

  // physical slots: without reuse one per logical slot that is written; with reuse a slot goes back to a LIFO
  // stack once the level after its last reader's has come (the level after its own if nothing reads it), and
  // the slots born at a level (in logical order) take the most recently returned ones first, then new numbers.
  // Slots due back at the same level are handed back in the order that level was first scheduled, so the
  // numbering (and the published slot counts) is that of the reference's plan.
  void physical(bool reuse) {
    const i64 N = 2 * nslots;
    V written(N, -1), read_to(N, -1);
    for (const Level& lev : levels) {
      if (!lev.present) continue;
      for (i64 s : lev.t_slot) written[s] = lev.m;
      for (i64 c : lev.col)
        if (c >= 0) read_to[c] = std::max(read_to[c], lev.m);
    }
    phys.assign(N, -1);
    if (!reuse) {
      for (i64 s = 0; s < N; ++s)
        if (written[s] >= 0) phys[s] = s;
      nphys = N;
    } else {
      V stack;
      std::vector<std::pair<i64, std::vector<V>>> returns;   // return level -> slot lists, scheduling order
      i64 made = 0;
      for (const Level& lev : levels) {
        if (!lev.present) continue;
        const i64 here = lev.m;
        for (size_t r = 0; r < returns.size();) {
          if (returns[r].first <= here) {
            for (const V& a : returns[r].second) stack.insert(stack.end(), a.begin(), a.end());
            returns.erase(returns.begin() + r);
          } else {
            ++r;
          }
        }
        V born = lev.t_slot;
        std::sort(born.begin(), born.end());
        const size_t k = std::min(born.size(), stack.size());
        V got(born.size());
        for (size_t i = 0; i < k; ++i) got[i] = stack[stack.size() - 1 - i];
        for (size_t i = k; i < born.size(); ++i) got[i] = made++;
        stack.resize(stack.size() - k);
        std::vector<i64> order_seen;
        std::unordered_map<i64, V> byr;
        for (size_t i = 0; i < born.size(); ++i) {
          phys[born[i]] = got[i];
          const i64 back = std::max(read_to[born[i]], here + 1);
          auto it = byr.find(back);
          if (it == byr.end()) {
            order_seen.push_back(back);
            it = byr.emplace(back, V()).first;
          }
          it->second.push_back(got[i]);
        }
        for (i64 r : order_seen) {
          auto it = std::find_if(returns.begin(), returns.end(), [&](const auto& e) { return e.first == r; });
          if (it == returns.end()) {
            returns.emplace_back(r, std::vector<V>());
            it = returns.end() - 1;
          }
          it->second.push_back(byr[r]);
        }
      }
      nphys = made;
    }
    for (Level& lev : levels) {
      if (!lev.present) continue;
      for (i64& s : lev.t_slot) s = phys[s];
      for (i64& c : lev.col) {
        const i64 v = c >= 0 ? phys[c] : -1;
        c = v >= 0 ? v : nphys;
      }
    }
  }
};

i64 put(const V& v, i64* out, i64 cap) {
  if (out) std::memcpy(out, v.data(), sizeof(i64) * (size_t)std::min<i64>(cap, (i64)v.size()));
  return (i64)v.size();
}

}  // namespace

extern "C" {

int b3d_plan_build(int two_j12, int two_j43, int two_j_max, int order, int pp0_left, int pp0_right, int pp1_left,
                   int pp1_right, int reuse, void** plan) {
  if (!plan) {
    b3d::set_error("b3d_plan_build: null plan pointer");
    return 2;
  }
  *plan = nullptr;
  if (two_j12 < 0 || two_j43 < 0 || two_j_max < 0 || order < 0 || order > 4096 || two_j_max > 1 << 16 ||
      two_j12 % 2 != two_j43 % 2) {
    b3d::set_error("b3d_plan_build: need two_j12, two_j43 >= 0 of one parity, 0 <= two_j_max, 0 <= order <= 4096");
    return 2;
  }
  const int pps[4] = {pp0_left, pp0_right, pp1_left, pp1_right};
  for (int x : pps)
    if (x != 0 && x != 1) {
      b3d::set_error("b3d_plan_build: pseudo-parities are 0 or 1");
      return 2;
    }
  std::unique_ptr<Plan> p(new Plan());
  p->tj12 = two_j12;
  p->tj43 = two_j43;
  p->tjmax = two_j_max;
  p->order = order;
  p->pp[0][0] = pp0_left;
  p->pp[0][1] = pp0_right;
  p->pp[1][0] = pp1_left;
  p->pp[1][1] = pp1_right;
  if (!p->build(reuse != 0)) return 1;
  *plan = p.release();
  return 0;
}

void b3d_plan_free(void* plan) { delete static_cast<Plan*>(plan); }

int64_t b3d_plan_get(const void* plan, int field, int level, int64_t* out, int64_t cap) {
  const Plan* p = static_cast<const Plan*>(plan);
  if (!p) return -1;
  if (field < B3D_PLAN_LEVEL) {
    switch (field) {
      case B3D_PLAN_SCALARS:
        return put({p->nslots, p->nphys, p->SLm, p->SRm, p->S2, p->Smax, p->Jh, p->dmax, p->order, p->b}, out, cap);
      case B3D_PLAN_SL: {
        V v(p->SL[0]);
        v.insert(v.end(), p->SL[1].begin(), p->SL[1].end());
        return put(v, out, cap);
      }
      case B3D_PLAN_SR: {
        V v(p->SR[0]);
        v.insert(v.end(), p->SR[1].begin(), p->SR[1].end());
        return put(v, out, cap);
      }
      case B3D_PLAN_PHYS: return put(p->phys, out, cap);
      case B3D_PLAN_MKEYS: return put(p->mkeys, out, cap);
      case B3D_PLAN_QKEYS: return put(p->qkeys, out, cap);
      case B3D_PLAN_LQPAIRS: return put(p->lqpairs, out, cap);
      case B3D_PLAN_LEVEL_TJS:
      case B3D_PLAN_LEVEL_NPOL:
      case B3D_PLAN_LEVEL_FIRST:
        if (level < 0 || level >= (int)p->ltjs.size()) return -1;
        return put(field == B3D_PLAN_LEVEL_TJS ? p->ltjs[level]
                                               : field == B3D_PLAN_LEVEL_NPOL ? p->lnpol[level] : p->lfirst[level],
                   out, cap);
      default: return -1;
    }
  }
  if (level < 0 || level >= (int)p->levels.size()) return -1;
  const Level& l = p->levels[level];
  switch (field) {
    case B3D_PLAN_LEVEL: return put({l.present ? 1 : 0, l.m, l.nb, l.Kc, l.nH, l.T, l.dmax}, out, cap);
    case B3D_PLAN_GSRC: return put(l.gsrc, out, cap);
    case B3D_PLAN_GTJP: return put(l.gtjp, out, cap);
    case B3D_PLAN_NCOL: return put(l.ncol, out, cap);
    case B3D_PLAN_COL: return put(l.col, out, cap);
    case B3D_PLAN_COLD: return put(l.cold, out, cap);
    case B3D_PLAN_RG: return put(l.rg, out, cap);
    case B3D_PLAN_RD: return put(l.rd, out, cap);
    case B3D_PLAN_T_SLOT: return put(l.t_slot, out, cap);
    case B3D_PLAN_T_HROW: return put(l.t_hrow, out, cap);
    case B3D_PLAN_T_L: return put(l.t_L, out, cap);
    case B3D_PLAN_T_R: return put(l.t_R, out, cap);
    case B3D_PLAN_T_Q: return put(l.t_q, out, cap);
    case B3D_PLAN_T_QL: return put(l.t_qL, out, cap);
    case B3D_PLAN_T_KEYS: {
      V v;
      for (const V* a : {&l.kt, &l.kk, &l.ktj, &l.kpl, &l.kpr, &l.ksign}) v.insert(v.end(), a->begin(), a->end());
      return put(v, out, cap);
    }
    default: return -1;
  }
}

}  // extern "C"
