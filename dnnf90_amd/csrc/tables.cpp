// tables.cpp -- host build of the multi-index tables and the GPU schedule.
//
// Reference semantics: lib/multi_index_bell_module.f90:128-351.  The
// construction here differs from the reference's O(|A|^2) scan: each
// carried a enumerates its own sub-multi-indices b <= a (a mixed-radix
// counter over the nonzero components) and locates b and a-b by a hashed
// key, which is O(sum_a prod_v (a_v+1)) and keeps force-field sized sets
// (D0 = 100, K = 2, |A| = 5151) to milliseconds.  Lists are then sorted
// into the reference's order, so every index and every coefficient matches.
#include "tables.hpp"

#include <algorithm>
#include <unordered_map>

namespace dnnf90 {

namespace {

std::string key_of(const int *a, int d0)
{
    return std::string(reinterpret_cast<const char *>(a), sizeof(int) * (size_t)d0);
}

// degree-p multi-indices in combinations-with-replacement order: choose
// p variable positions v_1 <= ... <= v_p, alpha_i = #{v_s == i}
void enum_degree(int d0, int p, std::vector<int> &out)
{
    std::vector<int> pos(p, 0);
    std::vector<int> a(d0);
    while (true) {
        std::fill(a.begin(), a.end(), 0);
        for (int s = 0; s < p; ++s) a[pos[s]]++;
        out.insert(out.end(), a.begin(), a.end());
        int s = p - 1;
        while (s >= 0 && pos[s] == d0 - 1) --s;
        if (s < 0) break;
        int v = pos[s] + 1;
        for (int r = s; r < p; ++r) pos[r] = v;
    }
}

}  // namespace

bool TableSet::build(int d0_, int k_, int nseed, const int *seeds, std::string *err)
{
    *this = TableSet();
    auto fail = [&](const char *m) {
        if (err) *err = m;
        *this = TableSet();
        return false;
    };
    if (d0_ < 1 || k_ < 1 || k_ > 15 || nseed < 0) return fail("d0 >= 1, 1 <= k <= 15, nseed >= 0");
    for (int s = 0; s < nseed; ++s) {
        int tot = 0;
        for (int v = 0; v < d0_; ++v) {
            int c = seeds[s * d0_ + v];
            if (c < 0 || c > k_) return fail("seed component outside [0,k]");
            tot += c;
        }
        if (tot > k_) return fail("seed multi-index exceeds k");
    }
    d0 = d0_;
    k = k_;

    // graded enumeration, then the downward-closure filter
    std::vector<int> full;
    full.insert(full.end(), (size_t)d0, 0);
    for (int p = 1; p <= k; ++p) enum_degree(d0, p, full);
    const size_t nfull = full.size() / d0;
    for (size_t i = 0; i < nfull; ++i) {
        const int *a = &full[i * d0];
        bool keep = nseed == 0;
        for (int s = 0; s < nseed && !keep; ++s) {
            bool le = true;
            for (int v = 0; v < d0; ++v)
                if (a[v] > seeds[s * d0 + v]) { le = false; break; }
            keep = le;
        }
        if (!keep) continue;
        int dg = 0;
        for (int v = 0; v < d0; ++v) dg += a[v];
        alpha.insert(alpha.end(), a, a + d0);
        deg.push_back(dg);
    }
    na = (int)deg.size();

    std::unordered_map<std::string, int> index;
    index.reserve((size_t)na * 2);
    for (int ia = 0; ia < na; ++ia) index.emplace(key_of(&alpha[(size_t)ia * d0], d0), ia);

    ind_e1.assign(d0, -1);
    for (int ia = 0; ia < na; ++ia)
        if (deg[ia] == 1)
            for (int v = 0; v < d0; ++v)
                if (alpha[(size_t)ia * d0 + v] == 1) ind_e1[v] = ia;

    double fact[17];
    fact[0] = 1.0;
    for (int i = 1; i <= 16; ++i) fact[i] = fact[i - 1] * (double)i;

    // sub-multi-index relation: for each a, every carried b <= a with b != 0
    struct Sub { int ib, id; double binom; };
    std::vector<std::vector<Sub>> subs(na);
    std::vector<int> b(d0), dvec(d0), nz;
    for (int ia = 1; ia < na; ++ia) {
        const int *a = &alpha[(size_t)ia * d0];
        nz.clear();
        for (int v = 0; v < d0; ++v)
            if (a[v] > 0) nz.push_back(v);
        std::fill(b.begin(), b.end(), 0);
        while (true) {
            // advance the mixed-radix counter over the nonzero components
            size_t r = 0;
            while (r < nz.size() && b[nz[r]] == a[nz[r]]) { b[nz[r]] = 0; ++r; }
            if (r == nz.size()) break;
            b[nz[r]]++;
            for (int v = 0; v < d0; ++v) dvec[v] = a[v] - b[v];
            Sub s;
            s.ib = index.at(key_of(b.data(), d0));
            s.id = index.at(key_of(dvec.data(), d0));
            double rr = 1.0;
            for (int v = 0; v < d0; ++v) rr = rr * fact[a[v]] / (fact[b[v]] * fact[a[v] - b[v]]);
            s.binom = rr;
            subs[ia].push_back(s);
        }
        std::sort(subs[ia].begin(), subs[ia].end(), [](const Sub &x, const Sub &y) { return x.ib < y.ib; });
    }

    // adjoint pair lists: for b = 1..na-1, every a >= b in ascending a
    pair_start.assign(na, 0);
    pair_num.assign(na, 0);
    {
        std::vector<std::vector<std::pair<int, Sub>>> byb(na);
        for (int ia = 1; ia < na; ++ia)
            for (const Sub &s : subs[ia]) byb[s.ib].push_back({ia, s});
        for (int ib = 1; ib < na; ++ib) {
            pair_start[ib] = (int)pair_ia.size();
            for (auto &e : byb[ib]) {
                pair_ia.push_back(e.first);
                pair_id.push_back(e.second.id);
                pair_binom.push_back(e.second.binom);
            }
            pair_num[ib] = (int)byb[ib].size();
        }
    }

    // forward lists: for a, q = 2..|a|, b ascending with |a|-|b| >= q-1
    fq_start.assign((size_t)na * (k + 1), 0);
    fq_num.assign((size_t)na * (k + 1), 0);
    for (int ia = 1; ia < na; ++ia) {
        const int p = deg[ia];
        for (int q = 2; q <= p; ++q) {
            fq_start[(size_t)ia * (k + 1) + q] = (int)fq_ib.size();
            for (const Sub &s : subs[ia]) {
                if (p - deg[s.ib] < q - 1) continue;
                fq_ib.push_back(s.ib);
                fq_id.push_back(s.id);
                fq_c.push_back(s.binom / (double)q);
            }
            fq_num[(size_t)ia * (k + 1) + q] = (int)fq_ib.size() - fq_start[(size_t)ia * (k + 1) + q];
        }
    }

    // tanh polynomials up to q = k+1 (degree q+1)
    const int ni = k + 3;
    tpc.assign((size_t)(k + 2) * ni, 0.0);
    tpc[1] = 1.0;  // P_0 = t
    std::vector<double> dc(ni);
    for (int q = 0; q < k + 1; ++q) {
        for (int i = 0; i < ni; ++i) dc[i] = (i + 1 < ni) ? (double)(i + 1) * tpc[(size_t)q * ni + i + 1] : 0.0;
        for (int i = 0; i < ni; ++i) {
            double v = dc[i];
            if (i >= 2) v -= dc[i - 2];
            tpc[(size_t)(q + 1) * ni + i] = v;
        }
    }
    return true;
}

void FwdSchedule::build(const TableSet &t)
{
    const int na = t.na, kq = t.k + 1;
    // which (a, q >= 2) does a later term read as its B(a-b, q-1) operand?
    std::vector<int> need((size_t)na * kq, 0);
    for (int ia = 1; ia < na; ++ia)
        for (int q = 3; q <= t.deg[ia]; ++q) {
            int st = t.fq_start[(size_t)ia * kq + q], nm = t.fq_num[(size_t)ia * kq + q];
            for (int it = st; it < st + nm; ++it) need[(size_t)t.fq_id[it] * kq + (q - 1)] = 1;
        }
    std::vector<int> entry((size_t)na * kq, -1);
    nentry = na;
    for (int ia = 0; ia < na; ++ia) {
        entry[(size_t)ia * kq + 1] = ia;
        for (int q = 2; q < kq; ++q)
            if (need[(size_t)ia * kq + q]) entry[(size_t)ia * kq + q] = nentry++;
    }
    slot_group.assign(na + 1, 0);
    grp_term.clear(); grp_store.clear(); grp_q.clear();
    term_s.clear(); term_b.clear(); term_c.clear();
    for (int ia = 0; ia < na; ++ia) {
        slot_group[ia] = (int)grp_store.size();
        for (int q = 2; q <= t.deg[ia] && ia > 0; ++q) {
            grp_term.push_back((int)term_s.size());
            grp_store.push_back(entry[(size_t)ia * kq + q]);
            grp_q.push_back(q);
            int st = t.fq_start[(size_t)ia * kq + q], nm = t.fq_num[(size_t)ia * kq + q];
            for (int it = st; it < st + nm; ++it) {
                term_s.push_back(t.fq_ib[it]);
                term_b.push_back(entry[(size_t)t.fq_id[it] * kq + (q - 1)]);
                term_c.push_back(t.fq_c[it]);
            }
        }
    }
    slot_group[na] = (int)grp_store.size();
    grp_term.push_back((int)term_s.size());
}

}  // namespace dnnf90
