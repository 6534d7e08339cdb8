// tables.hpp -- multi-index set, Faa di Bruno forward lists, adjoint pair
// lists and sigma^(q) polynomials for one (D0, K, closure) family, plus the
// flattened GPU schedules derived from them.
//
// Semantics follow lib/multi_index_bell_module.f90:128-351 (init_hod_tables):
// graded combinations-with-replacement order, downward closure of the seed
// multi-indices, slot 0 = the zero multi-index, pair lists for every b >= 2
// over all a >= b, forward lists for every (a, q >= 2) over b with
// |a|-|b| >= q-1 and coefficient C(a,b)/q, and tanh polynomials
// P_0 = t, P_{q+1} = (1 - t^2) P_q'.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace dnnf90 {

enum Activation : int {
    ACT_TANH = 0, ACT_SIN = 1, ACT_ERF = 2, ACT_BESSEL_J0 = 3, ACT_BESSEL_J1 = 4,
    ACT_GELU = 5,  // erf-GELU x*Phi(x); no counterpart in the reference
};

struct TableSet {
    int d0 = 0, k = 0, na = 0, iact = ACT_TANH;
    std::vector<int> alpha;           // [na][d0]
    std::vector<int> deg;             // [na]
    std::vector<int> ind_e1;          // [d0], -1 when e_i is not carried
    std::vector<int> pair_start, pair_num;  // [na]
    std::vector<int> pair_ia, pair_id;      // [P]
    std::vector<double> pair_binom;         // [P]
    std::vector<int> fq_start, fq_num;      // [na][k+1]
    std::vector<int> fq_ib, fq_id;          // [F]
    std::vector<double> fq_c;               // [F]
    std::vector<double> tpc;                // [k+2][k+3]: coefficient of t^i in P_q

    // Build; returns false (and leaves *this empty) on a bad argument:
    // d0 < 1, k outside [1,15], nseed < 0, a seed component outside [0,k]
    // or a seed of total order above k.
    bool build(int d0, int k, int nseed, const int *seeds, std::string *err = nullptr);
    int npair() const { return (int)pair_ia.size(); }
    int nfq() const { return (int)fq_ib.size(); }
};

// The forward Faa di Bruno step as a flat program over per-lane operand
// storage.  Entries 0..na-1 hold S_a (= B(a,1)); entries na.. hold the
// B(a,q), q >= 2, that a later term reads.  For slot a the groups
// slot_group[a] .. slot_group[a+1]-1 are q = 2..|a| in order; group g
// sums terms [grp_term[g], grp_term[g+1]) of coeff * E[ts] * E[tb] and,
// when grp_store[g] >= 0, keeps the sum in entry grp_store[g].
struct FwdSchedule {
    int nentry = 0;
    std::vector<int> slot_group;   // [na+1]
    std::vector<int> grp_term;     // [ngroup+1]
    std::vector<int> grp_store;    // [ngroup]
    std::vector<int> grp_q;        // [ngroup]
    std::vector<int> term_s;       // [F] operand entry of S_b
    std::vector<int> term_b;       // [F] operand entry of B(a-b, q-1)
    std::vector<double> term_c;    // [F]
    void build(const TableSet &t);
};

}  // namespace dnnf90
