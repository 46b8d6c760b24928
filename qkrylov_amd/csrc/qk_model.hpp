// Host-side model: sectors, bases with closed-form ranking tables, site
// operators compiled to bit actions, OpSums, and the OpSum -> group program
// that the HIP apply kernels execute.
#pragma once

#include "qk_common.hpp"

#include <memory>
#include <string>
#include <vector>

namespace qk {

// Quantum-number selection; same fields and meaning as the reference
// Sector (include/qkrylov/symmetry/sector.hpp:10-41).
struct Sector {
    bool use_sz = false;
    int sz2 = 0;
    bool use_nup = false;
    bool use_ndn = false;
    int nup = 0;
    int ndn = 0;
    bool use_n = false;
    int n = 0;
    bool use_nb = false;
    int nb = 0;
};

enum class SiteKind : int { SpinHalf = 0, Fermion = 1, Hubbard = 2, TJ = 3 };
const char* site_kind_name(SiteKind k);

// ---------------------------------------------------------------------------
// Basis.  The reference enumerates every admissible bit string in ascending
// integer order and indexes them by position (src/basis/spinhalf_basis.cpp:
// 86-104, fermion_basis.cpp:72-89, hubbard_basis.cpp:81-100,
// tj_basis.cpp:56-96).  Here the same order is produced without enumeration:
// a state s is split into lo = s mod 2^L and hi = s >> L.  Admissibility is a
// condition on popcounts of two counted masks (plus, for t-J, no doubly
// occupied site), so for a fixed hi the admissible lo values are exactly the
// lo values of one "class" (their counted popcounts are fixed by hi), and
// ascending integer order is "ascending hi, then ascending lo".  Hence
//     index(s) = jhi[hi] + jlo[lo]
// with jlo[lo] = rank of lo inside its class and jhi[hi] = number of basis
// states whose hi part is smaller.  L is even for Hubbard/t-J so that a site
// never straddles the split.
// ---------------------------------------------------------------------------
class Basis {
public:
    static constexpr u16 kLoInvalid = 0xFFFF;
    static constexpr u64 kHiInvalid = ~u64(0);

    Basis(SiteKind kind, int nsites, const Sector& sector);

    SiteKind kind() const { return kind_; }
    int nsites() const { return nsites_; }
    int nbits() const { return nbits_; }
    const Sector& sector() const { return sector_; }

    u64 size() const { return dim_; }
    // state(i); throws for i >= size() (reference: states_.at(i)).
    u64 state(u64 i) const;
    bool contains(u64 s) const;
    // index(s) or -1 when s is not a basis state.
    i64 index(u64 s) const;

    // Ranking-table view used by the device upload and by the kernels.
    int L() const { return L_; }
    int H() const { return H_; }
    u64 m1() const { return m1_; }
    u64 m2() const { return m2_; }
    int ncls2() const { return ncls2_; }
    int class_of_lo(u64 lo) const {
        return popc(lo & m1_ & lomask()) * ncls2_ + popc(lo & m2_ & lomask());
    }
    u64 lomask() const { return L_ >= 64 ? ~u64(0) : ((u64(1) << L_) - 1); }
    const std::vector<u16>& jlo() const { return jlo_; }
    const std::vector<u64>& jhi() const { return jhi_; }
    const std::vector<int32_t>& hi_cls() const { return hi_cls_; }
    const std::vector<u32>& cls_off() const { return cls_off_; }
    const std::vector<u32>& cls_cnt() const { return cls_cnt_; }
    const std::vector<u32>& lo_list() const { return lo_list_; }
    const std::vector<u32>& valid_hi() const { return valid_hi_; }

private:
    bool half_ok(u64 bits) const;  // no double occupancy inside this half (t-J)

    SiteKind kind_;
    int nsites_;
    int nbits_;
    Sector sector_;
    u64 m1_ = 0, m2_ = 0;   // counted masks
    int q1_ = 0, q2_ = 0;   // required counts
    bool no_double_ = false;
    bool empty_ = false;
    int L_ = 0, H_ = 0;
    int ncls1_ = 1, ncls2_ = 1;
    u64 dim_ = 0;
    std::vector<u16> jlo_;
    std::vector<u64> jhi_;
    std::vector<int32_t> hi_cls_;
    std::vector<u32> cls_off_, cls_cnt_, lo_list_;
    std::vector<u32> valid_hi_;   // ascending admissible hi values
};

// ---------------------------------------------------------------------------
// Site operators as bit actions.  One factor acts on the current state s:
//   valid iff (s & req1) == req1 and (s & req0) == 0
//   element = val[bit(s,bA) + 2*bit(s,bB)], times -1 when popcount(s & jw)
//   is odd (Jordan-Wigner phase of the state before the action)
//   new state = s ^ flip
// These reproduce SpinHalfSite/FermionSite/HubbardSite/TJSite::apply
// (src/sites/spinhalf_site.cpp:18-121, fermion_site.cpp:19-84,
//  hubbard_site.cpp:51-125, tj_site.cpp:47-113).
// ---------------------------------------------------------------------------
struct BitFactor {
    u64 req1 = 0, req0 = 0, flip = 0, jw = 0;
    int bA = 0, bB = 0;
    bool jw_sign = false;   // element gets the (+-1, 0) phase factor
    cplx val[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    u64 support() const;
};

// Throws qk::Error for an operator name the site does not define, the way
// the reference's Site::apply throws; also for a site whose bits fall outside
// the 64-bit state word.
BitFactor compile_factor(SiteKind kind, const std::string& op, int site);

// Evaluate one factor on a state (host reference used for tables).
bool apply_factor(const BitFactor& f, u64 s, u64* s_out, cplx* me_out);

// ---------------------------------------------------------------------------
// OpSum: append-only list of (coefficient, [(op, site), ...]) as in the
// reference (include/qkrylov/operators/operator_term.hpp:13-32, opsum.hpp).
// ---------------------------------------------------------------------------
struct OpFactor {
    std::string op;
    int site;
};
struct OpTerm {
    cplx coeff{0, 0};
    std::vector<OpFactor> factors;
};
struct OpSum {
    std::vector<OpTerm> terms;
};

// ---------------------------------------------------------------------------
// Compiled program.  Each term t of the OpSum maps a state s to
// s ^ F_t with amplitude amp_t(s) = sign(s & J_t) * A_t(config of s on the
// support S_t), where every factor's requirement, value and flip lives on
// S_t, and J_t collects the Jordan-Wigner bits outside S_t.  Terms with equal
// (F, J) share one group: their tables are added per configuration (the
// table stores conj(amp), the reference's gather value, cpp:75-84).
// ---------------------------------------------------------------------------
enum GroupKind : int { kDiag = 0, kLo = 1, kHi = 2, kMix = 3 };

struct Group {
    u64 flip = 0;
    u64 jw = 0;
    std::vector<int> pos;       // support bit positions, config bit k = bit(s, pos[k])
    std::vector<cplx> table;    // 2^|pos| entries: conj(amplitude) summed over terms
    int kind = kDiag;
    bool hi_only = false;       // support and jw inside the hi part
    std::vector<int> terms;     // OpSum term ids (debug / introspection)
};

struct Program {
    std::vector<Group> groups;       // diag groups first, then lo, hi, mixed
    std::vector<Group> diag_terms;   // one group per diagonal term, OpSum order
    int n_diag = 0, n_lo = 0, n_hi = 0, n_mix = 0;
    bool needs_jw = false;
    int max_support = 0;
};

// Compile an OpSum for a basis; operator names are those of `site`, which
// the reference does not require to match the basis kind (c_api.cpp:376-387).
// Every op name is validated up front.
Program compile_program(const Basis& basis, SiteKind site, const OpSum& ops);

// The amplitude of one term on one state, evaluated factor by factor as the
// reference does (cpp:47-68).  Used by host checks.
bool eval_term(const std::vector<BitFactor>& fs, cplx coeff, u64 s, u64* t, cplx* amp);

}  // namespace qk
