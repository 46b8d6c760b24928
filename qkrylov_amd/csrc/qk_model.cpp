// Host model: basis ranking tables, site-operator compile, OpSum compile.
// This file is compiled with -ffp-contract=off so table entries are the
// reference's matrix elements bit for bit.
#include "qk_model.hpp"

#include <algorithm>
#include <cstdlib>
#include <map>
#include <tuple>

namespace qk {

void fail(const std::string& msg) { throw Error(msg); }

const char* site_kind_name(SiteKind k) {
    switch (k) {
        case SiteKind::SpinHalf: return "SpinHalf";
        case SiteKind::Fermion: return "Fermion";
        case SiteKind::Hubbard: return "Hubbard";
        case SiteKind::TJ: return "TJ";
    }
    return "?";
}

static u64 even_mask(int nbits) {
    u64 m = 0;
    for (int b = 0; b < nbits; b += 2) m |= u64(1) << b;
    return m;
}

static u64 low_bits(int n) { return n >= 64 ? ~u64(0) : ((u64(1) << n) - 1); }

// ---------------------------------------------------------------------------
// Basis
// ---------------------------------------------------------------------------
Basis::Basis(SiteKind kind, int nsites, const Sector& sector)
    : kind_(kind), nsites_(nsites), sector_(sector) {
    const bool two_species = (kind == SiteKind::Hubbard || kind == SiteKind::TJ);
    // Range checks of the reference constructors (spinhalf_basis.cpp:19-24,
    // fermion_basis.cpp:19-24, hubbard_basis.cpp:19-24, tj_basis.cpp:18-23).
    if (two_species) {
        if (nsites <= 0 || nsites > 31)
            fail(std::string(kind == SiteKind::Hubbard ? "HubbardBasis" : "TJBasis") +
                 ": N must satisfy 1 <= N <= 31");
    } else {
        if (nsites <= 0 || nsites > 63)
            fail(std::string(kind == SiteKind::SpinHalf ? "SpinHalfBasis" : "FermionBasis") +
                 ": N must satisfy 1 <= N <= 63");
    }
    nbits_ = two_species ? 2 * nsites : nsites;
    const u64 all = low_bits(nbits_);

    // Admissibility as popcount conditions.
    switch (kind) {
        case SiteKind::SpinHalf:
            // compute_sz2(s) == sz2 with sz2 = 2*popcount - N (spinhalf_basis.cpp:72-84)
            if (sector.use_sz) {
                m1_ = all;
                const int t = sector.sz2 + nsites;
                if (t % 2 != 0 || t < 0 || t / 2 > nsites) empty_ = true;
                q1_ = t / 2;
            }
            break;
        case SiteKind::Fermion:
            if (sector.use_n) {  // fermion_basis.cpp:81
                m1_ = all;
                q1_ = sector.n;
                if (q1_ < 0 || q1_ > nsites) empty_ = true;
            }
            break;
        case SiteKind::Hubbard:
        case SiteKind::TJ:
            // hubbard_basis.cpp:81-90, tj_basis.cpp:64-87: each flag separately.
            if (sector.use_nup) {
                m1_ = even_mask(nbits_);
                q1_ = sector.nup;
                if (q1_ < 0 || q1_ > nsites) empty_ = true;
            }
            if (sector.use_ndn) {
                m2_ = even_mask(nbits_) << 1;
                q2_ = sector.ndn;
                if (q2_ < 0 || q2_ > nsites) empty_ = true;
            }
            no_double_ = (kind == SiteKind::TJ);
            break;
    }

    // Split point: the lo table (2^L u16) must fit in LDS; rows should be
    // long, since a tile holds the ranks of one row and a short row leaves
    // lanes idle (N=20 chain sectors: 95 % of the lanes busy with L = 14,
    // 57 % with L = 10; the FTLM sweep over them 5,012 against 4,137
    // samples/s).  At least 2 high bits above 8 bits.
    if (nbits_ <= 8) {
        L_ = nbits_;
    } else {
        L_ = std::min(14, nbits_ - 2);
        if (two_species && (L_ & 1)) L_ -= 1;
    }
    // Split override for tuning; the split never changes ranks or results.
    if (const char* e = std::getenv("QKRYLOV_SPLIT_L")) {
        const int l = std::atoi(e);
        if (l >= 1 && l <= 15 && l <= nbits_ && nbits_ - l <= 26 && !(two_species && (l & 1)))
            L_ = l;
    }
    H_ = nbits_ - L_;
    if (H_ > 26) fail("basis too large: more than 2^26 distinct high halves");

    const u64 lomask_v = low_bits(L_);
    const u64 m1lo = m1_ & lomask_v, m2lo = m2_ & lomask_v;
    const u64 m1hi = m1_ >> L_, m2hi = m2_ >> L_;
    ncls1_ = popc(m1lo) + 1;
    ncls2_ = popc(m2lo) + 1;
    const int ncls = ncls1_ * ncls2_;

    // Lo classes in ascending lo order.
    const u64 nlo = u64(1) << L_;
    jlo_.assign(nlo, kLoInvalid);
    cls_cnt_.assign(ncls, 0);
    for (u64 lo = 0; lo < nlo; ++lo) {
        if (!half_ok(lo)) continue;
        const int c = popc(lo & m1lo) * ncls2_ + popc(lo & m2lo);
        jlo_[lo] = static_cast<u16>(cls_cnt_[c]++);
    }
    cls_off_.assign(ncls, 0);
    for (int c = 1; c < ncls; ++c) cls_off_[c] = cls_off_[c - 1] + cls_cnt_[c - 1];
    lo_list_.assign(cls_off_.back() + cls_cnt_.back(), 0);
    for (u64 lo = 0; lo < nlo; ++lo) {
        if (jlo_[lo] == kLoInvalid) continue;
        const int c = popc(lo & m1lo) * ncls2_ + popc(lo & m2lo);
        lo_list_[cls_off_[c] + jlo_[lo]] = static_cast<u32>(lo);
    }

    // Hi offsets.
    const u64 nhi = u64(1) << H_;
    jhi_.assign(nhi, kHiInvalid);
    hi_cls_.assign(nhi, -1);
    u64 acc = 0;
    if (!empty_) {
        for (u64 hi = 0; hi < nhi; ++hi) {
            if (!half_ok(hi << L_)) continue;
            const int need1 = q1_ - popc(hi & m1hi);
            const int need2 = q2_ - popc(hi & m2hi);
            if (m1_ == 0 && need1 != 0) continue;
            if (m2_ == 0 && need2 != 0) continue;
            if (need1 < 0 || need1 >= ncls1_ || need2 < 0 || need2 >= ncls2_) continue;
            const int c = need1 * ncls2_ + need2;
            if (cls_cnt_[c] == 0) continue;
            jhi_[hi] = acc;
            hi_cls_[hi] = c;
            valid_hi_.push_back(static_cast<u32>(hi));
            acc += cls_cnt_[c];
        }
    }
    dim_ = acc;
}

bool Basis::half_ok(u64 bits) const {
    if (!no_double_) return true;
    // site i occupies bits 2i (up) and 2i+1 (down)
    return ((bits & (bits >> 1)) & even_mask(64)) == 0;
}

u64 Basis::state(u64 i) const {
    if (i >= dim_) fail("basis state index out of range");
    // last valid hi with jhi[hi] <= i
    auto it = std::upper_bound(valid_hi_.begin(), valid_hi_.end(), i,
                               [&](u64 v, u32 hi) { return v < jhi_[hi]; });
    const u32 hi = *(it - 1);
    const u64 r = i - jhi_[hi];
    const u32 lo = lo_list_[cls_off_[hi_cls_[hi]] + r];
    return (u64(hi) << L_) | lo;
}

bool Basis::contains(u64 s) const { return index(s) >= 0; }

i64 Basis::index(u64 s) const {
    if (dim_ == 0) return -1;
    if (nbits_ < 64 && (s >> nbits_) != 0) return -1;
    const u64 hi = s >> L_;
    const u64 lo = s & lomask();
    if (jhi_[hi] == kHiInvalid) return -1;
    const u16 r = jlo_[lo];
    if (r == kLoInvalid) return -1;
    if (class_of_lo(lo) != hi_cls_[hi]) return -1;
    return static_cast<i64>(jhi_[hi] + r);
}

// ---------------------------------------------------------------------------
// Site operators
// ---------------------------------------------------------------------------
u64 BitFactor::support() const {
    return req1 | req0 | flip | (u64(1) << bA) | (u64(1) << bB);
}

static BitFactor diag_factor(int b, double v0, double v1) {
    BitFactor f;
    f.bA = f.bB = b;
    f.val[0] = {v0, 0.0};
    f.val[3] = {v1, 0.0};
    return f;
}

static BitFactor const_factor(int b, u64 req1, u64 req0, u64 flip, cplx v) {
    BitFactor f;
    f.bA = f.bB = b;
    f.req1 = req1;
    f.req0 = req0;
    f.flip = flip;
    for (auto& x : f.val) x = v;
    return f;
}

BitFactor compile_factor(SiteKind kind, const std::string& op, int site) {
    const bool two = (kind == SiteKind::Hubbard || kind == SiteKind::TJ);
    const int maxbit = two ? 2 * site + 1 : site;
    if (site < 0 || maxbit > 63)
        fail("site index " + std::to_string(site) + " outside the 64-bit state word");
    const cplx one{1.0, 0.0};
    switch (kind) {
        case SiteKind::SpinHalf: {
            const u64 m = u64(1) << site;
            if (op == "Sz") return diag_factor(site, -0.5, 0.5);          // :32-41
            if (op == "Sp") return const_factor(site, 0, m, m, one);      // :46-59
            if (op == "Sm") return const_factor(site, m, 0, m, one);      // :64-77
            if (op == "Sx") return const_factor(site, 0, 0, m, {0.5, 0}); // :82-92
            if (op == "Sy") {                                              // :97-116
                BitFactor f = const_factor(site, 0, 0, m, {0, 0});
                f.val[0] = {0.0, 0.5};    // down: +0.5i
                f.val[3] = {0.0, -0.5};   // up:   -0.5i
                return f;
            }
            fail("Unknown spin operator: " + op);
        }
        case SiteKind::Fermion: {
            const u64 m = u64(1) << site;
            const u64 jw = m - 1;  // fermion_site.cpp:19-30
            if (op == "N") return diag_factor(site, 0.0, 1.0);
            if (op == "Id") return const_factor(site, 0, 0, 0, one);
            if (op == "C") {
                BitFactor f = const_factor(site, m, 0, m, one);
                f.jw = jw;
                f.jw_sign = true;
                return f;
            }
            if (op == "Cdag") {
                BitFactor f = const_factor(site, 0, m, m, one);
                f.jw = jw;
                f.jw_sign = true;
                return f;
            }
            fail("Unknown fermion operator: " + op);
        }
        case SiteKind::Hubbard:
        case SiteKind::TJ: {
            const int bu = 2 * site, bd = 2 * site + 1;
            const u64 mu = u64(1) << bu, md = u64(1) << bd;
            const u64 jwu = mu - 1, jwd = md - 1;  // hubbard_site.cpp:31-49
            const bool tj = (kind == SiteKind::TJ);
            if (op == "Nup") return diag_factor(bu, 0.0, 1.0);
            if (op == "Ndn") return diag_factor(bd, 0.0, 1.0);
            if (!tj && op == "Nupdn") {
                BitFactor f;
                f.bA = bu;
                f.bB = bd;
                f.val[3] = one;
                return f;
            }
            BitFactor f;
            if (op == "CUp") f = const_factor(bu, mu, 0, mu, one);
            else if (op == "CdagUp") f = const_factor(bu, 0, tj ? (mu | md) : mu, mu, one);
            else if (op == "CDn") f = const_factor(bd, md, 0, md, one);
            else if (op == "CdagDn") f = const_factor(bd, 0, tj ? (mu | md) : md, md, one);
            else fail(std::string(tj ? "Unknown TJ operator: " : "Unknown Hubbard operator: ") + op);
            f.jw = (op == "CUp" || op == "CdagUp") ? jwu : jwd;
            f.jw_sign = true;
            return f;
        }
    }
    fail("unknown site kind");
}

bool apply_factor(const BitFactor& f, u64 s, u64* s_out, cplx* me_out) {
    if ((s & f.req1) != f.req1 || (s & f.req0) != 0) return false;
    const int idx = int((s >> f.bA) & 1) + 2 * int((s >> f.bB) & 1);
    cplx me = f.val[idx];
    if (f.jw_sign) {
        // the reference's element is Complex(phase, 0) (fermion_site.cpp:61-78)
        const double phase = (popc(s & f.jw) % 2 == 0) ? 1.0 : -1.0;
        me = cmul(cplx{phase, 0.0}, me);
    }
    *s_out = s ^ f.flip;
    *me_out = me;
    return true;
}

bool eval_term(const std::vector<BitFactor>& fs, cplx coeff, u64 s, u64* t, cplx* amp) {
    u64 st = s;
    cplx a = coeff;
    for (const auto& f : fs) {
        cplx me;
        if (!apply_factor(f, st, &st, &me)) return false;
        a = cmul(a, me);
    }
    *t = st;
    *amp = a;
    return true;
}

// ---------------------------------------------------------------------------
// OpSum -> group program
// ---------------------------------------------------------------------------
namespace {

struct TermInfo {
    std::vector<BitFactor> fs;
    cplx coeff;
    u64 F = 0, J = 0, S = 0;
    bool outside = false;
};

constexpr int kMaxSupport = 8;

// Table over the support positions of `pos` for the terms in `ids`.
std::vector<cplx> build_table(const std::vector<int>& pos, const std::vector<int>& ids,
                              const std::vector<TermInfo>& ti) {
    const int k = static_cast<int>(pos.size());
    std::vector<cplx> tab(size_t(1) << k, cplx{0.0, 0.0});
    for (u64 c = 0; c < (u64(1) << k); ++c) {
        u64 s = 0;
        for (int b = 0; b < k; ++b)
            if ((c >> b) & 1) s |= u64(1) << pos[b];
        for (int id : ids) {
            u64 t;
            cplx amp;
            if (!eval_term(ti[id].fs, ti[id].coeff, s, &t, &amp)) continue;
            // zero-valued entries are kept by the reference (cpp:81-84); they
            // add 0 here as well.
            const cplx v = cconj(amp);
            tab[c] = cadd(tab[c], v);
        }
    }
    return tab;
}

std::vector<int> bits_of(u64 m) {
    std::vector<int> v;
    for (int b = 0; b < 64; ++b)
        if ((m >> b) & 1) v.push_back(b);
    return v;
}

}  // namespace

Program compile_program(const Basis& basis, SiteKind kind, const OpSum& ops) {
    Program prog;
    std::vector<TermInfo> ti(ops.terms.size());
    for (size_t t = 0; t < ops.terms.size(); ++t) {
        const auto& term = ops.terms[t];
        TermInfo& x = ti[t];
        x.coeff = term.coeff;
        u64 prior = 0;   // flips applied before the current factor
        u64 jw_all = 0;
        for (const auto& f : term.factors) {
            BitFactor bf = compile_factor(kind, f.op, f.site);
            x.S |= bf.support();
            if (bf.jw_sign) jw_all ^= bf.jw;
            prior ^= bf.flip;
            x.fs.push_back(bf);
        }
        // Bits at or above nbits are zero in every basis state: they carry no
        // configuration and no parity.  A flip that sets one of them leaves
        // the basis, and the reference drops such entries (cpp:71).
        const u64 all = low_bits(basis.nbits());
        x.F = prior;
        x.S &= all;
        x.J = jw_all & ~x.S & all;
        x.outside = (prior & ~all) != 0;
        if (popc(x.S) > kMaxSupport)
            fail("term " + std::to_string(t) + " touches more than " +
                 std::to_string(kMaxSupport) + " bits");
    }

    const int L = basis.L();
    const u64 lom = basis.lomask();
    auto kind_of = [&](u64 F) {
        if (F == 0) return int(kDiag);
        if ((F & ~lom) == 0) return int(kLo);
        if ((F & lom) == 0) return int(kHi);
        return int(kMix);
    };

    // Greedy grouping by (F, J) in OpSum order, support union capped.
    struct Acc {
        u64 F, J, S;
        std::vector<int> ids;
    };
    std::vector<Acc> accs;
    std::map<std::pair<u64, u64>, std::vector<size_t>> open;
    for (size_t t = 0; t < ti.size(); ++t) {
        if (ti[t].outside) continue;
        const auto key = std::make_pair(ti[t].F, ti[t].J);
        bool placed = false;
        for (size_t ai : open[key]) {
            if (popc(accs[ai].S | ti[t].S) <= (ti[t].F == 0 ? 6 : kMaxSupport)) {
                accs[ai].S |= ti[t].S;
                accs[ai].ids.push_back(int(t));
                placed = true;
                break;
            }
        }
        if (!placed) {
            accs.push_back({ti[t].F, ti[t].J, ti[t].S, {int(t)}});
            open[key].push_back(accs.size() - 1);
        }
    }

    std::vector<Group> by_kind[4];
    for (const auto& a : accs) {
        Group g;
        g.flip = a.F;
        g.jw = a.J;
        g.pos = bits_of(a.S);
        g.table = build_table(g.pos, a.ids, ti);
        g.kind = kind_of(a.F);
        g.hi_only = ((a.S | a.J) & lom) == 0;
        g.terms = a.ids;
        bool nonzero = false;
        for (const auto& v : g.table) nonzero |= (v.re != 0.0 || v.im != 0.0);
        if (!nonzero) continue;  // never contributes (e.g. Sp Sp on one site)
        prog.needs_jw |= (g.jw != 0);
        prog.max_support = std::max<int>(prog.max_support, int(g.pos.size()));
        by_kind[g.kind].push_back(std::move(g));
    }
    prog.n_diag = int(by_kind[kDiag].size());
    prog.n_lo = int(by_kind[kLo].size());
    prog.n_hi = int(by_kind[kHi].size());
    prog.n_mix = int(by_kind[kMix].size());
    for (auto& v : by_kind)
        for (auto& g : v) prog.groups.push_back(std::move(g));

    // Per-term diagonal list, OpSum order (for the exact diagonal()).
    for (size_t t = 0; t < ti.size(); ++t) {
        if (ti[t].outside || ti[t].F != 0) continue;
        Group g;
        g.flip = 0;
        g.jw = ti[t].J;
        g.pos = bits_of(ti[t].S);
        g.table = build_table(g.pos, {int(t)}, ti);
        g.kind = kDiag;
        g.terms = {int(t)};
        prog.diag_terms.push_back(std::move(g));
    }
    (void)L;
    return prog;
}

}  // namespace qk
