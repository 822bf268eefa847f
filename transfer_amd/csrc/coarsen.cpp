// coarsen.cpp -- the coarse operators of the multilevel preconditioner
// (extension; the reference preconditions with the angular-diagonal
// operator only, main.cpp:223-248).
//
// The operator is a sum of Kronecker products, A = sum_c Omega_c (x) K_c +
// sum_n O_n (x) B_n, spatial matrices K_c, B_n times angular matrices.  With
// a piecewise-constant aggregation P (x) I of the vertices, the Galerkin
// operator (P^T (x) I) A (P (x) I) = sum_c Omega_c (x) P^T K_c P + ... has the
// same form: the same angular tables and channel count, and a stencil whose
// channel coefficients are sums of the fine ones.  It is applied by the
// same kernels.  Aggregates are runs of up to `size` consecutive vertices of
// the (Morton) internal numbering, never straddling two problems of a batch.
#include <algorithm>
#include <array>
#include <unordered_map>
#include <vector>

#include "internal.h"
#include "stencil.h"

namespace tr {

int coarsen_operator(const HostOperator &f, const std::vector<uint32_t> &voff, uint32_t size, HostOperator &c,
                     std::vector<uint32_t> &agg, std::vector<uint32_t> &cvoff) {
    if (size < 2) return fail(TR_ERR_ARG, "coarsen_operator: aggregate size < 2");
    const uint32_t nP = f.nP, NC = f.NC;
    // aggregates: [agg[I], agg[I+1]) per problem range
    std::vector<uint32_t> ranges = voff.empty() ? std::vector<uint32_t>{0, nP} : voff;
    agg.clear();
    cvoff.clear();
    for (size_t p = 0; p + 1 < ranges.size(); p++) {
        cvoff.push_back((uint32_t)agg.size());
        for (uint32_t v = ranges[p]; v < ranges[p + 1]; v += size) agg.push_back(v);
    }
    cvoff.push_back((uint32_t)agg.size());
    const uint32_t nC = (uint32_t)agg.size();
    agg.push_back(nP);
    std::vector<uint32_t> of(nP);
    for (uint32_t I = 0; I < nC; I++)
        for (uint32_t v = agg[I]; v < agg[I + 1]; v++) of[v] = I;

    c = HostOperator();
    c.nP = nC;
    c.slm = f.slm;
    c.aslm = f.aslm;
    c.NC = NC;
    c.arow = f.arow;
    c.brow = f.brow;
    c.aidx = f.aidx;
    c.bidx = f.bidx;
    c.aval = f.aval;
    c.bval = f.bval;
    c.massfac = f.massfac;
    c.ddiag = f.ddiag;
    c.bdiag = f.bdiag;
    c.omc = f.omc;
    c.onc = f.onc;
    c.mom_val = f.mom_val;
    c.mom_pos = f.mom_pos;
    c.perm.resize(nC);
    c.iperm.resize(nC);
    for (uint32_t I = 0; I < nC; I++) c.perm[I] = c.iperm[I] = I;
    c.rhs.assign(nC, 0.0);
    if (!f.mf0.empty()) {
        c.mf0.resize(nC);
        for (uint32_t I = 0; I < nC; I++) c.mf0[I] = f.mf0[agg[I]];  // constant within a problem
    }
    // volume stencil: sum the fine couplings of every (I, K) pair; the mass
    // channel carries mf0 per fine row, which is constant within an aggregate
    std::vector<std::vector<std::pair<uint32_t, std::vector<double>>>> rows(nC);
    for (uint32_t I = 0; I < nC; I++) {
        std::unordered_map<uint32_t, size_t> at;
        auto &row = rows[I];
        row.push_back({I, std::vector<double>(NC, 0.0)});
        at[I] = 0;
        for (uint32_t i = agg[I]; i < agg[I + 1]; i++)
            for (uint32_t s = 0; s < f.W; s++) {
                const uint32_t K = of[f.nbr[(size_t)s * nP + i]];
                auto it = at.find(K);
                size_t q;
                if (it == at.end()) {
                    q = row.size();
                    at[K] = q;
                    row.push_back({K, std::vector<double>(NC, 0.0)});
                } else {
                    q = it->second;
                }
                for (uint32_t ch = 0; ch < NC; ch++) row[q].second[ch] += f.coef[((size_t)s * NC + ch) * nP + i];
            }
        std::sort(row.begin() + 1, row.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
    }
    uint32_t W = 1;
    for (auto &r : rows) W = std::max<uint32_t>(W, (uint32_t)r.size());
    c.W = W;
    c.nbr.assign((size_t)W * nC, 0);
    c.coef.assign((size_t)W * NC * nC, 0.0);
    for (uint32_t I = 0; I < nC; I++)
        for (uint32_t s = 0; s < W; s++) {
            c.nbr[(size_t)s * nC + I] = s < rows[I].size() ? rows[I][s].first : I;
            if (s < rows[I].size())
                for (uint32_t ch = 0; ch < NC; ch++) c.coef[((size_t)s * NC + ch) * nC + I] = rows[I][s].second[ch];
        }
    // boundary stencil
    std::vector<std::vector<std::pair<uint32_t, std::array<double, 3>>>> brows(nC);
    for (uint32_t I = 0; I < nC; I++) {
        std::unordered_map<uint32_t, size_t> at;
        for (uint32_t i = agg[I]; i < agg[I + 1]; i++) {
            const int b = f.bof[i];
            if (b < 0) continue;
            for (uint32_t s = 0; s < f.WB; s++) {
                const uint32_t K = of[f.bnbr[(size_t)s * f.nB + b]];
                auto it = at.find(K);
                size_t q;
                if (it == at.end()) {
                    q = brows[I].size();
                    at[K] = q;
                    brows[I].push_back({K, {0.0, 0.0, 0.0}});
                } else {
                    q = it->second;
                }
                for (int n = 0; n < 3; n++) brows[I][q].second[n] += f.bcoef[((size_t)s * 3 + n) * f.nB + b];
            }
        }
    }
    c.bof.assign(nC, -1);
    c.WB = 1;
    for (uint32_t I = 0; I < nC; I++)
        if (!brows[I].empty()) {
            c.bof[I] = (int32_t)c.bvert.size();
            c.bvert.push_back(I);
            c.WB = std::max<uint32_t>(c.WB, (uint32_t)brows[I].size());
        }
    c.nB = (uint32_t)c.bvert.size();
    c.bnbr.assign((size_t)c.WB * std::max<uint32_t>(c.nB, 1), 0);
    c.bcoef.assign((size_t)c.WB * 3 * std::max<uint32_t>(c.nB, 1), 0.0);
    for (uint32_t b = 0; b < c.nB; b++) {
        const auto &r = brows[c.bvert[b]];
        for (uint32_t s = 0; s < c.WB; s++) {
            c.bnbr[(size_t)s * c.nB + b] = s < r.size() ? r[s].first : c.bvert[b];
            for (int n = 0; n < 3; n++)
                c.bcoef[((size_t)s * 3 + n) * c.nB + b] = s < r.size() ? r[s].second[n] : 0.0;
        }
    }
    return TR_OK;
}

}  // namespace tr
