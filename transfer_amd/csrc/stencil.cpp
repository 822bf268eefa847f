// stencil.cpp -- renumbering and FP64 assembly of the device operator data
// (see stencil.h for the formulation).
#include "stencil.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <numeric>

#include "internal.h"

namespace tr {

static uint64_t spread3(uint64_t x) {
    x &= 0x1fffffULL;
    x = (x | x << 32) & 0x1f00000000ffffULL;
    x = (x | x << 16) & 0x1f0000ff0000ffULL;
    x = (x | x << 8) & 0x100f00f00f00f00fULL;
    x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
    x = (x | x << 2) & 0x1249249249249249ULL;
    return x;
}

// Neighbour sets (vertices sharing a tet) in caller numbering.
static void vertex_graph(const tr_mesh_view &mv, std::vector<uint32_t> &start,
                         std::vector<uint32_t> &adj) {
    start.assign(mv.nP + 1, 0);
    adj.clear();
    std::vector<uint32_t> tmp;
    for (uint32_t v = 0; v < mv.nP; v++) {
        tmp.clear();
        for (uint32_t j = mv.tetstart[v]; j < mv.tetstart[v + 1]; j++) {
            const tr_tetrahedron &t = mv.mesh[mv.tetidx[j]];
            for (int k = 0; k < 4; k++)
                if (t.p[k] != v) tmp.push_back(t.p[k]);
        }
        std::sort(tmp.begin(), tmp.end());
        tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
        adj.insert(adj.end(), tmp.begin(), tmp.end());
        start[v + 1] = (uint32_t)adj.size();
    }
}

// Reverse Cuthill-McKee over every connected component.
static void rcm_order(uint32_t nP, const std::vector<uint32_t> &start,
                      const std::vector<uint32_t> &adj, std::vector<uint32_t> &perm) {
    std::vector<uint32_t> deg(nP), order;
    for (uint32_t v = 0; v < nP; v++) deg[v] = start[v + 1] - start[v];
    std::vector<uint32_t> byDeg(nP);
    std::iota(byDeg.begin(), byDeg.end(), 0);
    std::stable_sort(byDeg.begin(), byDeg.end(),
                     [&](uint32_t a, uint32_t b) { return deg[a] < deg[b]; });
    std::vector<char> seen(nP, 0);
    order.reserve(nP);
    std::vector<uint32_t> nb;
    for (uint32_t root : byDeg) {
        if (seen[root]) continue;
        size_t head = order.size();
        order.push_back(root);
        seen[root] = 1;
        while (head < order.size()) {
            uint32_t v = order[head++];
            nb.clear();
            for (uint32_t j = start[v]; j < start[v + 1]; j++)
                if (!seen[adj[j]]) nb.push_back(adj[j]);
            std::stable_sort(nb.begin(), nb.end(),
                             [&](uint32_t a, uint32_t b) { return deg[a] < deg[b]; });
            for (uint32_t u : nb) {
                seen[u] = 1;
                order.push_back(u);
            }
        }
    }
    perm.assign(order.rbegin(), order.rend());
}

unsigned mirror_odd(int l, int m) {
    (void)l;  // l is even: the z-mirror parity (-1)^(l+|m|) depends on |m| only
    const int am = m < 0 ? -m : m;
    unsigned bits = 0;
    // x -> -x: phi -> pi - phi; cos(m(pi - phi)) = (-1)^m cos(m phi),
    // sin(k(pi - phi)) = -(-1)^k sin(k phi)
    if (m >= 0 ? (am & 1) : !(am & 1)) bits |= 1u;
    // y -> -y: phi -> -phi flips the sine terms
    if (m < 0) bits |= 2u;
    // z -> -z: cos(theta) -> -cos(theta), P_l^|m| picks up (-1)^(l+|m|)
    if (am & 1) bits |= 4u;
    return bits;
}

int assemble_operator(const tr_mesh_view &mv, const tr_angular_view &av,
                      HostOperator &op) {
    const uint32_t nP = mv.nP, slm = av.slm;
    if (nP == 0 || slm == 0 || slm > 128)
        return fail(TR_ERR_ARG, "assemble_operator: empty mesh or slm out of 1..128");
    op.nP = nP;
    op.slm = slm;
    op.aslm = (slm + 7) & ~7u;
    // the stiffness couplings divide by 9 kappa V (kernels.cu:252): an
    // opacity of zero (or a degenerate tet) has no operator
    for (uint32_t t = 0; t < mv.nT; t++) {
        const double kv = mv.mesh[t].kappa_volume;
        if (!(kv > 0) || !std::isfinite(kv)) {
            char msg[160];
            snprintf(msg, sizeof msg, "assemble_operator: tet %u has kappa*V = %g (must be positive and finite)", t, kv);
            return fail(TR_ERR_ARG, msg);
        }
        for (int k = 0; k < 4; k++)
            if (mv.mesh[t].p[k] >= nP) return fail(TR_ERR_ARG, "assemble_operator: tet vertex out of range");
    }

    // ---- renumbering
    std::vector<uint32_t> gstart, gadj;
    vertex_graph(mv, gstart, gadj);
    op.perm.resize(nP);
    if (mv.coords) {
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (uint32_t v = 0; v < nP; v++)
            for (int a = 0; a < 3; a++) {
                lo[a] = std::min(lo[a], mv.coords[3 * v + a]);
                hi[a] = std::max(hi[a], mv.coords[3 * v + a]);
            }
        double span = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-300});
        std::vector<uint64_t> code(nP);
        for (uint32_t v = 0; v < nP; v++) {
            uint64_t q[3];
            for (int a = 0; a < 3; a++)
                q[a] = (uint64_t)std::min(2097151.0, std::floor((mv.coords[3 * v + a] - lo[a]) / span * 2097151.0));
            code[v] = spread3(q[0]) | spread3(q[1]) << 1 | spread3(q[2]) << 2;
        }
        std::iota(op.perm.begin(), op.perm.end(), 0);
        std::stable_sort(op.perm.begin(), op.perm.end(),
                         [&](uint32_t a, uint32_t b) { return code[a] < code[b]; });
    } else {
        rcm_order(nP, gstart, gadj, op.perm);
    }
    // vertices on vacuum faces first, each group in the order above: their
    // wavefronts take the boundary couplings, and packing them together
    // leaves every other wavefront free of boundary work
    {
        std::vector<uint8_t> onb(nP, 0);
        for (uint32_t f = 0; f < mv.nF; f++)
            if (!mv.face_kind || mv.face_kind[f] == 0)
                for (int k = 0; k < 3; k++) onb[mv.bnd[f].p[k]] = 1;
        std::stable_partition(op.perm.begin(), op.perm.end(), [&](uint32_t v) { return onb[v] != 0; });
    }
    op.iperm.resize(nP);
    for (uint32_t i = 0; i < nP; i++) op.iperm[op.perm[i]] = i;

    // ---- angular channel matrices, [out][in] (kernels.cu:276-282 reads
    // omega[a][in][out] with b = omega_pos[a][in][out])
    std::vector<double> Om((size_t)9 * slm * slm, 0.0);  // [a][b][out][in]
    for (uint32_t a = 0; a < 3; a++)
        for (uint32_t in = 0; in < slm; in++)
            for (uint32_t out = 0; out < slm; out++) {
                size_t s = (size_t)a * slm * slm + (size_t)in * slm + out;
                uint32_t b = av.omega_pos[s];
                if (b > 2) return fail(TR_ERR_ARG, "omega_pos entry out of 0..2");
                Om[((size_t)(a * 3 + b) * slm + out) * slm + in] = av.omega[s];
            }
    auto O = [&](int a, int b, uint32_t out, uint32_t in) {
        return Om[((size_t)(a * 3 + b) * slm + out) * slm + in];
    };
    double asym = 0, omax = 0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            for (uint32_t o = 0; o < slm; o++)
                for (uint32_t i = 0; i < slm; i++) {
                    omax = std::max(omax, std::fabs(O(a, b, o, i)));
                    if (a < b) asym = std::max(asym, std::fabs(O(a, b, o, i) - O(b, a, o, i)));
                }
    op.NC = asym > 1e-12 * std::max(omax, 1e-300) ? 10 : 7;
    const uint32_t NC = op.NC;
    static const int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};
    op.omc.assign((size_t)(NC - 1) * slm * slm, 0.0);
    for (uint32_t o = 0; o < slm; o++)
        for (uint32_t i = 0; i < slm; i++) {
            for (int a = 0; a < 3; a++)
                op.omc[((size_t)a * slm + o) * slm + i] = O(a, a, o, i);
            for (int q = 0; q < 3; q++) {
                double ab = O(pa[q], pb[q], o, i), ba = O(pb[q], pa[q], o, i);
                op.omc[((size_t)(3 + q) * slm + o) * slm + i] = 0.5 * (ab + ba);
                if (NC == 10) op.omc[((size_t)(6 + q) * slm + o) * slm + i] = 0.5 * (ab - ba);
            }
        }
    op.onc.assign((size_t)3 * slm * slm, 0.0);
    const double *On[3] = {av.Ox, av.Oy, av.Oz};
    for (int n = 0; n < 3; n++)
        for (uint32_t o = 0; o < slm; o++)
            for (uint32_t i = 0; i < slm; i++)
                op.onc[((size_t)n * slm + o) * slm + i] = On[n][(size_t)i * slm + o];

    auto build_table = [&](const std::vector<double> &mats, uint32_t nmat, uint32_t &rows,
                           std::vector<uint16_t> &idx, std::vector<double> &val) {
        double mx = 0;
        for (double v : mats) mx = std::max(mx, std::fabs(v));
        const double thr = 1e-14 * std::max(mx, 1e-300);
        std::vector<std::vector<std::pair<uint16_t, double>>> R(slm);
        for (uint32_t o = 0; o < slm; o++)
            for (uint32_t c = 0; c < nmat; c++)
                for (uint32_t i = 0; i < slm; i++) {
                    double v = mats[((size_t)c * slm + o) * slm + i];
                    if (std::fabs(v) > thr) R[o].push_back({(uint16_t)(c * slm + i), v});
                }
        rows = 0;
        for (auto &r : R) rows = std::max<uint32_t>(rows, (uint32_t)r.size());
        // padded to a multiple of 8 entries (zero couplings) for the unrolled
        // contraction loop
        rows = std::max<uint32_t>((rows + 7) / 8 * 8, 8);
        idx.assign((size_t)rows * slm, 0);
        val.assign((size_t)rows * slm, 0.0);
        for (uint32_t o = 0; o < slm; o++)
            for (size_t e = 0; e < R[o].size(); e++) {
                idx[e * slm + o] = R[o][e].first;
                val[e * slm + o] = R[o][e].second;
            }
    };
    build_table(op.omc, NC - 1, op.arow, op.aidx, op.aval);
    build_table(op.onc, 3, op.brow, op.bidx, op.bval);

    op.massfac.assign(slm, 1.0);
    if (av.massfac)
        for (uint32_t lm = 0; lm < slm; lm++) op.massfac[lm] = av.massfac[lm];
    op.ddiag.resize((size_t)3 * slm);
    op.bdiag.resize((size_t)3 * slm);
    for (int a = 0; a < 3; a++)
        for (uint32_t lm = 0; lm < slm; lm++) {
            op.ddiag[(size_t)a * slm + lm] = av.omega[(size_t)a * slm * slm + (size_t)lm * slm + lm];
            op.bdiag[(size_t)a * slm + lm] = On[a][(size_t)lm * slm + lm];
        }

    // ---- volume stencil in the new numbering
    std::vector<uint32_t> cnt(nP);
    uint32_t W = 1;
    for (uint32_t i = 0; i < nP; i++) {
        uint32_t vo = op.perm[i];
        cnt[i] = 1 + gstart[vo + 1] - gstart[vo];
        W = std::max(W, cnt[i]);
    }
    // pad to the unrolled slot counts of the buffer-load kernel (15 = the
    // Kuhn interior degree + 1); extra slots point at the vertex itself with
    // zero couplings
    if (W < 15) W = 15;
    op.W = W;
    op.nbr.assign((size_t)W * nP, 0);
    op.coef.assign((size_t)W * NC * nP, 0.0);
    std::vector<uint32_t> slots;
    for (uint32_t i = 0; i < nP; i++) {
        uint32_t vo = op.perm[i];
        slots.clear();
        slots.push_back(i);
        for (uint32_t j = gstart[vo]; j < gstart[vo + 1]; j++) slots.push_back(op.iperm[gadj[j]]);
        std::sort(slots.begin() + 1, slots.end());
        for (uint32_t s = 0; s < W; s++)
            op.nbr[(size_t)s * nP + i] = s < slots.size() ? slots[s] : i;
        auto slot_of = [&](uint32_t v) {
            for (uint32_t s = 0; s < slots.size(); s++)
                if (slots[s] == v) return s;
            return (uint32_t)0;
        };
        for (uint32_t j = mv.tetstart[vo]; j < mv.tetstart[vo + 1]; j++) {
            const tr_tetrahedron &t = mv.mesh[mv.tetidx[j]];
            const uint32_t li = mv.tetpos[j];
            const double kv = t.kappa_volume, inv = 1.0 / (9.0 * kv);
            const double *si = t.s[li];
            for (int k = 0; k < 4; k++) {
                uint32_t s = slot_of(op.iperm[t.p[k]]);
                const double *sk = t.s[k];
                double ch[10];
                ch[0] = kv * (1.0 / 20.0) * (k == (int)li ? 2.0 : 1.0);
                ch[1] = si[0] * sk[0] * inv;
                ch[2] = si[1] * sk[1] * inv;
                ch[3] = si[2] * sk[2] * inv;
                ch[4] = (si[0] * sk[1] + si[1] * sk[0]) * inv;
                ch[5] = (si[0] * sk[2] + si[2] * sk[0]) * inv;
                ch[6] = (si[1] * sk[2] + si[2] * sk[1]) * inv;
                ch[7] = (si[0] * sk[1] - si[1] * sk[0]) * inv;
                ch[8] = (si[0] * sk[2] - si[2] * sk[0]) * inv;
                ch[9] = (si[1] * sk[2] - si[2] * sk[1]) * inv;
                for (uint32_t c = 0; c < NC; c++) op.coef[((size_t)s * NC + c) * nP + i] += ch[c];
            }
        }
    }

    // ---- vacuum-boundary stencil
    op.bvert.clear();
    op.gvert.clear();
    op.gstart.assign(1, 0);
    op.gface.clear();
    op.gpos.clear();
    std::vector<int64_t> gid(op.general_normals ? mv.nF : 0, -1);  // face -> general face index
    std::vector<uint32_t> gfaces;                                   // general face index -> face
    std::vector<std::vector<uint32_t>> bsl;
    std::vector<std::vector<double>> bco;
    for (uint32_t i = 0; i < nP; i++) {
        uint32_t vo = op.perm[i];
        std::vector<uint32_t> sl;
        std::vector<double> co;
        const size_t g0 = op.gface.size();
        for (uint32_t j = mv.facestart[vo]; j < mv.facestart[vo + 1]; j++) {
            uint32_t fi = mv.faceidx[j];
            if (mv.face_kind && mv.face_kind[fi] != 0) continue;
            const tr_face &F = mv.bnd[fi];
            const uint32_t lf = mv.facepos[j];
            if (op.general_normals) {
                // faces not along an axis take their exact angular matrix
                // (the reference's axis rule, kernels.cu:359-361, is exact
                // only along the axes)
                const double S = std::sqrt(F.s[0] * F.s[0] + F.s[1] * F.s[1] + F.s[2] * F.s[2]);
                const double mx = std::max({std::fabs(F.s[0]), std::fabs(F.s[1]), std::fabs(F.s[2])});
                if (S > 0 && mx < (1.0 - 1e-12) * S) {
                    if (gid[fi] < 0) {
                        gid[fi] = (int64_t)gfaces.size();
                        gfaces.push_back(fi);
                    }
                    op.gface.push_back((uint32_t)gid[fi]);
                    op.gpos.push_back(lf);
                    continue;
                }
            }
            double surf = std::fabs(F.s[0] + F.s[1] + F.s[2]);
            int axis = (2 * std::fabs(F.s[0]) > surf) ? 0 : ((2 * std::fabs(F.s[1]) > surf) ? 1 : 2);
            for (int k = 0; k < 3; k++) {
                uint32_t v = op.iperm[F.p[k]];
                size_t s = std::find(sl.begin(), sl.end(), v) - sl.begin();
                if (s == sl.size()) {
                    sl.push_back(v);
                    co.insert(co.end(), {0.0, 0.0, 0.0});
                }
                co[3 * s + axis] += surf * (1.0 / 12.0) * (k == (int)lf ? 2.0 : 1.0);
            }
        }
        if (!sl.empty()) {
            op.bvert.push_back(i);
            bsl.push_back(sl);
            bco.push_back(co);
        }
        if (op.gface.size() > g0) {
            op.gvert.push_back(i);
            op.gstart.push_back((uint32_t)op.gface.size());
        }
    }
    // general faces: vertices (new numbering), S/12 and <|Omega.n| Y Y'>
    op.gfv.resize(3 * gfaces.size());
    op.garea.resize(gfaces.size());
    op.gR.clear();
    op.gOd.assign(gfaces.size() * (size_t)slm, 0.0);
    op.gOz.clear();
    if (!gfaces.empty()) {
        int maxk = 0;
        while ((uint32_t)((1 + maxk) * (1 + 2 * maxk)) < slm) maxk++;
        if ((uint32_t)((1 + maxk) * (1 + 2 * maxk)) != slm) return fail(TR_ERR_ARG, "assemble_operator: slm is not (1+K)(1+2K)");
        const double zaxis[3] = {0.0, 0.0, 1.0};
        abs_normal_matrix(maxk, zaxis, op.gOz);
        const uint32_t nr = rotation_size(maxk);
        op.gR.resize(gfaces.size() * (size_t)nr);
        std::vector<double> Rf;
        for (size_t g = 0; g < gfaces.size(); g++) {
            const tr_face &F = mv.bnd[gfaces[g]];
            for (int k = 0; k < 3; k++) op.gfv[3 * g + k] = op.iperm[F.p[k]];
            op.garea[g] = std::sqrt(F.s[0] * F.s[0] + F.s[1] * F.s[1] + F.s[2] * F.s[2]) / 12.0;
            normal_rotation(maxk, F.s, Rf);
            std::copy(Rf.begin(), Rf.end(), op.gR.begin() + g * (size_t)nr);
            // diagonal of R O_z R^T: harmonic a of block l (start s, size nb)
            for (int l = 0, s = 0, off = 0; l <= 2 * maxk; l += 2) {
                const int nb = 2 * l + 1;
                for (int a = 0; a < nb; a++) {
                    double v = 0;
                    for (int c = 0; c < nb; c++)
                        for (int d = 0; d < nb; d++)
                            v += Rf[off + a * nb + c] * op.gOz[(size_t)(s + c) * slm + s + d] * Rf[off + a * nb + d];
                    op.gOd[g * (size_t)slm + s + a] = v;
                }
                s += nb;
                off += nb * nb;
            }
        }
    }
    op.nB = (uint32_t)op.bvert.size();
    op.bof.assign(nP, -1);
    for (uint32_t b = 0; b < op.nB; b++) op.bof[op.bvert[b]] = (int32_t)b;
    op.WB = 1;
    for (auto &s : bsl) op.WB = std::max<uint32_t>(op.WB, (uint32_t)s.size());
    op.bnbr.assign((size_t)op.WB * std::max<uint32_t>(op.nB, 1), 0);
    op.bcoef.assign((size_t)op.WB * 3 * std::max<uint32_t>(op.nB, 1), 0.0);
    for (uint32_t b = 0; b < op.nB; b++)
        for (uint32_t s = 0; s < op.WB; s++) {
            op.bnbr[(size_t)s * op.nB + b] = s < bsl[b].size() ? bsl[b][s] : op.bvert[b];
            for (int n = 0; n < 3; n++)
                op.bcoef[((size_t)s * 3 + n) * op.nB + b] = s < bsl[b].size() ? bco[b][3 * s + n] : 0.0;
        }

    // ---- specular reflection (extension).  In the even-parity weak form a
    // reflecting plane is not obtained by dropping its boundary term alone:
    // the trial and test functions must also be symmetric under the plane's
    // mirror there, i.e. the harmonics odd under that mirror vanish at the
    // plane's vertices.  Only axis-aligned planes are supported.
    op.mask.clear();
    {
        std::vector<uint8_t> refl(nP, 0);
        for (uint32_t f = 0; f < mv.nF; f++) {
            if (!mv.face_kind || mv.face_kind[f] != 1) continue;
            const tr_face &F = mv.bnd[f];
            int axis = 0;
            for (int a = 1; a < 3; a++)
                if (std::fabs(F.s[a]) > std::fabs(F.s[axis])) axis = a;
            for (int a = 0; a < 3; a++)
                if (a != axis && std::fabs(F.s[a]) > 1e-9 * std::fabs(F.s[axis]))
                    return fail(TR_ERR_ARG, "reflective boundary faces must lie in axis-aligned planes");
            for (int k = 0; k < 3; k++) refl[op.iperm[F.p[k]]] |= (uint8_t)(1u << axis);
        }
        std::vector<uint8_t> odd;
        for (int l = 0; odd.size() < slm; l += 2)
            for (int m = -l; m <= l; m++) odd.push_back((uint8_t)mirror_odd(l, m));
        for (uint32_t i = 0; i < nP; i++)
            if (refl[i])
                for (uint32_t n = 0; n < slm; n++)
                    if (refl[i] & odd[n]) op.mask.push_back(i * slm + n);
    }

    // ---- moments table: omega[i][lm][0] and its b index (MeshData.cpp:108-113)
    op.mom_val.assign(3 * (size_t)slm, 0.0);
    op.mom_pos.assign(3 * (size_t)slm, 0);
    for (int i = 0; i < 3; i++)
        for (uint32_t lm = 0; lm < slm; lm++) {
            op.mom_val[(size_t)i * slm + lm] = av.omega[(size_t)i * slm * slm + (size_t)lm * slm];
            op.mom_pos[(size_t)i * slm + lm] = av.omega_pos[(size_t)i * slm * slm + (size_t)lm * slm];
        }

    // ---- right-hand side (kernels.cu:157-163, same summation order)
    op.rhs.assign(nP, 0.0);
    for (uint32_t i = 0; i < nP; i++) {
        uint32_t vo = op.perm[i];
        double sum = 0;
        for (uint32_t j = mv.tetstart[vo]; j < mv.tetstart[vo + 1]; j++) {
            const tr_tetrahedron &t = mv.mesh[mv.tetidx[j]];
            sum += t.kappa_volume * t.I_p * (1. / 4.);
        }
        op.rhs[i] = sum;
    }

    // ---- per-vertex element sums and coordinates for the learned guess's
    // features (the same sums as tr_mlp_features)
    op.vsum.clear();
    op.xyz.clear();
    if (mv.coords) {
        op.vsum.assign((size_t)3 * nP, 0.0);
        op.xyz.resize((size_t)3 * nP);
        const double *X = mv.coords;
        for (uint32_t i = 0; i < nP; i++)
            for (int k = 0; k < 3; k++) op.xyz[(size_t)k * nP + i] = X[3 * (size_t)op.perm[i] + k];
        for (uint32_t t = 0; t < mv.nT; t++) {
            const tr_tetrahedron &T = mv.mesh[t];
            const double *a = X + 3 * (size_t)T.p[0], *b = X + 3 * (size_t)T.p[1], *c = X + 3 * (size_t)T.p[2],
                         *d = X + 3 * (size_t)T.p[3];
            double u[3], v[3], w[3];
            for (int k = 0; k < 3; k++) {
                u[k] = b[k] - a[k];
                v[k] = c[k] - a[k];
                w[k] = d[k] - a[k];
            }
            const double V = std::fabs(u[0] * (v[1] * w[2] - v[2] * w[1]) - u[1] * (v[0] * w[2] - v[2] * w[0]) +
                                       u[2] * (v[0] * w[1] - v[1] * w[0])) / 6.0;
            if (!(V > 0)) return fail(TR_ERR_ARG, "assemble_operator: degenerate tetrahedron");
            const double kap = T.kappa_volume / V;
            for (int k = 0; k < 4; k++) {
                const uint32_t i = op.iperm[T.p[k]];
                op.vsum[i] += V;
                op.vsum[(size_t)nP + i] += kap * V;
                op.vsum[(size_t)2 * nP + i] += T.I_p * kap * V;
            }
        }
    }
    return TR_OK;
}

}  // namespace tr

namespace tr {

int concat_operators(const std::vector<HostOperator> &ops, const std::vector<double> &albedo, HostOperator &u,
                     std::vector<uint32_t> &voff) {
    if (ops.empty() || albedo.size() != ops.size()) return fail(TR_ERR_ARG, "concat_operators: no problems");
    const HostOperator &o0 = ops[0];
    const uint32_t slm = o0.slm, NC = o0.NC;
    uint64_t nP = 0, nB = 0;
    uint32_t W = 0, WB = 1;
    for (const HostOperator &o : ops) {
        if (o.slm != slm || o.NC != NC || o.arow != o0.arow || o.brow != o0.brow || o.omc != o0.omc ||
            o.onc != o0.onc || o.massfac != o0.massfac || !o.mf0.empty())
            return fail(TR_ERR_ARG, "concat_operators: the problems' angular data differ");
        nP += o.nP;
        nB += o.nB;
        W = std::max(W, o.W);
        WB = std::max(WB, o.WB);
    }
    if (nP >= (1ull << 32) || (double)nP * slm >= 4294967296.0)
        return fail(TR_ERR_ARG, "concat_operators: batch too large for 32-bit vertex ids");
    u = HostOperator();
    u.nP = (uint32_t)nP;
    u.slm = slm;
    u.aslm = o0.aslm;
    u.NC = NC;
    u.W = W;
    u.WB = WB;
    u.nB = (uint32_t)nB;
    u.arow = o0.arow;
    u.brow = o0.brow;
    u.aidx = o0.aidx;
    u.bidx = o0.bidx;
    u.aval = o0.aval;
    u.bval = o0.bval;
    u.massfac = o0.massfac;
    u.ddiag = o0.ddiag;
    u.bdiag = o0.bdiag;
    u.omc = o0.omc;
    u.onc = o0.onc;
    u.mom_val = o0.mom_val;
    u.mom_pos = o0.mom_pos;
    u.perm.resize(nP);
    u.iperm.resize(nP);
    u.nbr.assign((size_t)W * nP, 0);
    u.coef.assign((size_t)W * NC * nP, 0.0);
    u.bof.assign(nP, -1);
    u.bnbr.assign((size_t)WB * std::max<uint64_t>(nB, 1), 0);
    u.bcoef.assign((size_t)WB * 3 * std::max<uint64_t>(nB, 1), 0.0);
    u.rhs.resize(nP);
    u.mf0.resize(nP);
    bool feats = true;
    for (const HostOperator &o : ops) feats = feats && !o.vsum.empty() && !o.xyz.empty();
    if (feats) {
        u.vsum.resize((size_t)3 * nP);
        u.xyz.resize((size_t)3 * nP);
    }
    voff.assign(ops.size() + 1, 0);
    u.gstart.assign(1, 0);
    uint32_t v0 = 0, b0 = 0;
    for (size_t p = 0; p < ops.size(); p++) {
        const HostOperator &o = ops[p];
        voff[p] = v0;
        // vacuum faces off the axes: vertices and faces shifted by the
        // problems before this one
        u.general_normals = u.general_normals || o.general_normals;
        const uint32_t f0 = (uint32_t)u.garea.size(), j0 = (uint32_t)u.gface.size();
        for (size_t b = 0; b < o.gvert.size(); b++) {
            u.gvert.push_back(v0 + o.gvert[b]);
            u.gstart.push_back(j0 + o.gstart[b + 1]);
        }
        for (size_t j = 0; j < o.gface.size(); j++) {
            u.gface.push_back(f0 + o.gface[j]);
            u.gpos.push_back(o.gpos[j]);
        }
        for (uint32_t w : o.gfv) u.gfv.push_back(v0 + w);
        u.garea.insert(u.garea.end(), o.garea.begin(), o.garea.end());
        u.gR.insert(u.gR.end(), o.gR.begin(), o.gR.end());
        u.gOd.insert(u.gOd.end(), o.gOd.begin(), o.gOd.end());
        if (u.gOz.empty()) u.gOz = o.gOz;
        for (uint32_t i = 0; i < o.nP; i++) {
            const uint32_t g = v0 + i;
            u.perm[g] = v0 + o.perm[i];
            u.iperm[v0 + o.perm[i]] = g;
            u.rhs[g] = o.rhs[i];
            u.mf0[g] = 1.0 - albedo[p];
            u.bof[g] = o.bof[i] >= 0 ? (int32_t)(b0 + (uint32_t)o.bof[i]) : -1;
            if (feats)
                for (int k = 0; k < 3; k++) {
                    u.vsum[(size_t)k * nP + g] = o.vsum[(size_t)k * o.nP + i];
                    u.xyz[(size_t)k * nP + g] = o.xyz[(size_t)k * o.nP + i];
                }
            for (uint32_t s = 0; s < W; s++) {
                // slots past a problem's own W repeat the vertex with zero couplings
                u.nbr[(size_t)s * nP + g] = s < o.W ? v0 + o.nbr[(size_t)s * o.nP + i] : g;
                if (s < o.W)
                    for (uint32_t c = 0; c < NC; c++)
                        u.coef[((size_t)s * NC + c) * nP + g] = o.coef[((size_t)s * NC + c) * o.nP + i];
            }
        }
        for (uint32_t b = 0; b < o.nB; b++) {
            const uint32_t gb = b0 + b;
            u.bvert.push_back(v0 + o.bvert[b]);
            for (uint32_t s = 0; s < WB; s++) {
                u.bnbr[(size_t)s * nB + gb] = s < o.WB ? v0 + o.bnbr[(size_t)s * o.nB + b] : v0 + o.bvert[b];
                if (s < o.WB)
                    for (int n = 0; n < 3; n++)
                        u.bcoef[((size_t)s * 3 + n) * nB + gb] = o.bcoef[((size_t)s * 3 + n) * o.nB + b];
            }
        }
        for (uint32_t e : o.mask) u.mask.push_back(v0 * slm + e);
        v0 += o.nP;
        b0 += o.nB;
    }
    voff[ops.size()] = v0;
    return TR_OK;
}

}  // namespace tr
